// Driver of the reference's own trilinterp (AD/trilinterp.cc, compiled from
// its source by oracle/ref.mk into oracle/_ref/).  Test infrastructure: it
// writes the golden vectors of tests/golden/interp/ and is never part of the
// product.  AD = autogrid_original_4_2_1_ref/autodock.
//
//   ref_trilinterp in.bin out.bin
//
// in.bin (native little-endian): int32 nx1 ny1 nz1 nslots elec_slot
// desolv_slot natoms npose outside; float64 lo[3] hi[3] center[3]
// inv_spacing; float32 map[nz1][ny1][nx1][nslots]; int32 type[natoms];
// float32 charge[natoms] abs_charge[natoms]; int32 ignore[natoms];
// float32 xyz[npose][natoms][3].
// out.bin: per pose float32 elec[natoms] emap[natoms], float32 elec_total
// emap_total (the reference's Real outputs) and float32 returned value.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "trilinterp.h"

// the caller-set map indices the reference reads as globals (AD/trilinterp.cc:40-41)
int ElecMap;
int DesolvMap;

namespace {

template <class T>
bool rd(FILE *f, T *p, size_t n)
{
    return fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: ref_trilinterp in.bin out.bin\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 1;
    int hdr[9];
    double geo[10];
    if (!rd(f, hdr, 9) || !rd(f, geo, 10)) return 1;
    const int nx1 = hdr[0], ny1 = hdr[1], nz1 = hdr[2], nslots = hdr[3], natoms = hdr[6], npose = hdr[7];
    if (nx1 > MAX_GRID_PTS || ny1 > MAX_GRID_PTS || nz1 > MAX_GRID_PTS || nslots > MAX_MAPS || natoms > MAX_ATOMS) {
        fprintf(stderr, "input exceeds the reference's static array bounds\n");
        return 1;
    }
    ElecMap = hdr[4];
    DesolvMap = hdr[5];
    const int outside = hdr[8];
    GridMapSetInfo info;
    memset(&info, 0, sizeof info);
    for (int k = 0; k < 3; ++k) {
        info.lo[k] = geo[k];
        info.hi[k] = geo[3 + k];
        info.center[k] = geo[6 + k];
    }
    info.inv_spacing = geo[9];
    info.num_points1[0] = nx1;
    info.num_points1[1] = ny1;
    info.num_points1[2] = nz1;
    typedef MapType Plane[MAX_GRID_PTS][MAX_GRID_PTS][MAX_MAPS];
    Plane *map = static_cast<Plane *>(calloc(MAX_GRID_PTS, sizeof(Plane)));
    std::vector<float> packed(size_t(nx1) * ny1 * nz1 * nslots);
    if (!map || !rd(f, packed.data(), packed.size())) return 1;
    for (int z = 0; z < nz1; ++z)
        for (int y = 0; y < ny1; ++y)
            for (int x = 0; x < nx1; ++x)
                for (int m = 0; m < nslots; ++m)
                    map[z][y][x][m] = packed[((size_t(z) * ny1 + y) * nx1 + x) * nslots + m];
    static int type[MAX_ATOMS], ignore[MAX_ATOMS];
    static Real charge[MAX_ATOMS], absq[MAX_ATOMS], tcoord[MAX_ATOMS][SPACE], elec[MAX_ATOMS], emap[MAX_ATOMS];
    std::vector<float> q(natoms), aq(natoms);
    if (!rd(f, type, natoms) || !rd(f, q.data(), natoms) || !rd(f, aq.data(), natoms) || !rd(f, ignore, natoms))
        return 1;
    for (int i = 0; i < natoms; ++i) {
        charge[i] = q[size_t(i)];
        absq[i] = aq[size_t(i)];
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 1;
    std::vector<float> xyz(size_t(natoms) * 3);
    for (int p = 0; p < npose; ++p) {
        if (!rd(f, xyz.data(), xyz.size())) return 1;
        for (int i = 0; i < natoms; ++i)
            for (int k = 0; k < 3; ++k) tcoord[i][k] = xyz[size_t(3 * i + k)];
        Real et = 0, mt = 0;
        const Real r = trilinterp(0, natoms, tcoord, charge, absq, type, map, &info, outside, ignore, elec, emap,
                                  &et, &mt);
        std::vector<float> rec;
        for (int i = 0; i < natoms; ++i) rec.push_back(elec[i]);
        for (int i = 0; i < natoms; ++i) rec.push_back(emap[i]);
        rec.push_back(et);
        rec.push_back(mt);
        rec.push_back(r);
        fwrite(rec.data(), sizeof(float), rec.size(), o);
    }
    fclose(o);
    fclose(f);
    free(map);
    return 0;
}
