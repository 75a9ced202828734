// Synthetic receptors for the workloads the maps are measured on.
//
// Atoms are placed by random sequential packing: a uniform proposal inside
// the region (sphere or centred box), rejected when closer than min_sep to an
// atom already placed (checked through a cell list of min_sep cells).  Types
// follow protein-like proportions by exact largest-remainder counts
// (C 55 %, A 5 %, N 15 %, OA 18 %, SA 2 %, HD 5 %), shuffled.  Charges are
// N(0, 0.25) clipped to +-0.7, rounded to 0.001 e and shifted so the printed
// charges sum to exactly zero.  Everything derives from one xoshiro256**
// stream seeded through splitmix64, so a seed gives the same files on every
// machine.  The GPF lists, in order: npts, gridfld, spacing, receptor_types,
// ligand_types, receptor, gridcenter auto, smooth, map..., elecmap, dsolvmap,
// dielectric -0.1465 (no parameter_file: AD4.1 defaults apply).  This is the
// AutoDockTools order; "receptor" must not directly follow "receptor_types",
// because the reference compares keywords in a buffer it never terminates
// and then reads the stale "_types" suffix (FG/InputDataLoader.cpp:803-821).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>
#include <unordered_map>

#include "host.hpp"

namespace fgh {
namespace {

struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed)
    {
        for (uint64_t &v : s) {  // splitmix64
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            v = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double normal()  // Box-Muller, one value per call
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

struct CellKey {
    int64_t operator()(int x, int y, int z) const
    {
        return (int64_t(x) + (1 << 20)) | ((int64_t(y) + (1 << 20)) << 21) | ((int64_t(z) + (1 << 20)) << 42);
    }
};

}  // namespace

int synthWrite(const fg_synth_spec &s, const std::string &dir, std::string &err)
{
    if (s.n_atoms <= 0 || s.min_sep <= 0 || s.spacing <= 0 || !s.stem || !s.ligand_types) {
        err = "invalid synthetic receptor spec";
        return -1;
    }
    for (int i = 0; i < 3; ++i)
        if (s.npts[i] < 1 || s.npts[i] % 2 == 0 || s.npts[i] > 1025) {
            err = "npts must be odd and at most 1025";
            return -1;
        }
    Xoshiro256 rng(s.seed);
    const int N = s.n_atoms;
    const double sep2 = s.min_sep * s.min_sep, cell = s.min_sep;
    std::vector<double> X, Y, Z;
    X.reserve(size_t(N));
    Y.reserve(size_t(N));
    Z.reserve(size_t(N));
    std::unordered_map<int64_t, std::vector<int>> grid;
    CellKey key;
    const int64_t maxTries = int64_t(N) * 2000 + 100000;
    int64_t tries = 0;
    while (int(X.size()) < N) {
        if (++tries > maxTries) {
            err = "random sequential packing did not converge (region too small for n_atoms)";
            return -1;
        }
        double p[3];
        if (s.shape == 0) {
            do {
                for (double &v : p) v = (2 * rng.uniform() - 1) * s.a;
            } while (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > s.a * s.a);
        } else {
            p[0] = (rng.uniform() - 0.5) * s.a;
            p[1] = (rng.uniform() - 0.5) * s.b;
            p[2] = (rng.uniform() - 0.5) * s.c;
        }
        // coordinates are printed with 3 decimals; test the printed values
        for (double &v : p) v = std::round(v * 1000.0) / 1000.0;
        int c[3] = {int(std::floor(p[0] / cell)), int(std::floor(p[1] / cell)), int(std::floor(p[2] / cell))};
        bool ok = true;
        for (int dz = -1; dz <= 1 && ok; ++dz)
            for (int dy = -1; dy <= 1 && ok; ++dy)
                for (int dx = -1; dx <= 1 && ok; ++dx) {
                    auto it = grid.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == grid.end()) continue;
                    for (int j : it->second) {
                        double ddx = X[size_t(j)] - p[0], ddy = Y[size_t(j)] - p[1], ddz = Z[size_t(j)] - p[2];
                        if (ddx * ddx + ddy * ddy + ddz * ddz < sep2) {
                            ok = false;
                            break;
                        }
                    }
                }
        if (!ok) continue;
        grid[key(c[0], c[1], c[2])].push_back(int(X.size()));
        X.push_back(p[0]);
        Y.push_back(p[1]);
        Z.push_back(p[2]);
    }
    // types by largest remainder
    static const char *const kTypes[6] = {"C", "A", "N", "OA", "SA", "HD"};
    static const double kFrac[6] = {0.55, 0.05, 0.15, 0.18, 0.02, 0.05};
    int count[6], assigned = 0;
    std::vector<std::pair<double, int>> rem;
    for (int t = 0; t < 6; ++t) {
        double want = kFrac[t] * N;
        count[t] = int(std::floor(want));
        assigned += count[t];
        rem.emplace_back(want - count[t], t);
    }
    std::stable_sort(rem.begin(), rem.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    for (int i = 0; assigned < N; ++i, ++assigned) ++count[rem[size_t(i) % 6].second];
    std::vector<int> types;
    types.reserve(size_t(N));
    for (int t = 0; t < 6; ++t) types.insert(types.end(), size_t(count[t]), t);
    for (int i = N - 1; i > 0; --i) std::swap(types[size_t(i)], types[size_t(rng.below(uint64_t(i) + 1))]);
    // charges in units of 0.001 e, summing to zero
    std::vector<int> q(static_cast<size_t>(N));
    long long sum = 0;
    for (int i = 0; i < N; ++i) {
        double v = std::max(-0.7, std::min(0.7, 0.25 * rng.normal()));
        q[size_t(i)] = int(std::lround(v * 1000.0));
        sum += q[size_t(i)];
    }
    for (int pass = 0; sum != 0 && pass < 1000; ++pass)
        for (int i = 0; i < N && sum != 0; ++i) {
            int step = sum > 0 ? -1 : 1;
            if (std::abs(q[size_t(i)] + step) > 700) continue;
            q[size_t(i)] += step;
            sum += step;
        }
    // PDBQT: x at column 30, y 38, z 46, charge 70, type 77
    std::string stem = s.stem, base = dir.empty() ? std::string(".") : dir;
    FILE *f = std::fopen((base + "/" + stem + ".pdbqt").c_str(), "w");
    if (!f) {
        err = "cannot write " + base + "/" + stem + ".pdbqt";
        return -3;
    }
    for (int i = 0; i < N; ++i) {
        const char *t = kTypes[types[size_t(i)]];
        char name[5];
        std::snprintf(name, sizeof name, "%s%d", t[0] == 'H' ? "H" : (t[0] == 'A' ? "C" : std::string(t, 1).c_str()),
                      (i % 9) + 1);
        std::fprintf(f, "ATOM  %5d %-4s SYN A%4d    %8.3f%8.3f%8.3f  1.00  0.00    %6.3f %-2s\n", (i + 1) % 100000,
                     name, (i / 10) % 10000, X[size_t(i)], Y[size_t(i)], Z[size_t(i)], q[size_t(i)] / 1000.0, t);
    }
    std::fprintf(f, "TER\n");
    std::fclose(f);
    // GPF
    std::vector<std::string> lig;
    {
        std::istringstream is(s.ligand_types);
        std::string w;
        while (is >> w) lig.push_back(w);
    }
    FILE *g = std::fopen((base + "/" + stem + ".gpf").c_str(), "w");
    if (!g) {
        err = "cannot write " + base + "/" + stem + ".gpf";
        return -3;
    }
    std::fprintf(g, "npts %d %d %d\n", s.npts[0] - 1, s.npts[1] - 1, s.npts[2] - 1);
    std::fprintf(g, "gridfld %s.maps.fld\n", stem.c_str());
    std::fprintf(g, "spacing %.3f\n", s.spacing);
    std::fprintf(g, "receptor_types A C HD N OA SA\n");
    std::fprintf(g, "ligand_types %s\n", s.ligand_types);
    std::fprintf(g, "receptor %s.pdbqt\n", stem.c_str());
    std::fprintf(g, "gridcenter auto\n");
    std::fprintf(g, "smooth 0.5\n");
    for (const std::string &t : lig) std::fprintf(g, "map %s.%s.map\n", stem.c_str(), t.c_str());
    std::fprintf(g, "elecmap %s.e.map\n", stem.c_str());
    std::fprintf(g, "dsolvmap %s.d.map\n", stem.c_str());
    std::fprintf(g, "dielectric -0.1465\n");
    std::fclose(g);
    return 0;
}

}  // namespace fgh
