// The maps' consumer on gfx950: AutoDock 4.2.1's float map set and its
// trilinear interpolation of ligand atoms (AD/trilinterp.cc:87-185, AD =
// autogrid_original_4_2_1_ref/autodock in the reference).  Included by
// capi_device.hip.
//
// The double arithmetic keeps the reference's operation order (no
// contraction in this file), so every output equals the reference's bit for
// bit.  Totals are the reference's running sums in atom order: one lane per
// (pose, atom) computes the atom's two contributions into LDS, one lane per
// pose adds them in order.  The map set is read from a corner-packed copy
// (k_pack_cells): the same floats, a cell's 8 corners of a slot side by side.
#pragma clang fp contract(off)

#include "kernels.hpp"

namespace fgk {

constexpr int kInterpThreads = 256;

// The float AutoDock reads back from a printed map value: "0" below 0.0005
// in magnitude, else "%.3f" of floor(1000 v + 0.5) * 0.001, which is the
// decimal k / 1000 (k = floor(1000 v + 0.5)); read with "%f" that is the
// float nearest to k / 1000, i.e. float(double(k) / 1000).
__device__ __forceinline__ float autodockValue(double v)
{
    if (fabs(v) < 0.0005) return 0.f;
    const double k = floor(v * 1000 + 0.5);
    return float(k / 1000.0);
}

// maps [V][P] (x fastest, as a job holds them) -> out [P][V] (AutoDock's
// [z][y][x][slot]).  One lane per grid point.
__global__ __launch_bounds__(kInterpThreads) void k_export_mapset(const double *maps, int64_t P, int V, float *out)
{
    const int64_t p = int64_t(blockIdx.x) * kInterpThreads + threadIdx.x;
    if (p >= P) return;
    for (int m = 0; m < V; ++m) out[p * V + m] = autodockValue(maps[int64_t(m) * P + p]);
}

// Corner-packed copy of a map set: for every cell (the grid point with the
// lowest indices of its 8 corners) and slot, the 8 corner values in the order
// u + 2v + 4w, 0 for a corner past the grid (as AutoDock's zero-padded array
// reads).  One lane per (cell, slot).
__global__ __launch_bounds__(kInterpThreads) void k_pack_cells(const float *map, int nx, int ny, int nz, int V,
                                                               float *cells)
{
    const int64_t q = int64_t(blockIdx.x) * kInterpThreads + threadIdx.x;
    const int64_t P = int64_t(nx) * ny * nz;
    if (q >= P * V) return;
    const int s = int(q % V);
    const int64_t p = q / V;
    const int i = int(p % nx), j = int((p / nx) % ny), k = int(p / (int64_t(nx) * ny));
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int u = i + (c & 1), w = j + ((c >> 1) & 1), z = k + (c >> 2);
        v[c] = (u < nx && w < ny && z < nz) ? map[((int64_t(z) * ny + w) * nx + u) * V + s] : 0.f;
    }
    float4 *out = reinterpret_cast<float4 *>(cells + (q << 3));
    out[0] = make_float4(v[0], v[1], v[2], v[3]);
    out[1] = make_float4(v[4], v[5], v[6], v[7]);
}

struct InterpArgs {
    double lo[3], hi[3], center[3], inv;
    int32_t nx, ny, nz, V;       // points per axis, slots
    int32_t natoms, npose, elecSlot, desolvSlot;
    const float *cells;          // [nz][ny][nx][V][8] corner values per cell (k_pack_cells)
    const float *xyz;            // [npose][natoms][3]
    const int32_t *type;         // [natoms]
    const float *charge, *absq;  // [natoms]
    const int32_t *ignore;       // [natoms] or null
    float *elec, *emap;          // [npose][natoms] or null
    double *elecTotal, *emapTotal;  // [npose] or null
    float *total;                // [npose] or null
};

// Atom i of pose `pose` (idx = pose * natoms + i): what it adds to the two
// totals (0 for an ignored atom), and its per-atom outputs.
__device__ __forceinline__ void interpAtom(const InterpArgs &a, int64_t idx, int i, double &addE, double &addM)
{
    if (a.ignore && a.ignore[i]) {
        if (a.elec) a.elec[idx] = 0.f;
        if (a.emap) a.emap[idx] = 0.f;
        addE = addM = 0.0;
        return;
    }
    const float fx = a.xyz[3 * idx], fy = a.xyz[3 * idx + 1], fz = a.xyz[3 * idx + 2];
    double x = fx, y = fy, z = fz;
    if (x <= a.lo[0] || x >= a.hi[0] || y <= a.lo[1] || y >= a.hi[1] || z <= a.lo[2] || z >= a.hi[2]) {
        // outside the grid: 500 x the squared distance from the centre
        x -= a.center[0];
        y -= a.center[1];
        z -= a.center[2];
        const double ep = (x * x + y * y + z * z) * 500.0;
        addE = addM = ep;
        if (a.elec) a.elec[idx] = float(ep);
        if (a.emap) a.emap[idx] = float(ep);
        return;
    }
    const int t = a.type[i];
    const double u = (double(fx) - a.lo[0]) * a.inv;
    const int u0 = int(u);
    const double p0u = u - double(u0), p1u = 1.0 - p0u;
    const double v = (double(fy) - a.lo[1]) * a.inv;
    const int v0 = int(v);
    const double p0v = v - double(v0), p1v = 1.0 - p0v;
    const double w = (double(fz) - a.lo[2]) * a.inv;
    const int w0 = int(w);
    const double p0w = w - double(w0), p1w = 1.0 - p0w;
    // the cell's 8 corner values of each slot side by side (corner u + 2v +
    // 4w, 0 one past the grid's last point): 3 x 32 bytes per atom
    const int64_t cell = (int64_t(w0) * a.ny + v0) * a.nx + u0;
    const float4 *c = reinterpret_cast<const float4 *>(a.cells + ((cell * a.V) << 3));
    const int es = a.elecSlot, ds = a.desolvSlot;
    const float4 e0 = c[2 * es], e1 = c[2 * es + 1];
    const float4 m0 = c[2 * t], m1 = c[2 * t + 1];
    const float4 d0 = c[2 * ds], d1 = c[2 * ds + 1];
    // corner weights as the reference forms them, p1u * p1v * p1w first
    const double wt[8] = {p1u * p1v * p1w, p0u * p1v * p1w, p1u * p0v * p1w, p0u * p0v * p1w,
                          p1u * p1v * p0w, p0u * p1v * p0w, p1u * p0v * p0w, p0u * p0v * p0w};
    const float ev[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
    const float mv[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
    const float dv[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
    // the three running sums over the corners in the reference's order
    double e = 0, m = 0, d = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        e += wt[q] * double(ev[q]);
        m += wt[q] * double(mv[q]);
        d += wt[q] * double(dv[q]);
    }
    addE = e * double(a.charge[i]);
    addM = m + d * double(a.absq[i]);
    if (a.elec) a.elec[idx] = float(addE);
    if (a.emap) a.emap[idx] = float(addM);
}

__device__ __forceinline__ void writeTotals(const InterpArgs &a, int pose, double et, double mt)
{
    if (a.elecTotal) a.elecTotal[pose] = et;
    if (a.emapTotal) a.emapTotal[pose] = mt;
    if (a.total) a.total[pose] = float(double(float(et)) + mt);
}

// Poses of up to kInterpThreads atoms: each workgroup takes
// kInterpThreads / natoms whole poses, one lane per atom; the contributions
// go through LDS and one lane per pose adds them in atom order (the
// reference's running sums), then writes the totals.
__global__ __launch_bounds__(kInterpThreads) void k_interp_poses(InterpArgs a)
{
    __shared__ double sE[kInterpThreads], sM[kInterpThreads];
    const int t = threadIdx.x, A = a.natoms;
    const int ppw = kInterpThreads / A;
    const int g = int(blockIdx.x);
    const int sp = g * ppw + t / A, i = t % A;
    if (t < ppw * A && sp < a.npose) interpAtom(a, int64_t(sp) * A + i, i, sE[t], sM[t]);
    __syncthreads();
    const int own = g * ppw + t;
    if (t < ppw && own < a.npose) {
        double et = 0, mt = 0;
        for (int k = 0; k < A; ++k) {
            et += sE[t * A + k];
            mt += sM[t * A + k];
        }
        writeTotals(a, own, et, mt);
    }
}

// Poses of more than kInterpThreads atoms: one workgroup per pose, atoms in
// chunks of kInterpThreads, lane 0 adding each chunk in atom order.
__global__ __launch_bounds__(kInterpThreads) void k_interp_big_poses(InterpArgs a)
{
    __shared__ double sE[kInterpThreads], sM[kInterpThreads];
    const int t = threadIdx.x, A = a.natoms, pose = int(blockIdx.x);
    double et = 0, mt = 0;
    for (int c0 = 0; c0 < A; c0 += kInterpThreads) {
        const int i = c0 + t, n = min(kInterpThreads, A - c0);
        if (i < A) interpAtom(a, int64_t(pose) * A + i, i, sE[t], sM[t]);
        __syncthreads();
        if (t == 0)
            for (int k = 0; k < n; ++k) {
                et += sE[k];
                mt += sM[k];
            }
        __syncthreads();
    }
    if (t == 0) writeTotals(a, pose, et, mt);
}

}  // namespace fgk
