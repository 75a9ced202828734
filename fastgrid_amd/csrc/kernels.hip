// CDNA4 (gfx950) map kernels.  See kernels.hpp for the decomposition.
// FG = fastgrid_master/fastgrid in the reference.
//
// FP64 code in this file keeps the reference's operation order: the file is
// compiled without floating-point contraction, and FP32 fused multiply-adds
// are written explicitly with __builtin_fmaf.
#pragma clang fp contract(off)

#include <type_traits>

#include "kernels.hpp"

namespace fgk {

// ------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------
__device__ __forceinline__ double dmin_(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax_(double a, double b) { return a > b ? a : b; }
// libm by precision: the FP64 forms call the double functions, the FP32
// plain form the float ones
__host__ __device__ __forceinline__ double acos_(double x) { return acos(x); }
__host__ __device__ __forceinline__ float acos_(float x) { return acosf(x); }
__host__ __device__ __forceinline__ double cos_(double x) { return cos(x); }
__host__ __device__ __forceinline__ float cos_(float x) { return cosf(x); }
__host__ __device__ __forceinline__ double sin_(double x) { return sin(x); }
__host__ __device__ __forceinline__ float sin_(float x) { return sinf(x); }
__host__ __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ float sqrt_(float x) { return sqrtf(x); }
// Math::Acos with the reference's clamp (FG/math/Math.h:51)
template <class T>
__host__ __device__ __forceinline__ T acosClamped(T f)
{
    return f > 1 ? T(0) : (f < -1 ? T(3.14159265358979323846) : acos_(f));
}
// angstromToIndex (FG/autogrid.h:120-125)
__device__ __forceinline__ int binOf(double r)
{
    int i = int(r * 100.0);
    return i < kTableLen - 1 ? i : kTableLen - 1;
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The FP32 table bin of the far field: r2 * rsq(r2) + (2^23 - 1/2) lands in
// [2^23, 2^24), where the FP32 spacing is 1, so its bit pattern is 2^23's
// plus floor(100 r) (coordinates in units of 0.01 A).
constexpr float kFloorBias = 8388607.5f;       // 2^23 - 1/2
constexpr unsigned kFloorBits = 0x4B000000u;   // bit pattern of 2^23
// Split of the compensated e-map: a pair whose FP32 bin is below kNearBins
// (100 r < 800) is the cutoff kernel's, in FP64; every other pair is the
// far-field kernel's, in FP32.  The far-field kernel drops its near pairs by
// reading zeros from table bins 0..kNearBins-1.
constexpr int kNearBins = 800;

// Added to dy^2 in the far-field arithmetic: every nonzero square of these
// scaled differences (>= ~1e-6) absorbs it exactly, so only an atom sitting
// exactly on a grid point sees it, as a finite 1/r (1e15) times a zero table
// entry instead of inf * 0.
constexpr float kTinyR2 = 1e-30f;

// The far-field kernel's FP32 squared distance (scaled by 1e4) between atom
// (x, y, z, scaled by 100) and a grid point: the operations in its order.
__device__ __forceinline__ float farR2(float ax, float ay, float az, float gx, float gy, float gz)
{
    const float dy = ay - gy;
    const float dy2 = __builtin_fmaf(dy, dy, kTinyR2);
    const float dx = ax - gx;
    const float dxy2 = __builtin_fmaf(dx, dx, dy2);
    const float dz = az - gz;
    return __builtin_fmaf(dz, dz, dxy2);
}
// The biased bin bits of a far-field pair with FP32 squared distance r2 >= 1.
__device__ __forceinline__ unsigned farBinBits(float r2)
{
    return __float_as_uint(__builtin_fmaf(r2, __builtin_amdgcn_rsqf(r2), kFloorBias));
}

// Read-only data of a launch seen through the constant address space:
// wave-uniform loads of it become scalar loads, also in persistent loops
// whose stores (elsewhere) would otherwise make the compiler reload it per
// lane.  The cast keeps the address (global and constant memory share it).
#define FG_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const FG_CONST T *constView(const T *p)
{
    return (const FG_CONST T *)p;
}
// A copy of *p (uniform address) loaded with scalar loads.
template <class T>
__device__ __forceinline__ T loadConst(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T out;
    const FG_CONST int32_t *src = (const FG_CONST int32_t *)p;
    int32_t *dst = reinterpret_cast<int32_t *>(&out);
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 4); ++i) dst[i] = src[i];
    return out;
}

// ------------------------------------------------------------------------
// Candidate lists: for each brick of the cutoff kernel, every atom whose
// distance to the brick's box is <= reach (8 A plus the band in which the
// cutoff kernel decides e-map pairs by the far-field kernel's FP32 bin), and
// the DS/D1 atoms within 8 A + 1e-4 A.  Order: coarse cells in z, y, x
// order, ascending atom index within a cell; the same on every run.
// ------------------------------------------------------------------------
constexpr int kListThreads = 256;
constexpr int kMaxCellsPerBrick = 1024;

struct ListArgs {
    GridGeom g;
    CellGrid cg;
    int32_t nbx, nby, nbz, N;
    double reach;             // list reach (A), >= 8 + 1e-4
    const double *xyz;        // [N][3]
    const int32_t *hbond;     // [N]
    const int32_t *cellStart; // [ncells+1]
    const int32_t *cellAtoms; // atoms sorted by cell, ascending within a cell
    int32_t *counts;          // [nb][2] all, H (count pass)
    const int32_t *start;     // [nb+1] (fill pass)
    const int32_t *hStart;    // [nb+1] (fill pass)
    int32_t *list, *hList;
    // type sort (k_type_sort): receptor type per atom, output list, run starts
    const int32_t *atomType;
    int32_t *sorted;
    int32_t *typeStart;       // [nb][T+1]
    int32_t T;
};

struct ListLds {
    int32_t cellOff[kMaxCellsPerBrick + 1];
    int32_t cellId[kMaxCellsPerBrick];
    int32_t scanBuf[kListThreads], scanH[kListThreads];
    int32_t base, baseH;
};

// Candidate list of brick b: counts it (kFill false) or writes it.
template <bool kFill>
__device__ __forceinline__ void brickList(const ListArgs &a, int b, ListLds &S)
{
    int32_t *cellOff = S.cellOff, *cellId = S.cellId, *scanBuf = S.scanBuf, *scanH = S.scanH;
    int32_t &base = S.base, &baseH = S.baseH;
    const int tid = threadIdx.x;
    const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
    const int ib[3] = {bx * kCutBX, by * kCutBY, bz * kCutBZ};
    const int ext[3] = {kCutBX, kCutBY, kCutBZ};
    const int n[3] = {a.g.nx, a.g.ny, a.g.nz}, d2[3] = {a.g.dx2, a.g.dy2, a.g.dz2};
    double lo[3], hi[3];
    int c0[3], c1[3];
    const double reach = a.reach, reachH = 8.0 + 1e-4;
    for (int k = 0; k < 3; ++k) {
        int i1 = ib[k] + ext[k] - 1;
        if (i1 > n[k] - 1) i1 = n[k] - 1;
        lo[k] = (ib[k] - d2[k]) * a.g.spacing;
        hi[k] = (i1 - d2[k]) * a.g.spacing;
        c0[k] = int(floor((lo[k] - reach - a.cg.origin[k]) / a.cg.size));
        c1[k] = int(floor((hi[k] + reach - a.cg.origin[k]) / a.cg.size));
        if (c0[k] < 0) c0[k] = 0;
        if (c1[k] > a.cg.nc[k] - 1) c1[k] = a.cg.nc[k] - 1;
    }
    const int ncx = c1[0] - c0[0] + 1, ncy = c1[1] - c0[1] + 1, ncz = c1[2] - c0[2] + 1;
    // at most 8 cells per axis (the cell edge is at least 4 A and the brick's)
    const int ncell = (ncx > 0 && ncy > 0 && ncz > 0) ? ncx * ncy * ncz : 0;
    const double reach2 = reach * reach, reachH2 = reachH * reachH;
    // cell table and prefix of atom counts, cells in z, y, x order: thread
    // t takes cells 4t .. 4t+3; a cell whose box lies wholly beyond the reach
    // of the brick box counts no atoms (atoms lie inside their cell's box;
    // 1e-6 A of slack covers the rounding of the cell index)
    constexpr int kCellsPerThread = kMaxCellsPerBrick / kListThreads;
    int cnt4[kCellsPerThread], mine = 0;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int m = kCellsPerThread * tid + q;
        int n = 0;
        if (m < ncell) {
            const int ci[3] = {c0[0] + m % ncx, c0[1] + (m / ncx) % ncy, c0[2] + m / (ncx * ncy)};
            const int c = (ci[2] * a.cg.nc[1] + ci[1]) * a.cg.nc[0] + ci[0];
            double dd = 0;
            for (int k = 0; k < 3; ++k) {
                const double clo = a.cg.origin[k] + ci[k] * a.cg.size, chi = clo + a.cg.size;
                const double t = fmax(fmax(clo - hi[k], lo[k] - chi), 0.0);
                dd += t * t;
            }
            const double r = reach + 1e-6;
            if (dd <= r * r) n = a.cellStart[c + 1] - a.cellStart[c];
            cellId[m] = c;
        }
        cnt4[q] = n;
        mine += n;
    }
    scanBuf[tid] = mine;
    if (tid == 0) {
        base = kFill ? a.start[b] : 0;
        baseH = kFill ? a.hStart[b] : 0;
    }
    __syncthreads();
    for (int s = 1; s < kListThreads; s <<= 1) {  // inclusive scan of the per-thread counts
        const int x = tid >= s ? scanBuf[tid - s] : 0;
        __syncthreads();
        scanBuf[tid] += x;
        __syncthreads();
    }
    {
        int off = scanBuf[tid] - mine;
#pragma unroll
        for (int q = 0; q < kCellsPerThread; ++q) {
            const int m = kCellsPerThread * tid + q;
            if (m < ncell) cellOff[m] = off;
            off += cnt4[q];
        }
        if (tid == kListThreads - 1) cellOff[ncell] = scanBuf[tid];
    }
    __syncthreads();
    const int m = ncell;
    const int total = cellOff[m];
    int cnt = 0, cntH = 0;
    for (int v0 = 0; v0 < total; v0 += kListThreads) {
        const int v = v0 + tid;
        int keep = 0, keepH = 0, ia = -1;
        if (v < total) {
            int lo_ = 0, hi_ = m - 1;  // last cell with cellOff <= v
            while (lo_ < hi_) {
                int mid = (lo_ + hi_ + 1) >> 1;
                if (cellOff[mid] <= v) lo_ = mid;
                else hi_ = mid - 1;
            }
            const int c = cellId[lo_];
            ia = a.cellAtoms[a.cellStart[c] + (v - cellOff[lo_])];
            double dd = 0;
            for (int k = 0; k < 3; ++k) {
                double p = a.xyz[3 * ia + k], t = 0;
                if (p < lo[k]) t = lo[k] - p;
                else if (p > hi[k]) t = p - hi[k];
                dd += t * t;
            }
            keep = dd <= reach2;
            keepH = dd <= reachH2 && (a.hbond[ia] == HB_DS || a.hbond[ia] == HB_D1);
        }
        if (!kFill) {
            cnt += keep;
            cntH += keepH;
            continue;
        }
        // block-wide exclusive scan of keep / keepH (Hillis-Steele in LDS)
        scanBuf[tid] = keep;
        scanH[tid] = keepH;
        __syncthreads();
        for (int s = 1; s < kListThreads; s <<= 1) {
            int x = tid >= s ? scanBuf[tid - s] : 0, y = tid >= s ? scanH[tid - s] : 0;
            __syncthreads();
            scanBuf[tid] += x;
            scanH[tid] += y;
            __syncthreads();
        }
        const int posA = base + scanBuf[tid] - keep, posH = baseH + scanH[tid] - keepH;
        if (keep) a.list[posA] = ia;
        if (keepH) a.hList[posH] = ia;
        __syncthreads();
        if (tid == kListThreads - 1) {
            base += scanBuf[tid];
            baseH += scanH[tid];
        }
        __syncthreads();
    }
    if (!kFill) {
        // block reduction of the two counts
        scanBuf[tid] = cnt;
        scanH[tid] = cntH;
        __syncthreads();
        for (int s = kListThreads / 2; s > 0; s >>= 1) {
            if (tid < s) {
                scanBuf[tid] += scanBuf[tid + s];
                scanH[tid] += scanH[tid + s];
            }
            __syncthreads();
        }
        if (tid == 0) {
            a.counts[2 * b] = scanBuf[0];
            a.counts[2 * b + 1] = scanH[0];
        }
    }
}

template <bool kFill>
__global__ __launch_bounds__(kListThreads) void k_brick_lists(ListArgs a)
{
    __shared__ ListLds S;
    brickList<kFill>(a, int(blockIdx.x), S);
}

// Exclusive scan of counts[i*2 + which] into out[0..n] (out[n] = total) by
// one workgroup of 1024 threads.
__device__ __forceinline__ void scanCounts(const int32_t *counts, int n, int which, int32_t *out, int32_t *part)
{
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = tid * per, b1 = min(n, b0 + per);
    int s = 0;
    for (int i = b0; i < b1; ++i) s += counts[2 * i + which];
    part[tid] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        int x = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += x;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int i = b0; i < b1; ++i) {
        out[i] = run;
        run += counts[2 * i + which];
    }
    if (tid == 1023) out[n] = part[1023];
}

// out[0..n] as scanCounts; the total out[n] also goes to *total (when given),
// so that the totals of many jobs can be read back with one copy.
__global__ __launch_bounds__(1024) void k_scan_counts(const int32_t *counts, int n, int which, int32_t *out,
                                                      int32_t *total)
{
    __shared__ int32_t part[1024];
    scanCounts(counts, n, which, out, part);
    if (total && threadIdx.x == 1023) *total = out[n];
}

// ------------------------------------------------------------------------
// Cutoff kernel: affinity maps + desolvation map + near-field electrostatics
// (FG/CalculateGridmaps.cpp:238-316, FG/electrostatics/Cpu.cpp:67-79 for the
// pairs with r^2 <= 64).
// ------------------------------------------------------------------------

// racc / rdon / Hramp (FG/CalculateGridmaps.cpp:84-197), in FP64 for the
// exact forms and FP32 for the plain form.  hb, rexp and disorder are
// wave-uniform in the cutoff kernel.
template <class T>
__host__ __device__ __forceinline__ void hbondAngular(int hb, int rexp, int disorder, bool isClosest, const T rv[3],
                                                      const T rv2[3], const T cHrv[3], T ux, T uy, T uz, T &racc,
                                                      T &rdon, T &Hramp)
{
    const T halfPi = T(3.14159265358979323846 / 2);
    racc = 1;
    rdon = 1;
    Hramp = 1;
    const T cosT = -(ux * rv[0] + uy * rv[1] + uz * rv[2]);
    if (hb == HB_D1) {
        racc = 0;
        if (cosT > 0) {
            T r = cosT;
            if (rexp == 4) r = r * r;
            if (rexp == 4 || rexp == 2) r = r * r;
            racc = r;
            if (!isClosest) {
                T th = acosClamped(cHrv[0] * rv[0] + cHrv[1] * rv[1] + cHrv[2] * rv[2]);
                Hramp = T(0.5) - T(0.5) * cos_(th * T(120.0 / 90.0));
            }
        }
    } else if (hb == HB_A1) {
        rdon = 0;
        if (cosT > 0) rdon = cosT * cosT;
    } else if (hb == HB_A2) {
        rdon = 0;
        if (disorder) {
            racc = 0;
            T th = acosClamped(cosT);
            if (th <= T(1.24791) + halfPi) {
                T c = cos_(th - T(1.24791));
                c = c * c;
                rdon = c * c;
                racc = rdon;
            }
        } else {
            T r = 0;
            if (cosT >= 0) {
                T cx = uy * rv2[2] - uz * rv2[1];
                T cy = uz * rv2[0] - ux * rv2[2];
                T cz = ux * rv2[1] - uy * rv2[0];
                T rd2 = cx * cx + cy * cy + cz * cz;
                if (rd2 < T(1.0E-6)) rd2 = T(1.0E-6);
                T inv = 1 / sqrt_(rd2);
                T ti = fabs(acosClamped(inv * (cx * rv[0] + cy * rv[1] + cz * rv[2])) - halfPi);
                r = T(0.9) + T(0.1) * sin_(ti + ti);
            } else if (cosT >= T(-0.34202)) {
                T t = T(0.116978) - cosT * cosT;
                r = T(562.25) * (t * t * t);
            }
            T t0 = halfPi - acosClamped(ux * rv2[0] + uy * rv2[1] + uz * rv2[2]);
            r *= cos_(t0);
            rdon = r;
        }
    }
}

// hbondAngular for one (point, bonded atom) pair, out of line.  The FP64
// acos/cos/sin it evaluates need far more registers than the pair loop; as a
// call their registers are not added to the loop's, so the loop keeps the
// occupancy of its plain terms and only pairs of bonded donors/acceptors pay
// for the call.
struct AngTerms {
    double racc, rdon, Hramp;
};
__device__ __noinline__ AngTerms pointAngular(const CutAtom &A, bool isClosest, double c0, double c1, double c2,
                                              double ux, double uy, double uz)
{
    const double rv[3] = {A.rv[0], A.rv[1], A.rv[2]}, rv2[3] = {A.rv2[0], A.rv2[1], A.rv2[2]};
    const double cHrv[3] = {c0, c1, c2};
    AngTerms t;
    hbondAngular(A.hbond, A.rexp, A.flags & kFlagDisorder, isClosest, rv, rv2, cHrv, ux, uy, uz, t.racc, t.rdon,
                 t.Hramp);
    return t;
}

// FP64 1/sqrt(x) for the cutoff kernel's pairs (1e-24 <= x <= e2Hi, a hair
// above 64): the hardware estimate, v_rsq_f64, measured at most 2^-24.2
// relative on 4 M inputs in [1e-6, 64] on MI355X.  Its continuous use (the
// near-field e term, w * min(1/r, 2) * eps) needs no more (a relative
// 2^-24 per term against the compensated form's 1e-3 kcal/mol bound);
// table-bin decisions made from it are guarded (see binGuarded).
__device__ __forceinline__ double rsqrtFast(double x) { return __builtin_amdgcn_rsq(x); }

// floor(100 r) for 0 <= r <= 8 (bins 0..800: no clamp) from an approximation
// r_a within 2^-23 relative of the exact r (100 r_a within 1e-4 of 100 r):
// equal to the reference's angstromToIndex unless 100 r_a lies within 2e-4
// of an integer, in which case `edge` is set and the caller recomputes r in
// the reference's exact operation order.
__device__ __forceinline__ int binGuarded(double r, bool &edge)
{
    const double t = r * 100.0;
    edge = __builtin_fabs(__builtin_amdgcn_fract(t) - 0.5) > 0.5 - 2e-4;
    return int(t);
}

// Initial value of a covalent map at (gx, gy, gz) (FG/main.cpp:39-46).
__device__ __forceinline__ double covalentValue(const CutoffArgs &a, double gx, double gy, double gz)
{
    const double dx = a.covPoint[0] - gx, dy = a.covPoint[1] - gy, dz = a.covPoint[2] - gz;
    double r = (dx * dx + dy * dy + dz * dz) * a.covHwSqInv;
    if (r < 1.0e-12) r = 1.0e-12;  // APPROX_ZERO^2
    return a.covBarrier * (1 - exp(-0.69314718055994529 * r));
}

constexpr int kChunk = 32;     // candidate atoms staged in LDS per round
constexpr int kRowBins = kPairBins;  // table bins of pairs with r <= 8 A: floor(100 r) <= 800
constexpr int kMaxTypes = 32;  // receptor types the type-sorted lists support

// LDS of one cutoff workgroup.  kTyped: the brick's candidates are sorted by
// receptor type and the pair-table entries of the current type for the first
// MM/2 maps (bins 0..800) are staged here; the other maps and the
// desolvation value are read through the vector L1.  The per-lane table
// reads of the pair loop (one bin per lane) are gathers either way; splitting
// them keeps the LDS and the L1 busy at the same time.
constexpr int ldsMapsOf(int MM) { return MM / 2 > 4 ? 4 : MM / 2 > 0 ? MM / 2 : 1; }
// Doubles per row of the compact table (CutoffArgs::pairHi) of a typed kernel
// compiled for MM maps: its maps past the LDS ones, padded to an even count.
constexpr int hiStrideOf(int MM) { return (MM - ldsMapsOf(MM) + 1) / 2 * 2; }
template <int MM>
constexpr int ldsMaps()
{
    return ldsMapsOf(MM);
}
template <int MM, bool kAffinity, bool kTyped>
struct CutoffLds {
    CutAtom atom[kChunk];
    double vs[kAffinity ? kChunk * 2 : 1];
    int32_t idx[kChunk];
    double row[kTyped ? kRowBins * ldsMaps<MM>() : 1];
    double eps[kTyped ? kRowBins : 1];
};

// Waves per SIMD a cutoff kernel is register-allocated for: the workgroups
// (one wave per SIMD each) that its LDS admits in a CU's 160 KB, at most 5
// (96 VGPRs) for up to 4 maps and 2 min/max maps, else at most 4 (128 VGPRs;
// the larger loops spill at 96).  Left to itself the compiler schedules the
// unrolled map loop for fewer waves than the LDS allows (up to 163 VGPRs
// with 8 maps).
template <int MM, int KH, bool kAffinity, bool kTyped>
constexpr int cutWaves()
{
    constexpr int lds = int((160 * 1024) / sizeof(CutoffLds<MM, kAffinity, kTyped>));
    constexpr int cap = (MM <= 4 && KH <= 2) ? 5 : 4;
    return lds < cap ? lds : cap;
}

// One brick of the cutoff kernel; L is the workgroup's LDS.
// kPointAng: the job has atoms whose directional terms depend on the point
// (bonded donors/acceptors); without them the FP64 transcendental code and its
// registers are left out.
// kCompact (typed kernel): the maps not in LDS come from the compact table
// (CutoffArgs::pairHi).  Its gathers touch half as many cache lines, which
// pays where the kernel waits on them (a single job's cutoff phase); in a
// batch, where the cutoff kernel shares the VALU with the e-kernel for most
// of its run, its extra address arithmetic costs more than that saves, and the
// rows of all maps are read instead.
template <int MM, int KH, bool kAffinity, bool kPointAng, bool kTyped, bool kCompact = false>
__device__ __forceinline__ void cutoffBrick(const CutoffArgs &a, int b, CutoffLds<MM, kAffinity, kTyped> &L)
{
    // the waves of a cutoff phase that carries a large share of the work
    // issue before the e-kernel's when both are ready (see CutoffArgs::prio)
    if (a.prio) __builtin_amdgcn_s_setprio(1);
    CutAtom *sAtom = L.atom;
    double *sVS = L.vs;
    int32_t *sIdx = L.idx;
    const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
    // wave = 4x4x4 cube of points, workgroup = 2x2x1 waves
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int ix = bx * kCutBX + (w & 1) * 4 + (l & 3);
    const int iy = by * kCutBY + (w >> 1) * 4 + ((l >> 2) & 3);
    const int iz = bz * kCutBZ + (l >> 4);
    const bool valid = ix < a.g.nx && iy < a.g.ny && iz < a.g.nz;
    // a lane past the grid sits 1e30 A away: every pair test rejects it
    const double gx = valid ? (ix - a.g.dx2) * a.g.spacing : 1e30;
    const double gy = (iy - a.g.dy2) * a.g.spacing;
    const double gz = (iz - a.g.dz2) * a.g.spacing;
    const int64_t p = valid ? (int64_t(iz) * a.g.ny + iy) * a.g.nx + ix : 0;
    const int M = a.M;
    // a wave with no point in the grid (half of a partial brick column) only
    // takes part in the staging and its barriers
    const bool live = __ballot(valid) != 0;

    if (kTyped) {
        // rows of the brick-independent tables for r <= 8 A (the barrier of
        // the first chunk orders these stores before their reads)
        __syncthreads();  // the previous brick of a persistent workgroup is done with them
        for (int q = tid; q < kRowBins; q += kCutThreads) L.eps[q] = a.eps[q];
    }

    // pass 1: nearest DS/D1 atom, lowest index on ties (FG/CalculateGridmaps.cpp:65-82)
    int cH = 0;
    double cHrv[3] = {0, 0, 0};
    if (kAffinity) {
        double rmin = 999999.0 * 999999.0;
        // uniform reads of read-only inputs: scalar loads
        const FG_CONST int32_t *hStart = constView(a.hStart), *hList = constView(a.hList);
        const FG_CONST CutAtom *atoms = constView(a.atoms);
        const int h0 = hStart[b], h1 = live ? hStart[b + 1] : h0;
        for (int c = h0; c < h1; ++c) {
            const int ia = uniform(hList[c]);
            const double ax = atoms[ia].x, ay = atoms[ia].y, az = atoms[ia].z;
            double dx = ax - gx, dy = ay - gy, dz = az - gz;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < rmin || (r2 == rmin && ia < cH)) {
                rmin = r2;
                cH = ia;
            }
        }
        cHrv[0] = atoms[cH].rv[0];
        cHrv[1] = atoms[cH].rv[1];
        cHrv[2] = atoms[cH].rv[2];
    }

    double e[MM], hmin[KH > 0 ? KH : 1], hmax[KH > 0 ? KH : 1];
    unsigned flags = 0;
    const double cov = (kAffinity && a.covOn && !a.covalent) ? covalentValue(a, gx, gy, gz) : 0.0;
#pragma unroll
    for (int m = 0; m < MM; ++m)
        e[m] = (kAffinity && m < M && a.mp[m].isCovalent)
                   ? (a.covalent ? a.covalent[int64_t(a.mp[m].out) * a.P + p] : cov)
                   : 0.0;
#pragma unroll
    for (int k = 0; k < (KH > 0 ? KH : 1); ++k) {
        hmin[k] = 999999;
        hmax[k] = -999999;
    }
    double dmap = 0, enear = 0, sv = 0, sq = 0;
    const double cut2 = a.doElecNear ? a.e2Hi : 64.0;  // e2Hi > 64

    // candidates: one run over the brick's list, or one run per receptor
    // type over the type-sorted list with that type's rows in LDS, in chunks
    // of kChunk atoms staged in LDS.  Each chunk's list entries and atom
    // records are loaded into registers while the chunk before it is
    // computed (one 16-byte piece of a record per lane, lane 8c also atom
    // c's desolvation factors), so their latency hides behind that work.
    const int nRuns = kTyped ? a.T : 1;
    auto runStart = [&](int r) {
        return kTyped ? constView(a.typeStart)[b * (a.T + 1) + r] : constView(a.brickStart)[b + r];
    };
    // the chunk after (run, cb) of a run ending at c1; run == nRuns: none
    int run = -1, cb = 0, c1 = 0;
    auto advance = [&]() {
        if (run >= 0 && cb + kChunk < c1) {
            cb += kChunk;
            return;
        }
        for (++run; run < nRuns; ++run) {
            cb = runStart(run);
            c1 = runStart(run + 1);
            if (cb < c1) return;
        }
    };
    float4 pRec = make_float4(0.f, 0.f, 0.f, 0.f);
    double2 pVS = make_double2(0.0, 0.0);
    int pIdx = -1;
    auto prefetch = [&]() {
        const int nc = min(kChunk, c1 - cb), at = tid >> 3;
        pIdx = at < nc ? a.brickList[cb + at] : -1;
        if (pIdx >= 0) {
            pRec = reinterpret_cast<const float4 *>(a.atoms + pIdx)[tid & 7];
            if (kAffinity && (tid & 7) == 0) pVS = reinterpret_cast<const double2 *>(a.vs)[pIdx];
        }
    };
    advance();
    if (run < nRuns) prefetch();
    int rowsOf = -1;  // the receptor type whose rows are in LDS
    while (run < nRuns) {
        const int nc = min(kChunk, c1 - cb);
        __syncthreads();  // every lane is done with the previous chunk (and type's rows)
        if (kTyped && run != rowsOf) {
            const double *src = a.pair + unsigned(run * kPairBins * M);
            constexpr int ML = ldsMaps<MM>();
            for (int q = tid; q < kRowBins * ML; q += kCutThreads) {
                const int bin = q / ML, m = q - bin * ML;
                L.row[q] = m < M ? src[bin * M + m] : 0.0;
            }
            rowsOf = run;
        }
        // stage this chunk, request the next one
        if (pIdx >= 0) {
            reinterpret_cast<float4 *>(sAtom)[tid] = pRec;
            if ((tid & 7) == 0) {
                sIdx[tid >> 3] = pIdx;
                if (kAffinity) reinterpret_cast<double2 *>(sVS)[tid >> 3] = pVS;
            }
        }
        advance();
        if (run < nRuns) prefetch();
        __syncthreads();
        if (live) {
            for (int c = 0; c < nc; ++c) {
                const CutAtom &A = sAtom[c];
                const double dx = A.x - gx, dy = A.y - gy, dz = A.z - gz;
                const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
                double rSq = dx2 + dy2 + dz2;
                // affinity and d maps: r^2 <= 64; e-map: the pairs the
                // far-field kernel leaves (FP32 bin < kNearBins), certain
                // up to e2Lo (< 64) and decided by its own FP32 arithmetic up
                // to e2Hi.  One test on the path most candidates take; the
                // band's pairs (rare) branch off.
                if (rSq > cut2) continue;
                double eW = A.w;  // the pair's e-map weight (0: the far field's pair)
                if (a.doElecNear && rSq > a.e2Lo) {
                    const float s100 = a.spacing100;
                    const float r2 = farR2(float(A.x * 100.0), float(A.y * 100.0), float(A.z * 100.0),
                                           float(ix - a.g.dx2) * s100, float(iy - a.g.dy2) * s100,
                                           float(iz - a.g.dz2) * s100);
                    const bool farOwns = farBinBits(r2) >= kFloorBits + unsigned(kNearBins);
                    if (rSq > 64.0) {
                        // past the affinity cutoff: the e term alone, if the far field left it
                        if (!farOwns) {
                            const double y = rsqrtFast(rSq);
                            bool edgeE;
                            int iE = binGuarded(rSq * y, edgeE);
                            if (edgeE) iE = binOf(sqrt(dx2 + (dy2 + dz2)));
                            enear = __builtin_fma(A.w * __builtin_fmin(y, 2.0), kTyped ? L.eps[iE] : a.eps[iE],
                                                  enear);
                        }
                        continue;
                    }
                    if (farOwns) eW = 0.0;
                }
                // from here r^2 <= 64
                rSq = __builtin_fmax(rSq, 1.0E-24);  // rSq >= 0, never NaN
                // guarded fast geometry (see binGuarded)
                const double yinv = rsqrtFast(rSq);
                bool edge;
                int idx = binGuarded(rSq * yinv, edge);
                int idxE = idx;
                if (edge) {
                    // exact: the cutoff loop's 1/(1/sqrt(r^2)) and the e-map's
                    // sqrt(dx^2 + (dy^2 + dz^2))  (FG/CalculateGridmaps.cpp:286-289,
                    // FG/electrostatics/Cpu.cpp:70-75)
                    const double invR = 1 / sqrt(rSq);
                    idx = binOf(1 / invR);
                    idxE = binOf(sqrt(dx2 + (dy2 + dz2)));
                }
                // r <= 8 A: both bins are <= 800
                if (a.doElecNear)
                    enear = __builtin_fma(eW * __builtin_fmin(yinv, 2.0), kTyped ? L.eps[idxE] : a.eps[idxE], enear);
                if (!kAffinity) continue;
                // atom fields are the same in every lane: keep them in scalar registers
                const int fl = uniform(A.flags);
                const int dis = fl & kFlagDisorder, kind = (fl >> 1) & 3, ahb = (fl >> 3) & 7, type = fl >> 8;
                if (type == a.hydrogen && dis) continue;
                double racc = 1, rdon = 1, Hramp = 1;
                if (kind == kAngConst) {
                    racc = A.ang[0];
                    rdon = A.ang[1];
                    Hramp = A.ang[2];
                } else if (kPointAng && kind == kAngPoint) {
                    // directional terms use the unit vector exactly as the reference forms it
                    const double invR = 1 / sqrt(rSq);
                    const AngTerms t = pointAngular(A, uniform(sIdx[c]) == cH, cHrv[0], cHrv[1], cHrv[2], dx * invR,
                                                    dy * invR, dz * invR);
                    racc = t.racc;
                    rdon = t.rdon;
                    Hramp = t.Hramp;
                }
                const double S = a.desolv[idx];
                // 32-bit offsets (job creation bounds the table below 2^31 entries)
                const unsigned rowIdx = unsigned(type * kPairBins + idx);
                const unsigned rowOff = __umul24(rowIdx, unsigned(M));
                const double *grow = a.pair + rowOff;
                const double *lrow = kTyped ? &L.row[idx * ldsMaps<MM>()] : grow;
                const double *ghi = kCompact ? a.pairHi + rowIdx * unsigned(hiStrideOf(MM)) : grow;
                // map m's entry: LDS for the first MM/2 maps of the typed kernel
                // (at most 4), the compact table (or the full row) for the others
#define FG_ROW(m)                                                                        \
    ((kTyped && (m) < ldsMaps<MM>()) ? lrow[(m)]                                          \
                                     : (kCompact ? ghi[(m) - ldsMaps<MM>()] : grow[(m)]))
                const double *row = grow;
                const double *rowH =
                    (dis && a.hydrogen >= 0)
                        ? a.pair + unsigned((a.hydrogen * kPairBins + (idx - 110 > 0 ? idx - 110 : 0)) * M)
                        : row;
                const double vol = sVS[2 * c], solQ = sVS[2 * c + 1];
                // one addend per map and a single update of every e[m] after
                // the two paths join, so the accumulators never sit in
                // divergent branches (a map that keeps a min/max, or a
                // covalent map, adds 0: e + 0 == e)
                double add[MM];
                if (((a.addOnlyClasses >> ahb) & 1) && kTyped) {
                    // every map adds the table value for this atom (no H-bond
                    // term, no covalent map): no per-map control flow.  Maps
                    // m >= M (the kernel is compiled for 4, 6, 8 or 16 maps)
                    // accumulate values that are never stored.
#pragma unroll
                    for (int m = 0; m < MM; ++m) add[m] = FG_ROW(m);
                } else {
                const unsigned acts = unsigned(uniform(A.acts));
#pragma unroll
                for (int m = 0; m < MM; ++m) {
                    add[m] = 0.0;
                    if (m >= M || ((a.covMask >> m) & 1)) continue;
                    const int act = int(acts >> (2 * m)) & 3;
                    const double E = FG_ROW(m);
                    add[m] = E;
                    if (act != kActAdd) {
                        // min(max(E / 100, 0), 1), formed on the host per table entry
                        const double rsph = a.rsph[rowOff + m];
                        if (act == kActDonor) {
                            add[m] = rowH[m] * Hramp * (racc + (1 - racc) * rsph);
                        } else {
                            add[m] = 0.0;
                            if (m < KH) {
                                const double f = act == kActMinRacc ? racc : rdon;
                                const double v = E * (f + (1 - f) * rsph);
                                hmin[m < KH ? m : 0] = dmin_(hmin[m < KH ? m : 0], v);
                                hmax[m < KH ? m : 0] = dmax_(hmax[m < KH ? m : 0], v);
                                flags |= 1u << m;
                            }
                        }
                    }
                }
                }
#pragma unroll
                for (int m = 0; m < MM; ++m) e[m] += add[m];
                sv = __builtin_fma(vol, S, sv);
                sq = __builtin_fma(solQ, S, sq);
                dmap = __builtin_fma(A.qvol, S, dmap);
#undef FG_ROW
            }
        }
    }
    if (!valid) return;
    if (kAffinity) {
#pragma unroll
        for (int m = 0; m < MM; ++m) {
            if (m >= M) continue;
            double v = e[m];
            // the desolvation term summed over atoms with the probe's factors
            // taken out: solProbe * sum(vol S) + volProbe * sum(solQ S)
            if (!a.mp[m].isCovalent) v += a.mp[m].solProbe * sv + a.mp[m].volProbe * sq;
            if (m < KH && (flags & (1u << m))) v += hmin[m < KH ? m : 0] + hmax[m < KH ? m : 0];
            a.maps[int64_t(a.mp[m].out) * a.P + p] = v;
        }
        a.maps[int64_t(M + 1) * a.P + p] = dmap;
    }
    if (a.doElecNear) a.maps[int64_t(M) * a.P + p] = enear;
}

template <int MM, int KH, bool kAffinity, bool kPointAng, bool kTyped>
__global__ __launch_bounds__(kCutThreads)
__attribute__((amdgpu_waves_per_eu(cutWaves<MM, KH, kAffinity, kTyped>()))) void k_cutoff(CutoffArgs a)
{
    __shared__ CutoffLds<MM, kAffinity, kTyped> L;
    // XCD-aware brick order: consecutive workgroups go to different XCDs, so
    // hand each XCD a contiguous run of bricks (neighbours share atoms in L2)
    // (bijection on the first 8*floor(nb/8) workgroups, identity on the tail)
    const int nb = a.nbx * a.nby * a.nbz;
    const int full = (nb / 8) * 8, g = int(blockIdx.x);
    const int b = g < full ? (g % 8) * (full / 8) + g / 8 : g;
    cutoffBrick<MM, KH, kAffinity, kPointAng, kTyped, kTyped>(a, b, L);
}

// ------------------------------------------------------------------------
// Floating grid (FG/main.cpp:54-79): float(sqrt(min over atoms of |a - p|^2)).
// The nearest atom of a point within 8 A is in its brick's candidate list,
// so the list decides every such point exactly; a wave with a point farther
// than 8 A from every listed atom scans all atoms.
// ------------------------------------------------------------------------
__device__ __forceinline__ void fmapBrick(const CutoffArgs &a, int b)
{
    const int N = a.N;
    float *fmap = a.fmap;
    const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int ix = bx * kCutBX + (w & 1) * 4 + (l & 3);
    const int iy = by * kCutBY + (w >> 1) * 4 + ((l >> 2) & 3);
    const int iz = bz * kCutBZ + (l >> 4);
    const bool valid = ix < a.g.nx && iy < a.g.ny && iz < a.g.nz;
    const double gx = (ix - a.g.dx2) * a.g.spacing;
    const double gy = (iy - a.g.dy2) * a.g.spacing;
    const double gz = (iz - a.g.dz2) * a.g.spacing;
    double m2 = 1.0e24;  // BIG^2
    for (int c = a.brickStart[b]; c < a.brickStart[b + 1]; ++c) {
        const CutAtom &A = a.atoms[uniform(a.brickList[c])];
        const double dx = A.x - gx, dy = A.y - gy, dz = A.z - gz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < m2) m2 = d2;
    }
    if (__ballot(valid && m2 > 64.0))
        for (int ia = 0; ia < N; ++ia) {
            const CutAtom &A = a.atoms[ia];
            const double dx = A.x - gx, dy = A.y - gy, dz = A.z - gz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < m2) m2 = d2;
        }
    if (valid) fmap[(int64_t(iz) * a.g.ny + iy) * a.g.nx + ix] = float(sqrt(m2));
}

// Covalent map values alone (fg_covalent_map), one point per lane.
__global__ __launch_bounds__(256) void k_covalent(CutoffArgs a, double *out)
{
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= a.P) return;
    const int ix = int(p % a.g.nx), iy = int((p / a.g.nx) % a.g.ny), iz = int(p / (int64_t(a.g.nx) * a.g.ny));
    out[p] = covalentValue(a, (ix - a.g.dx2) * a.g.spacing, (iy - a.g.dy2) * a.g.spacing, (iz - a.g.dz2) * a.g.spacing);
}

__global__ __launch_bounds__(kCutThreads) void k_fmap(CutoffArgs a)
{
    fmapBrick(a, int(blockIdx.x));
}

// Stable counting sort of brick b's candidate list by receptor type: list ->
// sorted, typeStart[b*(T+1) + t] = first entry of type t (typeStart[..+T] =
// end).  Ranks within a chunk come from wave ballots, so the order inside a
// type is the list order on every run.
__device__ __forceinline__ void typeSortBrick(const ListArgs &a, int b, int32_t *cnt, int32_t *runStart,
                                              int32_t *wcnt)
{
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int T = a.T;
    const int s0 = a.start[b], s1 = a.start[b + 1];
    if (tid < T) cnt[tid] = 0;
    __syncthreads();
    for (int i = s0 + tid; i < s1; i += kListThreads) atomicAdd(&cnt[a.atomType[a.list[i]]], 1);
    __syncthreads();
    if (tid == 0) {
        int run = s0;
        for (int t = 0; t < T; ++t) {
            runStart[t] = run;
            a.typeStart[b * (T + 1) + t] = run;
            run += cnt[t];
        }
        a.typeStart[b * (T + 1) + T] = s1;
    }
    __syncthreads();
    for (int base = s0; base < s1; base += kListThreads) {
        const int i = base + tid;
        const int ia = i < s1 ? a.list[i] : -1;
        const int t = ia >= 0 ? a.atomType[ia] : -1;
        int rank = 0;
        for (int tt = 0; tt < T; ++tt) {
            const uint64_t m = __ballot(t == tt);
            if (t == tt)
                rank = int(__builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u)));
            if (lane == 0) wcnt[w * kMaxTypes + tt] = int(__popcll(m));
        }
        __syncthreads();
        if (ia >= 0) {
            int pos = runStart[t] + rank;
            for (int v = 0; v < w; ++v) pos += wcnt[v * kMaxTypes + t];
            a.sorted[pos] = ia;
        }
        __syncthreads();
        if (tid < T) runStart[tid] += wcnt[tid] + wcnt[kMaxTypes + tid] + wcnt[2 * kMaxTypes + tid] +
                                      wcnt[3 * kMaxTypes + tid];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kListThreads) void k_type_sort(ListArgs a)
{
    __shared__ int32_t cnt[kMaxTypes], runStart[kMaxTypes], wcnt[4 * kMaxTypes];
    typeSortBrick(a, int(blockIdx.x), cnt, runStart, wcnt);
}

// ------------------------------------------------------------------------
// Far-field electrostatics in FP32 (compensated) or all-pairs FP32.
// Coordinates are in units of 0.01 A, so 100*r = r2 * rsq(r2) indexes the
// dielectric table directly and w is stored as 100*w.
// ------------------------------------------------------------------------

constexpr int kFold = 64;  // atoms per FP32 partial sum of the compensated form

// Loads the dielectric table into LDS (all lanes of the workgroup), with
// bins below `zeroBelow` read as 0: the compensated form's near pairs (bins
// below kNearBins, which its far loops never reach) then add nothing.
__device__ __forceinline__ void loadEps(float *eps, const float *src, int n, int zeroBelow)
{
    for (int i = threadIdx.x; i < n; i += kElecThreads) eps[i] = i < zeroBelow ? 0.f : src[i];
}

// One work item of the electrostatic map: 256 lanes, each owning PX points
// along x times PZ consecutive points along z.  2D bricks (bxShift > 0):
// 1 << bxShift lanes along x, the rest along y, a lane's x points 1 <<
// bxShift apart.  Linear items (bxShift 0): 256 consecutive units of PX
// points 8 apart along x, adjacent in the last tile (ElecArgs); `region` 1
// is the partial
// slab at the top of the grid, whose units stand zStride planes above the
// last full slab.  `eps` is the table in LDS.  The FP32 arithmetic of a
// (point, atom) pair does not depend on the shape, and the compensated form
// folds its FP32 partial into the two-float sum after every kFold atoms, so
// every shape gives the same bits.
template <int PX, int PZ, bool kComp, bool kClamp>
__device__ __forceinline__ void elecBrick(const ElecArgs &a, int b, const float *eps, int zStride = PZ,
                                          int region = 0)
{
    constexpr int NP = PX * PZ;
    const int tid = threadIdx.x;
    const int sh = a.bxShift;
    // (ix, iy, iz0): the lane's first point, for the stores; (cx, cy, cz):
    // the points it computes.  A padding lane computes a copy of a real
    // point (and stores nothing), so every lane's points lie in its wave's
    // box of real points and the far tests against that box hold for all.
    int ix, iy, iz0, xs = 1;  // xs: x distance between the lane's points
    bool real = true;
    int cx[PX], cy, cz[PZ];
    if (sh == 0) {
        int u, nU, slab0;
        if (region == 0) {
            u = b * kElecThreads + tid;
            nU = a.slabs * a.units;
            slab0 = 0;
        } else {
            u = (b - a.itemsA) * kElecThreads + tid;
            nU = a.units;
            slab0 = a.slabs;
        }
        real = u < nU;
        if (!real) u = nU - 1;
        const int s = slab0 + u / a.units, w = u - (u / a.units) * a.units;
        // within the slab: tile row ty (8 rows, fewer in the last), tile tx
        // in it (8 x h units, fewer lane columns in the last), lane (lx, ly)
        const int ty = w / (8 * a.ux), r = w - ty * 8 * a.ux, h = min(8, a.g.ny - 8 * ty);
        const int tx = r / (8 * h), s2 = r - tx * 8 * h;
        const bool lastTile = tx == a.tilesX - 1;
        const int wt = lastTile ? a.lastLanes : 8;
        const int ly = s2 / wt, lx = s2 - ly * wt;
        // two points per lane: 8 apart, or adjacent in the last tile, which
        // may be narrower than 16 points
        xs = PX == 2 && lastTile ? 1 : 8;
        ix = tx * 8 * PX + (xs == 1 ? 2 * lx : lx);
        iy = 8 * ty + ly;
        iz0 = s * zStride;
#pragma unroll
        for (int i = 0; i < PX; ++i) cx[i] = min(ix + xs * i, a.g.nx - 1);
        cy = iy;
    } else {
        // a lane's PX points are bxl apart along x, so that each table read
        // of a wave (one point per lane) covers a compact bxl x (64 / bxl)
        // patch: nearby points read nearby bins, fewer LDS bank conflicts
        const int bxl = 1 << sh, byl = kElecThreads >> sh;
        const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
        const int ix0 = bx * bxl * PX, iy0 = by * byl;
        iz0 = bz * zStride;
        ix = ix0 + (tid & (bxl - 1));
        iy = iy0 + (tid >> sh);
        xs = bxl;
        real = iy < a.g.ny && ix < a.g.nx;
        // past the grid: the brick's last point on that axis
        const int bx1 = min(ix0 + bxl * PX - 1, a.g.nx - 1), by1 = min(iy0 + byl - 1, a.g.ny - 1);
#pragma unroll
        for (int i = 0; i < PX; ++i) cx[i] = min(ix + i * bxl, bx1);
        cy = min(iy, by1);
    }
    // a wave of padding lanes only has nothing to do (no barrier follows)
    if (!__ballot(real)) return;
    const int bz1 = min(iz0 + PZ - 1, a.g.nz - 1);
#pragma unroll
    for (int k = 0; k < PZ; ++k) cz[k] = min(iz0 + k, bz1);
    const float s100 = a.spacing100;
    const float gy = float(cy - a.g.dy2) * s100;
    float gx[PX], gz[PZ];
#pragma unroll
    for (int i = 0; i < PX; ++i) gx[i] = float(cx[i] - a.g.dx2) * s100;
#pragma unroll
    for (int k = 0; k < PZ; ++k) gz[k] = float(cz[k] - a.g.dz2) * s100;
    // the wave's box (its 64 lanes' points): tiles and sub-tiles are tested
    // against it, so a wave far from a sub-tile skips the per-pair
    // exclusion of the near loop
    float wlo[3], whi[3];
    {
        float v[6] = {gx[0], gy, gz[0], gx[PX - 1], gy, gz[PZ - 1]};
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = fminf(v[k], __shfl_xor(v[k], o));
                v[3 + k] = fmaxf(v[3 + k], __shfl_xor(v[3 + k], o));
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            wlo[k] = __int_as_float(uniform(__float_as_int(v[k])));
            whi[k] = __int_as_float(uniform(__float_as_int(v[3 + k])));
        }
    }
    const unsigned clampMax = unsigned(a.clampMax);
    const unsigned clampBits = kFloorBits + clampMax;
    // byte offset that turns a biased bit pattern into a table address; read
    // from the arguments so that it stays in a scalar register
    const unsigned biasBytes = unsigned(a.floorBitsNeg4);
    const char *epsBytes = reinterpret_cast<const char *>(eps);

    // atom tiles and their boxes: read-only, uniform loads
    const FG_CONST float *tiles = constView(reinterpret_cast<const float *>(a.tiles));
    const FG_CONST float *subBox = constView(a.subBox);

    float hi[NP], lo[NP], part[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) hi[q] = lo[q] = part[q] = 0.f;

    // Atoms come kGroup at a time in one scalar load: scalar loads may
    // return out of order, so a wave that uses one waits for every LDS read
    // it has in flight, once per group.  Groups of 2 (8 points per lane) or
    // 4 (16 points, whose per-atom work is twice as long); ranges are
    // multiples of kSubTile atoms.
    constexpr int kGroup = NP >= 16 ? 4 : 2;
    auto forAtoms = [&](int c0, int c1, auto &&fn) {
        for (int j = c0; j < c1; j += kGroup) {
            float v[4 * kGroup];
#pragma unroll
            for (int i = 0; i < 4 * kGroup; ++i) v[i] = tiles[4 * j + i];
#pragma unroll
            for (int t = 0; t < kGroup; ++t) fn(make_float4(v[4 * t], v[4 * t + 1], v[4 * t + 2], v[4 * t + 3]));
        }
    };
    // Every pair of a tile: the compensated form's table reads 0 at FP32
    // bins below kNearBins, so the pairs within 8 A add nothing (they are the
    // cutoff kernel's), and an atom on a grid point has a finite 1/r
    // (kTinyR2).  Clamp: bins past the table's end are clamped to its last
    // entry (the clamped kernel variant, for tiles that may reach beyond the
    // table).
    auto farAtoms = [&](int c0, int c1, auto clampTag) {
        constexpr bool kClampHere = kClamp && decltype(clampTag)::value;
        forAtoms(c0, c1, [&](const float4 at) {
            const float dy = at.y - gy;
            const float dy2 = __builtin_fmaf(dy, dy, kTinyR2);
            float dxy2[PX], dz[PZ];
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                const float dx = at.x - gx[i];
                dxy2[i] = __builtin_fmaf(dx, dx, dy2);
            }
#pragma unroll
            for (int k = 0; k < PZ; ++k) dz[k] = at.z - gz[k];
            float ri[NP], ev[NP];
            unsigned bin[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const float r2 = __builtin_fmaf(dz[q % PZ], dz[q % PZ], dxy2[q / PZ]);
                ri[q] = __builtin_amdgcn_rsqf(r2);
                if (kComp) {
                    // floor(100 r) in one operation: r2 * ri + (2^23 - 1/2)
                    // lands in [2^23, 2^24), where the FP32 spacing is 1,
                    // so its bit pattern is 2^23's plus round(100 r - 1/2)
                    // (far pairs have 100 r > 800)
                    bin[q] = __float_as_uint(__builtin_fmaf(r2, ri[q], kFloorBias));
                    if (kClampHere) bin[q] = min(bin[q], clampBits);
                } else {
                    bin[q] = unsigned(r2 * ri[q]);
                    if (kClampHere) bin[q] = min(bin[q], clampMax);
                }
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                // eps[bits - 2^23's bits] as one shift-add with a uniform byte offset
                ev[q] = kComp ? *reinterpret_cast<const float *>(epsBytes + ((bin[q] << 2) + biasBytes))
                              : eps[bin[q]];
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                // min(1/r, 2) in scaled units; the far-field pairs of the
                // compensated form are all beyond 8 A
                const float rc = kComp ? ri[q] : fminf(ri[q], 0.02f);
                part[q] = __builtin_fmaf(at.w * rc, ev[q], part[q]);
            }
        });
    };
    // in the table's constant tail: every lane's pair has a bin from which on
    // the FP32 table holds its last entry (100 r past ~5842 for the
    // distance-dependent dielectric), so the bin, its address and the read
    // drop out; the rest is the far loop's arithmetic, so the bits are the same
    const float tLast = eps[kTableLen - 1];
    auto lastBinAtoms = [&](int c0, int c1) {
        forAtoms(c0, c1, [&](const float4 at) {
            const float dy = at.y - gy;
            const float dy2 = __builtin_fmaf(dy, dy, kTinyR2);
            float dxy2[PX], dz[PZ];
#pragma unroll
            for (int i = 0; i < PX; ++i) {
                const float dx = at.x - gx[i];
                dxy2[i] = __builtin_fmaf(dx, dx, dy2);
            }
#pragma unroll
            for (int k = 0; k < PZ; ++k) dz[k] = at.z - gz[k];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const float r2 = __builtin_fmaf(dz[q % PZ], dz[q % PZ], dxy2[q / PZ]);
                part[q] = __builtin_fmaf(at.w * __builtin_amdgcn_rsqf(r2), tLast, part[q]);
            }
        });
    };
    // squared shortest and longest distances between sub-tile q0's box and
    // the box (lo, hi)
    auto boxDist2 = [&](int q0, const float *lo, const float *hi, float &dmin2, float &dmax2) {
        const FG_CONST float *tb = subBox + 6 * (q0 / kSubTile);
        dmin2 = dmax2 = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float d = fmaxf(fmaxf(tb[k] - hi[k], lo[k] - tb[3 + k]), 0.f);
            const float e = fmaxf(tb[3 + k] - lo[k], hi[k] - tb[k]);
            dmin2 = __builtin_fmaf(d, d, dmin2);
            dmax2 = __builtin_fmaf(e, e, dmax2);
        }
    };
    using kYes = std::integral_constant<bool, true>;
    using kNo = std::integral_constant<bool, false>;

    for (int j0 = 0; j0 < a.nAtoms; j0 += a.tile) {
        const int j1 = min(j0 + a.tile, a.nAtoms);
        // the tile's class against the wave, from its sub-tile boxes: within
        // the table (1), possibly past the table's end (2, clamped variant
        // only), in the table's constant tail (3)
        int kind = 1;
        if (kComp) {
            bool last = true, inTable = true;
            for (int q = j0; q < j1; q += kSubTile) {
                float dmin2, dmax2;
                boxDist2(q, wlo, whi, dmin2, dmax2);
                last = last && dmin2 > a.lastCut2;
                inTable = inTable && dmax2 < a.tableCut2;
            }
            kind = uniform(last ? 3 : (kClamp && !inTable ? 2 : 1));
        }
        for (int c0 = j0; c0 < j1; c0 += kFold) {
            const int c1 = min(c0 + kFold, j1);
            if (kind == 3) {
                lastBinAtoms(c0, c1);
            } else if (kind == 1) {
                // the FP32 form (kComp false) clamps as its variant says
                if (kComp) farAtoms(c0, c1, kNo());
                else farAtoms(c0, c1, kYes());
            } else {
                farAtoms(c0, c1, kYes());
            }
            if (kComp && ((c1 & (kFold - 1)) == 0 || c1 == a.nAtoms)) {
                // fold the FP32 partial into the two-float accumulator (TwoSum)
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const float s = hi[q] + part[q];
                    const float bp = s - hi[q];
                    const float err = (hi[q] - (s - bp)) + (part[q] - bp);
                    hi[q] = s;
                    lo[q] += err;
                    part[q] = 0.f;
                }
            }
        }
    }
    if (!real) return;
#pragma unroll
    for (int i = 0; i < PX; ++i) {
        if (ix + i * xs >= a.g.nx) break;
#pragma unroll
        for (int k = 0; k < PZ; ++k) {
            const int iz = iz0 + k;
            if (iz >= a.g.nz) break;
            const int q = i * PZ + k;
            const int64_t p = (int64_t(iz) * a.g.ny + iy) * a.g.nx + ix + i * xs;
            if (kComp) a.far[p] = double(hi[q]) + double(lo[q]);
            else a.e[p] = double(part[q]);
        }
    }
}

// The compensated e-map: the near part k_cutoff wrote plus the far part
// k_elec wrote (the two kernels run concurrently into separate buffers).
__global__ __launch_bounds__(256) void k_add_far(int64_t P, const double *far, double *e)
{
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p < P) e[p] = e[p] + far[p];
}

// roundOutput (FG/autogrid.h:128-131) in FP64, as the writers apply it.
__device__ __forceinline__ double round3(double a)
{
    return __builtin_fabs(a) < 0.0005 ? 0.0 : floor(a * 1000 + 0.5) * 0.001;
}

// The maps as the writers print them, in integer thousandths: roundOutput's
// floor(1000 v + 0.5) (0 below 0.0005 in magnitude).  A value whose
// thousandths do not fit 32 bits sets *overflow (and stores 0).
__global__ __launch_bounds__(256) void k_round_milli(const double *maps, int64_t n, int32_t *out, int32_t *overflow)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double v = maps[i];
    const double k = __builtin_fabs(v) < 0.0005 ? 0.0 : floor(v * 1000 + 0.5);
    if (!(__builtin_fabs(k) <= 2147483647.0)) {
        *overflow = 1;
        out[i] = 0;
        return;
    }
    out[i] = int32_t(k);
}

// The job's summary: the minimum and maximum of each map's rounded values
// (GridMapList::calculateEnergyMinMax over the stored, rounded energies,
// FG/GridMap.cpp:214-224, starting from +-BIG = 1e12) and the receptor's total
// charge summed in atom order, the reference's own loop
// (FG/InputDataLoader.cpp:325).  Min and max are exact, so the reduction
// order does not matter: workgroup (c, m) of k_job_summary scans chunk c of
// map m into part[m][c], and k_job_summary_final combines the chunks.
// Workgroup (0, nMaps) sums the charges: 256 at a time staged in LDS and added
// in order by one lane, the next 256 loading meanwhile.
constexpr int kSumChunks = 64;  // workgroups per map

__global__ __launch_bounds__(256) void k_job_summary(const double *maps, int64_t P, int nMaps, const double *charge,
                                                      int N, double *part, double *out)
{
    const int m = int(blockIdx.y), c = int(blockIdx.x), tid = threadIdx.x;
    if (m == nMaps) {
        if (c != 0) return;
        __shared__ double sQ[256];
        double q = 0, next = tid < N ? charge[tid] : 0.0;
        for (int i0 = 0; i0 < N; i0 += 256) {
            __syncthreads();  // lane 0 is done with the previous 256
            sQ[tid] = next;
            __syncthreads();
            if (i0 + 256 + tid < N) next = charge[i0 + 256 + tid];
            if (tid == 0) {
                const int n = min(256, N - i0);
                for (int k = 0; k < n; ++k) q += sQ[k];
            }
        }
        if (tid == 0) out[2 * nMaps] = q;
        return;
    }
    __shared__ double sLo[4], sHi[4];
    const int64_t len = (P + kSumChunks - 1) / kSumChunks;
    const int64_t p0 = int64_t(c) * len, p1 = min(P, p0 + len);
    double lo = 1.0e12, hi = -1.0e12;
    const double *v = maps + int64_t(m) * P;
    for (int64_t p = p0 + tid; p < p1; p += 256) {
        const double r = round3(v[p]);
        lo = __builtin_fmin(lo, r);
        hi = __builtin_fmax(hi, r);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = __builtin_fmin(lo, __shfl_xor(lo, o));
        hi = __builtin_fmax(hi, __shfl_xor(hi, o));
    }
    if ((tid & 63) == 0) {
        sLo[tid >> 6] = lo;
        sHi[tid >> 6] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = __builtin_fmin(lo, sLo[w]);
            hi = __builtin_fmax(hi, sHi[w]);
        }
        part[2 * (int64_t(m) * kSumChunks + c)] = lo;
        part[2 * (int64_t(m) * kSumChunks + c) + 1] = hi;
    }
}

// out[2m], out[2m + 1]: map m's minimum and maximum over its chunks (one lane per map)
__global__ __launch_bounds__(256) void k_job_summary_final(const double *part, int nMaps, double *out)
{
    const int m = int(threadIdx.x);
    if (m >= nMaps) return;
    double lo = 1.0e12, hi = -1.0e12;
    for (int c = 0; c < kSumChunks; ++c) {
        lo = __builtin_fmin(lo, part[2 * (m * kSumChunks + c)]);
        hi = __builtin_fmax(hi, part[2 * (m * kSumChunks + c) + 1]);
    }
    out[2 * m] = lo;
    out[2 * m + 1] = hi;
}

// ------------------------------------------------------------------------
// Batches: many jobs in one launch.  Work items are (job, brick) pairs,
// numbered job after job; base[j] is the first item of job j and
// base[nJobs] the total.  The persistent kernels take items from a counter
// until none are left; the counter is cleared by k_fill_i32 before them.
// ------------------------------------------------------------------------
__device__ __forceinline__ int jobOfItem(const int32_t *base, int nJobs, int item)
{
    int lo = 0, hi = nJobs - 1;  // last job with base <= item
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Next work item of the workgroup (every lane gets the same value).
__device__ __forceinline__ int nextItem(int *counter, int *slot)
{
    __syncthreads();  // every lane has read the previous item
    if (threadIdx.x == 0) *slot = atomicAdd(counter, 1);
    __syncthreads();
    return *slot;
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *p, int n, int32_t v)
{
    const int i = int(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// One job as a persistent grid: workgroups take bricks from a counter
// (cleared by k_fill_i32 before the launch) until none are left.  A
// workgroup that starts late, its slot held by a kernel running
// concurrently, takes fewer bricks instead of stretching the launch by a
// whole brick's time; which workgroup computes a brick changes no bits.
// Item b at the kernel's shape, except items whose slab has at most PZ / 2
// planes left (the top z-brick of a 2D column, region 1 of linear items):
// they run with PZ / 2 points per lane on the same z origin (same bits: a
// pair's arithmetic does not depend on the shape).
template <int PX, int PZ, bool kComp, bool kClamp>
__device__ __forceinline__ void elecItem(const ElecArgs &a, int b, const float *eps)
{
    constexpr int PH = PZ > 4 ? PZ / 2 : PZ;
    if (a.bxShift == 0) {
        if (b < a.itemsA) {
            elecBrick<PX, PZ, kComp, kClamp>(a, b, eps, PZ, 0);
        } else if (PZ > 4 && a.g.nz - a.slabs * PZ <= PZ / 2) {
            elecBrick<PX, PH, kComp, kClamp>(a, b, eps, PZ, 1);
        } else {
            elecBrick<PX, PZ, kComp, kClamp>(a, b, eps, PZ, 1);
        }
        return;
    }
    if (PZ > 4 && a.g.nz - (b / (a.nbx * a.nby)) * PZ <= PZ / 2) {
        elecBrick<PX, PH, kComp, kClamp>(a, b, eps, PZ);
        return;
    }
    elecBrick<PX, PZ, kComp, kClamp>(a, b, eps);
}

// Register budget of the e-kernels: 4 waves per SIMD (128 VGPRs), the
// workgroups per CU that a table of up to 10,240 entries leaves in LDS.
constexpr int kElecWaves = 4;

template <int PX, int PZ, bool kComp, bool kClamp>
__global__ __launch_bounds__(kElecThreads) __attribute__((amdgpu_waves_per_eu(kElecWaves))) void k_elec(ElecArgs a,
                                                                                                         int *counter)
{
    extern __shared__ float eps[];
    __shared__ int slot;
    loadEps(eps, a.eps32, a.tableLen, kComp ? kNearBins : 0);
    const int total = a.items;
    for (;;) {
        const int item = nextItem(counter, &slot);  // its barrier also orders the table load
        if (item >= total) break;
        elecItem<PX, PZ, kComp, kClamp>(a, item, eps);
    }
}

// sharedTableLen > 0: every job uses the same table (entries past 8191
// repeat entry 8191), loaded once from jobs[0] with that length; else each
// job's table is loaded when the workgroup moves to it.
template <int PX, int PZ, bool kComp, bool kClamp>
__global__ __launch_bounds__(kElecThreads) __attribute__((amdgpu_waves_per_eu(kElecWaves))) void k_elec_batch(
    const ElecArgs *jobs, const int32_t *base, int nJobs, int *counter, int sharedTableLen)
{
    extern __shared__ float eps[];
    __shared__ int slot;
    const int total = base[nJobs];
    int loaded = -1;
    if (sharedTableLen > 0) loadEps(eps, loadConst(jobs).eps32, sharedTableLen, kComp ? kNearBins : 0);
    for (;;) {
        const int item = nextItem(counter, &slot);
        if (item >= total) break;
        const int job = uniform(jobOfItem(base, nJobs, item));
        const ElecArgs a = loadConst(jobs + job);
        if (sharedTableLen <= 0 && job != loaded) {
            loadEps(eps, a.eps32, a.tableLen, kComp ? kNearBins : 0);
            loaded = job;
        }
        __syncthreads();  // table loaded
        elecItem<PX, PZ, kComp, kClamp>(a, item - base[job], eps);
    }
}

// Cutoff maps of every job: one workgroup per (job, brick) item, items in
// queue order (largest jobs first), so the hardware's in-order dispatch
// balances the load.
template <int MM, int KH, bool kAffinity, bool kPointAng, bool kTyped>
__global__ __launch_bounds__(kCutThreads)
__attribute__((amdgpu_waves_per_eu(cutWaves<MM, KH, kAffinity, kTyped>()))) void k_cutoff_batch(
    const CutoffArgs *jobs, const int32_t *base, int nJobs)
{
    __shared__ CutoffLds<MM, kAffinity, kTyped> L;
    const int item = int(blockIdx.x);
    const int job = uniform(jobOfItem(base, nJobs, item));
    cutoffBrick<MM, KH, kAffinity, kPointAng, kTyped>(loadConst(jobs + job), item - base[job], L);
}

// Candidate lists (kFill) and their type-sorted copies of every job: one
// workgroup per (job, brick) item, the items of the cutoff kernels.
template <bool kFill>
__global__ __launch_bounds__(kListThreads) void k_brick_lists_batch(const ListArgs *jobs, const int32_t *base, int nJobs)
{
    __shared__ ListLds S;
    const int item = int(blockIdx.x);
    const int job = uniform(jobOfItem(base, nJobs, item));
    brickList<kFill>(loadConst(jobs + job), item - base[job], S);
}

__global__ __launch_bounds__(kListThreads) void k_type_sort_batch(const ListArgs *jobs, const int32_t *base, int nJobs)
{
    __shared__ int32_t cnt[kMaxTypes], runStart[kMaxTypes], wcnt[4 * kMaxTypes];
    const int item = int(blockIdx.x);
    const int job = uniform(jobOfItem(base, nJobs, item));
    const ListArgs a = loadConst(jobs + job);
    if (a.sorted) typeSortBrick(a, item - base[job], cnt, runStart, wcnt);
}

// Floating grids of every job that has one: one workgroup per (job, brick) item.
__global__ __launch_bounds__(kCutThreads) void k_fmap_batch(const CutoffArgs *jobs, const int32_t *base, int nJobs)
{
    const int item = int(blockIdx.x);
    const int job = uniform(jobOfItem(base, nJobs, item));
    const CutoffArgs a = loadConst(jobs + job);
    if (a.fmap) fmapBrick(a, item - base[job]);
}


// k_add_far for every job: items are (job, 256-point chunk) pairs.
__global__ __launch_bounds__(256) void k_add_far_batch(const ElecArgs *jobs, const int32_t *base, int nJobs)
{
    const int item = int(blockIdx.x);
    const int job = uniform(jobOfItem(base, nJobs, item));
    const FG_CONST ElecArgs *a = constView(jobs + job);
    const int64_t p = int64_t(item - base[job]) * 256 + threadIdx.x;
    if (p < a->P) a->e[p] = a->e[p] + a->far[p];
}

// ------------------------------------------------------------------------
// Electrostatic map in the reference's FP64 arithmetic, all atoms in index
// order (FG/electrostatics/Cpu.cpp:27-88).  One point per lane.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_elec_fp64(GridGeom g, int N, const CutAtom *atoms, const double *epsG,
                                                   int ddd, double *e)
{
    __shared__ double eps[kTableLen];
    for (int i = threadIdx.x; i < kTableLen; i += 256) eps[i] = epsG[i];
    __syncthreads();
    const int ix = blockIdx.x * 16 + (threadIdx.x & 15);
    const int iy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int iz = blockIdx.z;
    const double gx = (ix - g.dx2) * g.spacing, gy = (iy - g.dy2) * g.spacing, gz = (iz - g.dz2) * g.spacing;
    double en = 0;
    for (int ia = 0; ia < N; ++ia) {
        const CutAtom &A = atoms[ia];
        const double tz = A.z - gz, ty = A.y - gy, tx = A.x - gx;
        const double dzSq = tz * tz;
        const double dydz = ty * ty + dzSq;
        const double r = sqrt(tx * tx + dydz);
        const double inv = 1 / r;
        if (ddd) en += A.w * dmin_(inv, 2) * eps[binOf(r)];
        else en += A.w * dmin_(inv, 2);
    }
    if (ix < g.nx && iy < g.ny) e[(int64_t(iz) * g.ny + iy) * g.nx + ix] = en;
}

// Number of (point, atom) pairs with r^2 <= 64, per brick candidate list.
__global__ __launch_bounds__(kCutThreads) void k_count_pairs(CutoffArgs a, unsigned long long *out)
{
    const int b = blockIdx.x;
    const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int ix = bx * kCutBX + (w & 1) * 4 + (l & 3);
    const int iy = by * kCutBY + (w >> 1) * 4 + ((l >> 2) & 3);
    const int iz = bz * kCutBZ + (l >> 4);
    const bool valid = ix < a.g.nx && iy < a.g.ny && iz < a.g.nz;
    const double gx = (ix - a.g.dx2) * a.g.spacing, gy = (iy - a.g.dy2) * a.g.spacing,
                 gz = (iz - a.g.dz2) * a.g.spacing;
    unsigned long long n = 0;
    for (int c = a.brickStart[b]; c < a.brickStart[b + 1]; ++c) {
        const CutAtom &A = a.atoms[uniform(a.brickList[c])];
        double dx = A.x - gx, dy = A.y - gy, dz = A.z - gz;
        n += (valid && dx * dx + dy * dy + dz * dz <= 64.0) ? 1 : 0;
    }
    atomicAdd(out, n);
}

}  // namespace fgk
