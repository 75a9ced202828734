// The plain form: the reference's GPU method in FP32, one job per launch.
// FG = fastgrid_master/fastgrid in the reference.
//
//  * k_plain_elec restates the reference's CUDA electrostatic kernel
//    (FG/electrostatics/cuda_internal/Kernels.inl:174-220 with the
//    global-memory dielectric table, :73-76): every lane owns 8 points along
//    z, the receptor atoms are read as uniform (scalar) loads in chunks of
//    4000 atoms per launch (STD_NUM_ATOMS_PER_KERNEL,
//    FG/electrostatics/cuda_internal/StandardKernels.h:27), and each launch
//    adds its chunk's FP32 sum into an FP32 grid
//    (FG/electrostatics/Cuda.cpp:161-181).  All pairs, min(1/r, 2), no
//    near/far split, no compensation.
//  * k_plain_cutoff restates the reference's per-point cutoff loop
//    (FG/CalculateGridmaps.cpp:199-316) in FP32: one lane per grid point,
//    scanning the candidate list of the point's cell, FP32 tables in global
//    memory, the H-bond terms in FP32.
// Both exist to be measured against; their error against the FP64 forms is
// reported per map.
#pragma once

#include "kernels.hpp"

namespace fgk {

constexpr int kPlainAtomsPerLaunch = 4000;
constexpr int kPlainBX = 16, kPlainBY = 8, kPlainPZ = 8;  // 128 lanes, 8 z-points each
constexpr int kPlainThreads = kPlainBX * kPlainBY;
constexpr int kPlainCutThreads = 256;

// Dielectric of the plain e-kernel, the reference's kernel variants
// (Kernels.inl:73-119): the binned table in global memory (:73-76, also its
// constant-memory variant), the analytic value in place (:100-104,
// calculateDistDepDielInv, FG/autogrid.h:106-112) and the table through a
// normalized, point-sampled, clamped texture (:93-97, :224-228), here a
// global load at the texel that sampling picks; kPlainConst: constant
// dielectric (no factor).
enum { kPlainConst = 0, kPlainTable = 1, kPlainInPlace = 2, kPlainTexture = 3 };

struct PlainElecArgs {
    GridGeom g;
    int32_t a0, a1;        // atom range of this launch
    int32_t pad0, pad1;
    const float4 *atoms;   // (x, y, z, w) in A, FP32
    const float *eps;      // [8192]
    float *acc;            // [P] FP32 sum over launches
};

template <int kDiel>
__device__ __forceinline__ float plainDiel(const float *eps, float r2, float ri)
{
    const float r = fmaxf(r2 * ri, 0.f);  // r = 0: 0 * inf = NaN -> 0
    if (kDiel == kPlainTable) return eps[int(fminf(r * 100.f, float(kTableLen - 1)))];
    if (kDiel == kPlainTexture) {
        // normalized coordinate u = 100 r / 8191; point sampling reads texel
        // floor(u * 8192), clamped to the table
        const float u = (100.f / float(kTableLen - 1)) * r;
        return eps[int(fminf(u * float(kTableLen), float(kTableLen - 1)))];
    }
    if (kDiel == kPlainInPlace) {
        const float E = expf(-0.3153767175f * r);
        return (332.f + float(332 * 7.7839) * E) / (78.4f + -66.57180475f * E);
    }
    return 1.f;
}

template <int kDiel>
__global__ __launch_bounds__(kPlainThreads) void k_plain_elec(PlainElecArgs a)
{
    const int ix = blockIdx.x * kPlainBX + (threadIdx.x % kPlainBX);
    const int iy = blockIdx.y * kPlainBY + (threadIdx.x / kPlainBX);
    const int iz0 = blockIdx.z * kPlainPZ;
    const float s = float(a.g.spacing);
    const float gx = float(ix - a.g.dx2) * s, gy = float(iy - a.g.dy2) * s, gz = float(iz0 - a.g.dz2) * s;
    float en[kPlainPZ];
#pragma unroll
    for (int k = 0; k < kPlainPZ; ++k) en[k] = 0.f;
    for (int ia = a.a0; ia < a.a1; ++ia) {
        const float4 at = a.atoms[ia];
        const float dx = gx - at.x, dy = gy - at.y;
        const float dxy = __builtin_fmaf(dx, dx, dy * dy);
        float dz = gz - at.z;
#pragma unroll
        for (int k = 0; k < kPlainPZ; ++k) {
            const float r2 = __builtin_fmaf(dz, dz, dxy);
            const float ri = __builtin_amdgcn_rsqf(r2);
            float d = fminf(ri, 2.f);
            if (kDiel != kPlainConst) d *= plainDiel<kDiel>(a.eps, r2, ri);
            en[k] = __builtin_fmaf(at.w, d, en[k]);
            dz += s;
        }
    }
    if (ix >= a.g.nx || iy >= a.g.ny) return;
#pragma unroll
    for (int k = 0; k < kPlainPZ; ++k) {
        const int iz = iz0 + k;
        if (iz < a.g.nz) {
            const int64_t p = (int64_t(iz) * a.g.ny + iy) * a.g.nx + ix;
            a.acc[p] += en[k];
        }
    }
}

// FP32 sum grid -> FP64 electrostatic map, and clears the sum for the next run.
__global__ __launch_bounds__(256) void k_plain_elec_finish(int64_t P, float *acc, double *e)
{
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= P) return;
    e[p] = double(acc[p]);
    acc[p] = 0.f;
}

// One receptor atom of the plain cutoff kernel (64 bytes, FP32).
struct PlainAtom {
    float x, y, z, vol;
    float solparQ;        // solpar + solparQ * |q|  (FG/CalculateGridmaps.cpp:235)
    float rv[3], rv2[3];
    int32_t type, hbond, rexp, disorder, pad;
};
static_assert(sizeof(PlainAtom) == 64, "PlainAtom must stay 64 bytes");

struct PlainMap {
    int32_t hbond, isHBonder, isCovalent, pad;
    float solparProbe, volProbe;
};

struct PlainCutArgs {
    GridGeom g;
    int32_t nbx, nby, nbz, M, hydrogen, pad;
    int64_t P;
    float solparQ;
    const PlainAtom *atoms;
    const int32_t *brickStart, *brickList;  // the candidate lists of k_cutoff serve as cells
    const int32_t *hStart, *hList;
    const float *pair;     // [T][kPairBins][M], output map order
    const float *desolv;   // [8192]
    double *maps;          // [M+2][P]
    const double *covalent;
    int32_t covOn, pad2;   // covalent maps from their parameters when covalent is null
    double covPoint[3], covHwSqInv, covBarrier;
    PlainMap mp[16];
};

template <int MM>
__global__ __launch_bounds__(kPlainCutThreads) void k_plain_cutoff(PlainCutArgs a)
{
    const int64_t p = int64_t(blockIdx.x) * kPlainCutThreads + threadIdx.x;
    if (p >= a.P) return;
    const int ix = int(p % a.g.nx), iy = int((p / a.g.nx) % a.g.ny), iz = int(p / (int64_t(a.g.nx) * a.g.ny));
    const int b = ((iz / kCutBZ) * a.nby + iy / kCutBY) * a.nbx + ix / kCutBX;
    const float s = float(a.g.spacing);
    const float gx = float(ix - a.g.dx2) * s, gy = float(iy - a.g.dy2) * s, gz = float(iz - a.g.dz2) * s;
    const int M = a.M;

    // closest DS/D1 atom (FG/CalculateGridmaps.cpp:65-82)
    int cH = 0;
    float rmin = 999999.0f * 999999.0f;
    for (int c = a.hStart[b]; c < a.hStart[b + 1]; ++c) {
        const int ia = a.hList[c];
        const float dx = a.atoms[ia].x - gx, dy = a.atoms[ia].y - gy, dz = a.atoms[ia].z - gz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < rmin || (r2 == rmin && ia < cH)) {
            rmin = r2;
            cH = ia;
        }
    }
    const float cHrv[3] = {a.atoms[cH].rv[0], a.atoms[cH].rv[1], a.atoms[cH].rv[2]};

    float e[MM], hmin[MM], hmax[MM];
    bool flag[MM];
#pragma unroll
    for (int m = 0; m < MM; ++m) {
        e[m] = 0.f;
        hmin[m] = 999999.f;
        hmax[m] = -999999.f;
        flag[m] = false;
    }
    float dmap = 0.f;
    for (int c = a.brickStart[b]; c < a.brickStart[b + 1]; ++c) {
        const int ia = a.brickList[c];
        const PlainAtom &A = a.atoms[ia];
        const float dx = A.x - gx, dy = A.y - gy, dz = A.z - gz;
        float rSq = dx * dx + dy * dy + dz * dz;
        if (rSq > 64.f) continue;
        if (A.type == a.hydrogen && A.disorder) continue;
        if (rSq < 1.0e-24f) rSq = 1.0e-24f;
        const float invR = 1.f / sqrtf(rSq);
        const int idx = min(int((1.f / invR) * 100.f), kTableLen - 1);
        float racc = 1.f, rdon = 1.f, Hramp = 1.f;
        hbondAngular(A.hbond, A.rexp, A.disorder, ia == cH, A.rv, A.rv2, cHrv, dx * invR, dy * invR, dz * invR, racc,
                     rdon, Hramp);
        const float S = a.desolv[idx];
        const bool aDon = A.hbond == HB_DS || A.hbond == HB_D1;
#pragma unroll
        for (int m = 0; m < MM; ++m) {
            if (m >= M || a.mp[m].isCovalent) continue;
            const float E = a.pair[(int64_t(A.type) * kPairBins + min(idx, kPairBins - 1)) * M + m];
            const int mh = a.mp[m].hbond;
            if (a.mp[m].isHBonder) {
                const float rsph = fminf(fmaxf(E / 100.f, 0.f), 1.f);
                if ((mh == HB_AS || mh == HB_A2) && aDon) {
                    const float Eh = (A.disorder && a.hydrogen >= 0)
                                         ? a.pair[(int64_t(a.hydrogen) * kPairBins + max(0, min(idx, kPairBins - 1) - 110)) * M + m]
                                         : E;
                    e[m] += Eh * Hramp * (racc + (1.f - racc) * rsph);
                } else if (mh == HB_A1 && aDon) {
                    const float v = E * (racc + (1.f - racc) * rsph);
                    hmin[m] = fminf(hmin[m], v);
                    hmax[m] = fmaxf(hmax[m], v);
                    flag[m] = true;
                } else if ((mh == HB_DS || mh == HB_D1) && A.hbond >= HB_AS) {
                    const float v = E * (rdon + (1.f - rdon) * rsph);
                    hmin[m] = fminf(hmin[m], v);
                    hmax[m] = fmaxf(hmax[m], v);
                    flag[m] = true;
                } else {
                    e[m] += E;
                }
            } else {
                e[m] += E;
            }
            e[m] += a.mp[m].solparProbe * A.vol * S + A.solparQ * a.mp[m].volProbe * S;
        }
        dmap += a.solparQ * A.vol * S;
    }
#pragma unroll
    for (int m = 0; m < MM; ++m) {
        if (m >= M) continue;
        double v = double(e[m] + (flag[m] ? hmin[m] + hmax[m] : 0.f));
        if (a.mp[m].isCovalent) {
            if (a.covalent) {
                v = a.covalent[int64_t(m) * a.P + p];
            } else {
                // FG/main.cpp:39-46
                const double gxd = (ix - a.g.dx2) * a.g.spacing, gyd = (iy - a.g.dy2) * a.g.spacing,
                             gzd = (iz - a.g.dz2) * a.g.spacing;
                const double dx = a.covPoint[0] - gxd, dy = a.covPoint[1] - gyd, dz = a.covPoint[2] - gzd;
                double r = (dx * dx + dy * dy + dz * dz) * a.covHwSqInv;
                if (r < 1.0e-12) r = 1.0e-12;
                v = a.covOn ? a.covBarrier * (1 - exp(-0.69314718055994529 * r)) : 0.0;
            }
        }
        a.maps[int64_t(m) * a.P + p] = v;
    }
    a.maps[int64_t(M + 1) * a.P + p] = double(dmap);
}

}  // namespace fgk
