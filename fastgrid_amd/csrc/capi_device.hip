// Device half of the libfgmaps C ABI (include/fgmaps.h): contexts, device-
// resident jobs, launches and timing.  The host data layouts the kernels need
// (atom records, spatially sorted electrostatic tiles, coarse cells, guard
// bands) are built here from an fg_job_desc.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "capi_common.hpp"
#include "fgmaps.h"
#include "kernels.hip"
#include "kernels_plain.hip"
#include "kernels_interp.hip"

using fgk::CutAtom;

namespace {


struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

#define FG_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fgc::fail(FG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

}  // namespace

// Staging buffer of fg_trilinterp_noalloc: one fixed array per GPU, held
// for the life of the process, so that the call allocates nothing.  On the
// host this name only stands for the symbol: host code takes the array's
// device address from hipGetSymbolAddress, once per device (scratchSetUp,
// from fg_create with that device current), and uses that address only.
static __device__ __align__(256) unsigned char scratch[FG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (which is `device`).
int scratchSetUp(int device)
{
    if (device < 0 || device >= kMaxDevices) return FG_OK;  // fg_trilinterp_noalloc then fails on it
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratch[device]) return FG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fgc::fail(FG_E_DEVICE, std::string("fg_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return FG_OK;
}

void *scratchOf(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return g_scratch[device];
}

}  // namespace

struct fg_ctx {
    int device = 0, numCUs = 0;
    size_t ldsPerCU = 0;  // LDS bytes per CU (0: unknown)
    hipStream_t stream = nullptr;
    // The electrostatic kernels run on `side`, concurrently with the cutoff
    // kernels on `stream` (both at the default priority): fork is recorded
    // on stream before them, e0/e1 bracket them on side, and stream waits
    // for e1.
    hipStream_t side = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, e0 = nullptr, e1 = nullptr, join = nullptr;
    bool haveTiming = false;
    bool splitTiming = true;  // ev[0]..ev[1]: cutoff phase; e0..e1: electrostatic phase
    bool elecTimed = false;   // e0..e1 were recorded by the last run
    std::string tuneReport;
    // live jobs and batches: fg_destroy detaches them, so destroying them
    // afterwards only frees their memory
    std::vector<fg_job *> jobs;
    std::vector<fg_batch *> batches;
    std::vector<fg_mapset *> mapsets;
    // pinned staging of device-to-host copies (fetches), made on first use
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t pinEv[2] = {nullptr, nullptr};
    // pinned staging of job creation's uploads, grown to the largest creation
    unsigned char *upPin = nullptr;
    size_t upPinBytes = 0;
    // the distinct dielectric tables (first kTableLen FP32 entries) of the
    // jobs made on this context, at most kMaxEpsTables: a job's epsId indexes
    // them (-1: not kept), so a batch compares ids instead of tables
    std::mutex epsMu;
    std::vector<std::vector<float>> epsTables;
};

// After a failure inside a capture that may have forked the side stream:
// join it, so that ending the capture takes both streams out of capture mode.
static void joinSide(fg_ctx *ctx)
{
    (void)hipEventRecord(ctx->join, ctx->side);
    (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
}

// An AutoDock map set in HBM: floats [z][y][x][slot].
struct fg_mapset {
    fg_ctx *ctx = nullptr;
    fg_map_grid g{};
    int nslots = 0;
    float *d = nullptr;
    float *cells = nullptr;  // corner-packed copy (k_pack_cells), 8x the floats
};

struct fg_job {
    fg_ctx *ctx = nullptr;
    fgk::GridGeom g{};
    int64_t P = 0;
    int N = 0, M = 0, T = 0;
    int distDepDiel = 1, hydrogen = -1;
    bool hasGeom = false;
    std::vector<fgk::MapParam> mp;  // kernel order
    std::vector<int> order;         // kernel map k computes output map order[k]
    int KH = 0;                     // maps keeping an H-bond min/max (kernel positions 0..KH-1)
    bool pointAng = false;          // some atom's directional terms depend on the point
    // cutoff inputs
    CutAtom *dAtoms = nullptr;
    double *dVS = nullptr;
    double *dFar = nullptr;  // [P] far-field e of the compensated form
    double *dCharge = nullptr;   // [N] partial charges (when given), for the summary
    double *dSummary = nullptr;  // [M + 2][2] min/max of the rounded maps, then the total charge
    double *dSummaryPart = nullptr;  // [M + 2][kSumChunks][2] the maps' min/max per chunk
    int32_t *dMilli = nullptr;   // [M + 2][P] rounded thousandths (fg_job_fetch_milli), then an overflow flag
    int32_t *dCounter = nullptr;  // work counter of the persistent e-kernel
    int elecPerCU[2][2][2] = {};  // e-kernel workgroups per CU by [px - 1][pz == 8][compensated]
    // the graph of the last compensated or FP32 run and what it was made for
    hipGraph_t runGraph = nullptr;
    hipGraphExec_t runExec = nullptr;
    unsigned runWhat = 0;
    int runKey[5] = {-1, 0, 0, 0, 0};  // precision, px, pz, tile, bxShift
    double *dPair = nullptr, *dRsph = nullptr, *dDesolv = nullptr, *dEps = nullptr, *dCovalent = nullptr;
    // compact table of the maps the job's typed cutoff kernel (compiled for
    // hiMM maps) does not stage in LDS (CutoffArgs::pairHi)
    double *dPairHi = nullptr;
    int hiMM = 4;
    double *dXYZ = nullptr;
    int32_t *dHbond = nullptr, *dCellStart = nullptr, *dCellAtoms = nullptr;
    fgk::CellGrid cg{};
    int nbx = 0, nby = 0, nbz = 0, nb = 0;
    int32_t *dCounts = nullptr, *dStart = nullptr, *dHStart = nullptr, *dList = nullptr, *dHList = nullptr;
    int64_t listTotal = 0;  // candidates over all bricks (the count pass's total)
    // type-sorted candidate lists for the typed cutoff kernel (pair-table rows in LDS)
    bool typed = false;
    int32_t addOnlyClasses = 0;  // bit h: H-bond class h only adds (see CutoffArgs)
    int32_t *dType = nullptr, *dSorted = nullptr, *dTypeStart = nullptr;
    // electrostatic tiles
    float4 *dTiles = nullptr;
    float *dSubBox = nullptr, *dEps32 = nullptr;
    int epsId = -1;  // the context's id of the job's dielectric table (-1: none)
    int nAtoms32 = 0, tableLen = fgk::kTableLen;
    bool clamp = true;
    // k_elec shape: points per lane along x and z, atoms per tile, lanes
    // along x (defaults; the autotuner may change them)
    int px = 1, pz = 4, tile = 64, bxShift = 4;
    bool shaped = false;  // the shape came from fg_autotune or fg_job_set_shape
    int plainDiel = FG_PLAIN_DIEL_TABLE;  // dielectric variant of the plain e-kernel
    float lastCut2 = 0, tableCut2 = 0, spacing100 = 0;
    double elecEd = 0;  // bound on the FP32 coordinates' error per axis (scaled units)
    // e-map band of the cutoff kernel (A^2) and the candidate lists' reach (A)
    double e2Lo = 0, e2Hi = 0, listReach = 8.0 + 1e-4;
    // plain form (FP32 throughout)
    float4 *dPlainQ = nullptr;          // (x, y, z, w) per atom, index order
    fgk::PlainAtom *dPlainAtoms = nullptr;
    float *dPair32 = nullptr, *dDesolv32 = nullptr, *dAcc32 = nullptr;
    fgk::PlainMap pmp[16] = {};
    float solparQ = 0;
    // covalent maps from their parameters, floating grid
    int32_t covOn = 0;
    double covPoint[3] = {0, 0, 0}, covHwSqInv = 1, covBarrier = 1000;
    bool floating = false;
    float *dFmap = nullptr;
    // outputs: [M affinity][e][d]
    double *dMaps = nullptr;
    // the device allocations the job's buffers are carved from; jobs created
    // together share them, and the last job to go frees them
    std::vector<std::shared_ptr<void>> owned;
    // Device buffers are carved out of a few arenas: devAlloc/devUpload
    // record a slot (and stage the upload's bytes); jobsFinishDevice places
    // every recorded slot, copies the staged bytes in and sets the recorded
    // pointers.
    struct Slot {
        void **dst;          // the job's pointer to set
        size_t off, bytes;   // place in the arena
        const void *direct;  // caller memory copied straight in (not staged)
    };
    std::vector<Slot> slots;
    std::vector<unsigned char> stage;  // zero-filled between staged uploads
    size_t arenaBytes = 0;
};

namespace {

// A device buffer of `count` elements in the job's next arena; `ptr` is set
// when the arena is placed.
template <class T>
int devAlloc(fg_job *j, T *&ptr, size_t count)
{
    ptr = nullptr;
    if (count == 0) count = 1;
    const size_t off = (j->arenaBytes + 255) & ~size_t(255);
    j->arenaBytes = off + count * sizeof(T);
    j->slots.push_back(fg_job::Slot{reinterpret_cast<void **>(&ptr), off, count * sizeof(T), nullptr});
    return FG_OK;
}

// A device buffer holding a copy of host[0..count) (staged now, so `host`
// may be a temporary).
template <class T>
int devUpload(fg_job *j, T *&ptr, const T *host, size_t count)
{
    if (int rc = devAlloc(j, ptr, count)) return rc;
    if (count) {
        const fg_job::Slot &sl = j->slots.back();
        if (j->stage.size() < sl.off + sl.bytes) j->stage.resize(sl.off + sl.bytes);
        std::memcpy(j->stage.data() + sl.off, host, count * sizeof(T));
    }
    return FG_OK;
}

// As devUpload for caller memory that stays valid until the job's device
// half has run: copied from `host` then (large arrays are not staged twice).
template <class T>
int devUploadDirect(fg_job *j, T *&ptr, const T *host, size_t count)
{
    if (int rc = devAlloc(j, ptr, count)) return rc;
    if (count) j->slots.back().direct = host;
    return FG_OK;
}

uint32_t spread3(uint32_t v)  // 10 bits -> every third bit
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Launches the candidate-list count pass and the two scans; the list totals
// also go to tot[0] (all candidates) and tot[1] (H-bond candidates).
int runListCount(fg_job *j, const fgk::ListArgs &la, int32_t *tot)
{
    hipLaunchKernelGGL(fgk::k_brick_lists<false>, dim3(j->nb), dim3(fgk::kListThreads), 0, j->ctx->stream, la);
    hipLaunchKernelGGL(fgk::k_scan_counts, dim3(1), dim3(1024), 0, j->ctx->stream, j->dCounts, j->nb, 0, j->dStart,
                       tot);
    hipLaunchKernelGGL(fgk::k_scan_counts, dim3(1), dim3(1024), 0, j->ctx->stream, j->dCounts, j->nb, 1, j->dHStart,
                       tot + 1);
    FG_HIP(hipGetLastError());
    return FG_OK;
}

fgk::ListArgs listArgs(const fg_job *j)
{
    fgk::ListArgs la{};
    la.g = j->g;
    la.cg = j->cg;
    la.nbx = j->nbx;
    la.nby = j->nby;
    la.nbz = j->nbz;
    la.N = j->N;
    la.reach = j->listReach;
    la.xyz = j->dXYZ;
    la.hbond = j->dHbond;
    la.cellStart = j->dCellStart;
    la.cellAtoms = j->dCellAtoms;
    la.counts = j->dCounts;
    la.start = j->dStart;
    la.hStart = j->dHStart;
    la.list = j->dList;
    la.hList = j->dHList;
    la.atomType = j->dType;
    la.sorted = j->dSorted;
    la.typeStart = j->dTypeStart;
    la.T = j->T;
    return la;
}

// Whether the cutoff phase's waves raise their issue priority above the
// persistent e-kernel's beside them (CutoffArgs::prio), from an estimate of
// the two phases' VALU wave-instructions: about 45 per candidate and wave of
// the cutoff kernel (4 waves visit every candidate of their brick) and 7.5
// per 64 pairs of the e-kernel.  The cutoff kernels wait on memory far more
// than the e-kernel, which holds 3 of the 4 slots per CU, so with a large
// share of the work they end long after it: on the 256-pocket batch (cutoff
// ~21 % of the instructions) the raised priority took a replay from 62.4 ms
// to 60.0 ms; on the 50,000-atom job (~12 %) the two phases end together
// without it (22.45 ms per step, 22.7 ms with it).
int cutoffPriority(double candidates, double elecPairs)
{
    const double cut = 180.0 * candidates, elec = 7.5 / 64.0 * elecPairs;
    return cut > 0.15 * (cut + elec) ? 1 : 0;
}

// Maps the single-job cutoff kernel of j is compiled for (launchCutoff).
int cutKernelMM(const fg_job *j)
{
    return j->M <= 4 ? 4 : (j->M <= 6 && j->KH <= 4) ? 6 : (j->M <= 8 && j->KH <= 4) ? 8 : 16;
}

fgk::CutoffArgs cutoffArgs(const fg_job *j, int doElecNear)
{
    fgk::CutoffArgs a{};
    a.g = j->g;
    a.nbx = j->nbx;
    a.nby = j->nby;
    a.nbz = j->nbz;
    a.M = j->M;
    a.hydrogen = j->hydrogen;
    a.doElecNear = doElecNear;
    a.P = j->P;
    a.atoms = j->dAtoms;
    a.vs = j->dVS;
    a.brickStart = j->dStart;
    a.brickList = j->typed ? j->dSorted : j->dList;
    a.hStart = j->dHStart;
    a.hList = j->dHList;
    a.typeStart = j->dTypeStart;
    a.T = j->T;
    a.addOnlyClasses = j->addOnlyClasses;
    a.covMask = 0;
    for (int m = 0; m < j->M && m < 16; ++m) a.covMask |= j->mp[size_t(m)].isCovalent ? 1 << m : 0;
    a.pair = j->dPair;
    a.pairHi = j->dPairHi;
    a.rsph = j->dRsph;
    a.desolv = j->dDesolv;
    a.eps = j->dEps;
    a.maps = j->dMaps;
    a.covalent = j->dCovalent;
    a.covOn = j->covOn;
    for (int k = 0; k < 3; ++k) a.covPoint[k] = j->covPoint[k];
    a.covHwSqInv = j->covHwSqInv;
    a.covBarrier = j->covBarrier;
    a.N = j->N;
    a.fmap = j->floating ? j->dFmap : nullptr;
    for (int m = 0; m < j->M && m < 16; ++m) a.mp[m] = j->mp[size_t(m)];
    a.e2Lo = j->e2Lo;
    a.e2Hi = j->e2Hi;
    a.spacing100 = j->spacing100;
    a.prio = cutoffPriority(double(j->listTotal), double(j->N) * double(j->P));
    return a;
}

template <int MM, bool PA, bool TY>
void launchCutoffKH(fg_job *j, const fgk::CutoffArgs &a, dim3 grid, dim3 block)
{
    hipStream_t s = j->ctx->stream;
    if (j->KH == 0) hipLaunchKernelGGL((fgk::k_cutoff<MM, 0, true, PA, TY>), grid, block, 0, s, a);
    else if (j->KH <= 2) hipLaunchKernelGGL((fgk::k_cutoff<MM, 2, true, PA, TY>), grid, block, 0, s, a);
    else if (MM < 16 || j->KH <= 4) hipLaunchKernelGGL((fgk::k_cutoff<MM, 4, true, PA, TY>), grid, block, 0, s, a);
    // more than 4 maps with a min/max (repeated donor or A1 ligand types):
    // a slot for every map (register spills; no GPF of AutoDock's own types
    // needs it)
    else if constexpr (MM == 16) hipLaunchKernelGGL((fgk::k_cutoff<16, 16, true, PA, TY>), grid, block, 0, s, a);
}

template <int MM, bool TY>
void launchCutoffM(fg_job *j, const fgk::CutoffArgs &a, dim3 grid, dim3 block)
{
    if (j->pointAng) launchCutoffKH<MM, true, TY>(j, a, grid, block);
    else launchCutoffKH<MM, false, TY>(j, a, grid, block);
}

// Candidate lists of the job: the fill pass at the offsets of the count pass
// (made when the job is created, to size the lists), and the type-sorted copy
// the typed cutoff kernel reads; the unsorted lists serve the floating grid
// and the untyped kernels.  Every run makes them again before its cutoff
// phase, as the reference builds its cutoff grid in every calculateGridmaps
// (FG/CalculateGridmaps.cpp:332-415); job creation makes them once too, for
// the queries that read them outside a run (fg_job_cutoff_pairs).
int fillLists(fg_job *j, bool sorted = true)
{
    fgk::ListArgs la = listArgs(j);
    hipLaunchKernelGGL(fgk::k_brick_lists<true>, dim3(j->nb), dim3(fgk::kListThreads), 0, j->ctx->stream, la);
    if (sorted && j->typed)
        hipLaunchKernelGGL(fgk::k_type_sort, dim3(j->nb), dim3(fgk::kListThreads), 0, j->ctx->stream, la);
    FG_HIP(hipGetLastError());
    return FG_OK;
}

// Floating grid from the (unsorted) candidate lists of this run.
int launchFmap(fg_job *j)
{
    fgk::CutoffArgs a = cutoffArgs(j, 0);
    a.brickList = j->dList;
    hipLaunchKernelGGL(fgk::k_fmap, dim3(unsigned(j->nb)), dim3(fgk::kCutThreads), 0, j->ctx->stream, a);
    FG_HIP(hipGetLastError());
    return FG_OK;
}

int launchCutoff(fg_job *j, bool affinity, int doElecNear)
{
    fgk::CutoffArgs a = cutoffArgs(j, doElecNear);
    dim3 grid(unsigned(j->nb)), block(fgk::kCutThreads);
    if (!affinity) {
        a.brickList = j->dList;
        hipLaunchKernelGGL((fgk::k_cutoff<1, 0, false, false, false>), grid, block, 0, j->ctx->stream, a);
    } else if (j->M <= 4) {
        if (j->typed) launchCutoffM<4, true>(j, a, grid, block);
        else launchCutoffM<4, false>(j, a, grid, block);
    } else if (j->M <= 6 && j->KH <= 4) {
        if (j->typed) launchCutoffM<6, true>(j, a, grid, block);
        else launchCutoffM<6, false>(j, a, grid, block);
    } else if (j->M <= 8 && j->KH <= 4) {
        if (j->typed) launchCutoffM<8, true>(j, a, grid, block);
        else launchCutoffM<8, false>(j, a, grid, block);
    } else {
        if (j->typed) launchCutoffM<16, true>(j, a, grid, block);
        else launchCutoffM<16, false>(j, a, grid, block);
    }
    FG_HIP(hipGetLastError());
    return FG_OK;
}

fgk::ElecArgs elecArgs(const fg_job *j, int px, int pz, int tile, int bxShift)
{
    fgk::ElecArgs a{};
    a.g = j->g;
    a.nAtoms = j->nAtoms32;
    a.tableLen = j->tableLen;
    a.tile = tile;
    a.bxShift = bxShift;
    if (bxShift == 0) {
        // linear items: units of px points 8 apart in a row times the pz
        // planes of a slab, in 8 px x 8-column tiles, 256 units to an item;
        // region 0 the full slabs, region 1 the rest (job creation bounds the
        // grid so that these counts fit in 32 bits)
        const int64_t T = fgk::kElecThreads;
        a.tilesX = (j->g.nx + 8 * px - 1) / (8 * px);
        const int lastWidth = j->g.nx - 8 * px * (a.tilesX - 1);
        a.lastLanes = px == 2 ? (lastWidth + 1) / 2 : lastWidth;
        a.ux = 8 * (a.tilesX - 1) + a.lastLanes;
        a.units = a.ux * j->g.ny;
        a.slabs = j->g.nz / pz;
        a.itemsA = int32_t((int64_t(a.slabs) * a.units + T - 1) / T);
        a.items = a.itemsA + (j->g.nz > a.slabs * pz ? int32_t((a.units + T - 1) / T) : 0);
        a.nbx = a.nby = a.nbz = 0;
    } else {
        const int bw = (1 << bxShift) * px, byl = fgk::kElecThreads >> bxShift;
        a.nbx = (j->g.nx + bw - 1) / bw;
        a.nby = (j->g.ny + byl - 1) / byl;
        a.nbz = (j->g.nz + pz - 1) / pz;
        a.items = a.nbx * a.nby * a.nbz;
    }
    a.clampMax = j->tableLen - 1;
    a.floorBitsNeg4 = int32_t(0u - 4u * fgk::kFloorBits);
    a.P = j->P;
    a.tiles = j->dTiles;
    a.subBox = j->dSubBox;
    a.eps32 = j->dEps32;
    a.e = j->dMaps + int64_t(j->M) * j->P;
    a.far = j->dFar;
    a.lastCut2 = j->lastCut2;
    a.tableCut2 = j->tableCut2;
    a.spacing100 = j->spacing100;
    return a;
}


// The persistent e-kernel: the workgroups its occupancy admits per CU (found
// when the job is made, so that nothing here queries the device), less
// `reserve` per CU left to the kernels running concurrently (a persistent
// workgroup keeps its slot until the bricks run out), after its counter is
// cleared on the same stream.
template <int PX, int PZ, bool kComp, bool kClamp>
int launchElecK(fg_job *j, const fgk::ElecArgs &a, size_t lds, int reserve, hipStream_t s)
{
    auto k = fgk::k_elec<PX, PZ, kComp, kClamp>;
    const int perCU = j->elecPerCU[PX - 1][PZ == 8][kComp];
    const int wgs = std::max(1, std::min(a.items, std::max(perCU - reserve, 1) * j->ctx->numCUs));
    hipLaunchKernelGGL(fgk::k_fill_i32, dim3(1), dim3(256), 0, s, j->dCounter, 1, 0);
    hipLaunchKernelGGL(k, dim3(unsigned(wgs)), dim3(fgk::kElecThreads), lds, s, a, j->dCounter);
    return FG_OK;
}

template <int PX, int PZ>
int launchElecV(fg_job *j, const fgk::ElecArgs &a, bool comp, size_t lds, int reserve, hipStream_t s)
{
    if (comp)
        return j->clamp ? launchElecK<PX, PZ, true, true>(j, a, lds, reserve, s)
                        : launchElecK<PX, PZ, true, false>(j, a, lds, reserve, s);
    return j->clamp ? launchElecK<PX, PZ, false, true>(j, a, lds, reserve, s)
                    : launchElecK<PX, PZ, false, false>(j, a, lds, reserve, s);
}

// k_elec of shape (px, pz, tile, bxShift) on stream s.
int launchElecShape(fg_job *j, int px, int pz, int tile, int bxShift, bool comp, int reserve, hipStream_t s)
{
    const fgk::ElecArgs a = elecArgs(j, px, pz, tile, bxShift);
    const size_t lds = size_t(j->tableLen) * sizeof(float);
    if (px == 2)
        return pz == 8 ? launchElecV<2, 8>(j, a, comp, lds, reserve, s) : launchElecV<2, 4>(j, a, comp, lds, reserve, s);
    return pz == 8 ? launchElecV<1, 8>(j, a, comp, lds, reserve, s) : launchElecV<1, 4>(j, a, comp, lds, reserve, s);
}

// The electrostatic kernel of the precision on stream s (the compensated
// form writes its far part into dFar, for k_add_far); reserve: workgroup
// slots per CU left to the cutoff phase running concurrently.
int launchElec(fg_job *j, int precision, int reserve, hipStream_t s)
{
    double *e = j->dMaps + int64_t(j->M) * j->P;
    if (precision == FG_PREC_FP64) {
        dim3 grid(unsigned((j->g.nx + 15) / 16), unsigned((j->g.ny + 15) / 16), unsigned(j->g.nz));
        hipLaunchKernelGGL(fgk::k_elec_fp64, grid, dim3(256), 0, s, j->g, j->N, j->dAtoms, j->dEps, j->distDepDiel, e);
        FG_HIP(hipGetLastError());
        return FG_OK;
    }
    if (int rc = launchElecShape(j, j->px, j->pz, j->tile, j->bxShift, precision == FG_PREC_COMPENSATED, reserve, s))
        return rc;
    FG_HIP(hipGetLastError());
    return FG_OK;
}

// Electrostatic far-field tiles: atoms in Morton order of 2 A cells, padded to
// a multiple of kSubTile, coordinates and w scaled by 100 (FP32), plus FP64
// coordinates for the exact near/far predicate and outward-rounded boxes of
// every kSubTile atoms.  Also the guard band of the near/far predicate (see
// ElecArgs).
constexpr int kTableMaxNoClamp = 12288;

constexpr size_t kMaxEpsTables = 64;

// The id of dielectric table t[0..kTableLen) among ctx's tables (bit-equal
// tables share an id); -1 once kMaxEpsTables distinct ones are kept.
int internEps(fg_ctx *ctx, const float *t)
{
    std::lock_guard<std::mutex> lock(ctx->epsMu);
    const size_t bytes = size_t(fgk::kTableLen) * sizeof(float);
    for (size_t i = 0; i < ctx->epsTables.size(); ++i)
        if (std::memcmp(ctx->epsTables[i].data(), t, bytes) == 0) return int(i);
    if (ctx->epsTables.size() >= kMaxEpsTables) return -1;
    ctx->epsTables.emplace_back(t, t + fgk::kTableLen);
    return int(ctx->epsTables.size() - 1);
}

// Squared reach (scaled units) within which a sub-tile's farthest corner
// never indexes past a table of tableLen entries (0: no clamp, no limit),
// for FP32 coordinates within Ed of the FP64 ones.
float tableCut2Of(int tableLen, double Ed)
{
    const double tableCut = tableLen > 0 ? (tableLen - 3.0) - 4 * Ed : 1e18;
    return float(tableCut * tableCut);
}

int buildElecTiles(fg_ctx *ctx, fg_job *j, const fg_job_desc &d, double gridMaxAbs)
{
    const int N = j->N, TT = fgk::kSubTile;
    const fg_atoms &A = d.atoms;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N; ++i) {
        const double c[3] = {A.x[i], A.y[i], A.z[i]};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], c[k]);
            hi[k] = std::max(hi[k], c[k]);
        }
    }
    std::vector<std::pair<uint32_t, int>> order(static_cast<size_t>(N));
    for (int i = 0; i < N; ++i) {
        uint32_t c[3];
        const double v[3] = {A.x[i], A.y[i], A.z[i]};
        for (int k = 0; k < 3; ++k) c[k] = uint32_t(std::min(1023.0, std::floor((v[k] - lo[k]) / 2.0)));
        order[size_t(i)] = {spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2), i};
    }
    std::stable_sort(order.begin(), order.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    const int nSub = (N + TT - 1) / TT;
    j->nAtoms32 = nSub * TT;
    const size_t NT = size_t(j->nAtoms32);
    std::vector<float4> tiles(std::max<size_t>(NT, 1));
    std::vector<float> box(std::max<size_t>(size_t(nSub) * 6, 1));
    const double s = j->g.spacing;
    for (size_t q = 0; q < NT; ++q) {
        double c[3], w;
        if (q < size_t(N)) {
            int i = order[q].second;
            c[0] = A.x[i];
            c[1] = A.y[i];
            c[2] = A.z[i];
            w = A.w ? A.w[i] : 0.0;  // no weights: a floating-grid-only job
        } else {
            // padding: zero charge at the centre of a grid cell next to the
            // last atom, never on a grid point, inside the table's range
            int i = order[size_t(N) - 1].second;
            const double v[3] = {A.x[i], A.y[i], A.z[i]};
            for (int k = 0; k < 3; ++k) c[k] = (std::floor(v[k] / s) + 0.5) * s;
            w = 0;
        }
        tiles[q] = make_float4(float(c[0] * 100.0), float(c[1] * 100.0), float(c[2] * 100.0), float(w * 100.0));
    }
    for (int t = 0; t < nSub; ++t) {
        float bl[3] = {1e30f, 1e30f, 1e30f}, bh[3] = {-1e30f, -1e30f, -1e30f};
        for (int q = t * TT; q < (t + 1) * TT; ++q) {
            const float v[3] = {tiles[size_t(q)].x, tiles[size_t(q)].y, tiles[size_t(q)].z};
            for (int k = 0; k < 3; ++k) {
                bl[k] = std::min(bl[k], v[k]);
                bh[k] = std::max(bh[k], v[k]);
            }
        }
        for (int k = 0; k < 3; ++k) {
            box[size_t(6 * t + k)] = std::nextafter(bl[k], -1e30f);
            box[size_t(6 * t + 3 + k)] = std::nextafter(bh[k], 1e30f);
        }
    }
    // Bound on |r2' - 1e4 * r2_fp64| for r near 8 A, in scaled units.
    const double u = std::ldexp(1.0, -24);
    double C = gridMaxAbs * 100.0;
    for (int k = 0; k < 3; ++k) C = std::max(C, std::max(std::fabs(lo[k]), std::fabs(hi[k])) * 100.0);
    const double dmax = 900.0;  // |d'| near the boundary (<= 9 A)
    const double Ed = u * (2 * C + dmax) + 1e-9 * C;
    const double B2 = 3 * u * 6.5e5 + 2 * dmax * std::sqrt(3.0) * Ed + 3 * Ed * Ed;
    // The e-map's near pairs are those whose FP32 bin is below 800, i.e.
    // whose r2' * rsq(r2') < 800, within 2 ulp of sqrt(r2'): below
    // 6.4e5 - 4 B2 - 1 for certain (scaled r^2), above 6.4e5 + 4 B2 + 1 not.
    // The cutoff kernel takes the first for certain and decides the band by
    // the far field's own FP32 arithmetic; its lists reach past the band.
    j->e2Lo = (6.4e5 - 4 * B2 - 1.0) * 1e-4;
    j->e2Hi = (6.4e5 + 4 * B2 + 1.0) * 1e-4;
    if (j->e2Hi >= 64.08)  // bins up to 800 only in the cutoff kernel's tables
        return fgc::fail(FG_E_ARG, "fg_job_create: coordinates too far from the origin (FP32 near-field band)");
    j->listReach = std::sqrt(j->e2Hi) + 1e-4;
    // The FP32 coordinates and outward-rounded sub-tile boxes differ from
    // the FP64 ones by at most Ed per axis (< 1e-5 A for atoms within 100 A
    // of the origin): the margin of the table-range tests below.
    j->elecEd = Ed;
    j->spacing100 = float(s * 100.0);
    // dielectric table for k_elec: extended past 8191 (no clamp) when it fits
    double span2 = 0;
    for (int k = 0; k < 3; ++k) {
        double gl = -gridMaxAbs, gh = gridMaxAbs;
        double ext = std::max(std::fabs(hi[k] - gl), std::fabs(gh - lo[k]));
        span2 += ext * ext;
    }
    // padding atoms sit up to half a cell outside the atom box: add 2 cells
    const double tmax = (std::sqrt(span2) + 2 * s) * 100.0 * (1 + 1e-5) + 4;
    std::vector<float> eps32;
    const double *eps = d.tables.eps;
    // Entries past 8191 repeat entry 8191, so clamping the index to the end of
    // the (possibly extended) table is the reference's min(k, 8191).
    // With the table covering every reachable index no clamp is needed: the
    // FP32 index exceeds 100 r by at most a relative 1e-6 plus ~1e-3 from
    // coordinate rounding, inside the slack of tmax.  Past kTableMaxNoClamp
    // entries (4 bytes each in LDS) the table would cost occupancy, so it
    // stays at 8192 entries and the clamped variant runs: tiles past the last
    // bin skip the table and only tiles that may reach past it clamp.
    if (tmax + 2 <= kTableMaxNoClamp) {
        j->tableLen = std::max(fgk::kTableLen, int(std::ceil(tmax)) + 2);
        j->clamp = false;
    } else {
        j->tableLen = fgk::kTableLen;
        j->clamp = true;
    }
    eps32.resize(size_t(j->tableLen));
    for (int k = 0; k < j->tableLen; ++k) {
        double v = d.dist_dep_diel ? eps[std::min(k, fgk::kTableLen - 1)] : 1.0;
        eps32[size_t(k)] = float(v);
    }
    // The FP32 table is constant from bin `tail` on (332/eps(r) stops changing
    // in FP32 near 58.4 A; a constant dielectric is constant everywhere), at
    // or past the compensated form's first far bin.  Sub-tiles whose boxes lie
    // farther than lastCut from a wave's box (a margin of 2 bins plus 4 Ed)
    // read only that value, as the last table bin: the bin, its address and
    // the read drop out and every pair keeps its bits.  Sub-tiles whose
    // farthest corner lies within tableCut never index past the table (the
    // clamped variant skips its clamp for them).
    int tail = fgk::kTableLen - 1;
    while (tail > fgk::kNearBins && eps32[size_t(tail - 1)] == eps32[size_t(fgk::kTableLen - 1)]) --tail;
    const double lastCut = (tail + 2.0) + 4 * Ed;
    j->lastCut2 = float(lastCut * lastCut);
    j->tableCut2 = tableCut2Of(j->clamp ? j->tableLen : 0, Ed);
    j->epsId = internEps(ctx, eps32.data());
    if (int rc = devUpload(j, j->dTiles, tiles.data(), tiles.size())) return rc;
    if (int rc = devUpload(j, j->dSubBox, box.data(), box.size())) return rc;
    if (int rc = devUpload(j, j->dEps32, eps32.data(), eps32.size())) return rc;
    return FG_OK;
}

int buildCutoff(fg_job *j, const fg_job_desc &d)
{
    const int N = j->N;
    const fg_atoms &A = d.atoms;
    std::vector<CutAtom> atoms(size_t(std::max(N, 1)));
    std::vector<double> xyz(size_t(3) * size_t(std::max(N, 1)));
    std::vector<int32_t> hb(size_t(std::max(N, 1)), 0);
    for (int i = 0; i < N; ++i) {
        CutAtom c{};
        c.x = A.x[i];
        c.y = A.y[i];
        c.z = A.z[i];
        c.w = A.w ? A.w[i] : 0.0;
        if (j->hasGeom) {
            c.qvol = d.solpar_q * A.vol[i];
            bool zeroVectors = true;
            for (int k = 0; k < 3; ++k) {
                c.rv[k] = d.geom.rvector[3 * i + k];
                c.rv2[k] = d.geom.rvector2[3 * i + k];
                zeroVectors = zeroVectors && c.rv[k] == 0 && c.rv2[k] == 0;
            }
            c.hbond = A.hbond[i];
            c.rexp = d.geom.rexp[i];
            c.acts = 0;
            for (int k = 0; k < j->M && k < 16; ++k)
                c.acts |= int32_t(uint32_t((j->mp[size_t(k)].actions >> (2 * c.hbond)) & 3) << (2 * k));
            const int dis = d.geom.disorder[i] ? 1 : 0;
            int kind = fgk::kAngNone;
            c.ang[0] = c.ang[1] = c.ang[2] = 1.0;
            if (c.hbond == fgk::HB_D1 || c.hbond == fgk::HB_A1 || c.hbond == fgk::HB_A2) {
                if (zeroVectors) {
                    // with zero bond vectors racc/rdon/Hramp do not depend on the
                    // point: evaluate the reference's formulas once, with the
                    // host's libm as the reference does
                    kind = fgk::kAngConst;
                    const double zero[3] = {0, 0, 0};
                    fgk::hbondAngular(c.hbond, c.rexp, dis, true, zero, zero, zero, 1.0, 0.0, 0.0, c.ang[0], c.ang[1],
                                      c.ang[2]);
                } else {
                    kind = fgk::kAngPoint;
                    j->pointAng = true;
                }
            }
            // the kernel reads the atom's uniform fields with one scalar move
            c.flags = dis | (kind << 1) | (c.hbond << 3) | (A.type[i] << 8);
        }
        atoms[size_t(i)] = c;
        xyz[size_t(3 * i)] = c.x;
        xyz[size_t(3 * i + 1)] = c.y;
        xyz[size_t(3 * i + 2)] = c.z;
        hb[size_t(i)] = c.hbond;
    }
    if (int rc = devUpload(j, j->dAtoms, atoms.data(), atoms.size())) return rc;
    // per atom the two factors of the desolvation term (FG/CalculateGridmaps.cpp:234-235):
    // vol and solpar + solparQ |q|; the probe's factors are per map (MapParam)
    std::vector<double> vs(size_t(std::max(N, 1)) * 2, 0.0);
    if (j->hasGeom)
        for (int i = 0; i < N; ++i) {
            vs[size_t(2 * i)] = A.vol[i];
            vs[size_t(2 * i + 1)] = A.solpar[i] + d.solpar_q * std::fabs(A.charge[i]);
        }
    if (int rc = devUpload(j, j->dVS, vs.data(), vs.size())) return rc;
    if (int rc = devUpload(j, j->dXYZ, xyz.data(), xyz.size())) return rc;
    if (int rc = devUpload(j, j->dHbond, hb.data(), hb.size())) return rc;
    // coarse cells: edge >= 4 A and >= the brick edge, so a brick's 8 A
    // neighbourhood spans at most 8 cells per axis
    const double s = j->g.spacing;
    double size = std::max({4.0, fgk::kCutBX * s, fgk::kCutBY * s, fgk::kCutBZ * s});
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (N > 0) {
        lo[0] = hi[0] = A.x[0];
        lo[1] = hi[1] = A.y[0];
        lo[2] = hi[2] = A.z[0];
    }
    for (int i = 0; i < N; ++i) {
        const double c[3] = {A.x[i], A.y[i], A.z[i]};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], c[k]);
            hi[k] = std::max(hi[k], c[k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        j->cg.origin[k] = lo[k] - 1e-3;
        j->cg.nc[k] = int(std::floor((hi[k] - j->cg.origin[k]) / size)) + 1;
    }
    j->cg.size = size;
    const int nc = j->cg.nc[0] * j->cg.nc[1] * j->cg.nc[2];
    std::vector<int32_t> cstart(size_t(nc) + 1, 0), catoms(size_t(std::max(N, 1)));
    std::vector<int32_t> cell(size_t(std::max(N, 1)));
    for (int i = 0; i < N; ++i) {
        const double c[3] = {A.x[i], A.y[i], A.z[i]};
        int ci[3];
        for (int k = 0; k < 3; ++k)
            ci[k] = std::min(j->cg.nc[k] - 1, std::max(0, int(std::floor((c[k] - j->cg.origin[k]) / size))));
        cell[size_t(i)] = (ci[2] * j->cg.nc[1] + ci[1]) * j->cg.nc[0] + ci[0];
        cstart[size_t(cell[size_t(i)]) + 1]++;
    }
    for (int c = 0; c < nc; ++c) cstart[size_t(c) + 1] += cstart[size_t(c)];
    std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
    for (int i = 0; i < N; ++i) catoms[size_t(fill[size_t(cell[size_t(i)])]++)] = i;
    if (int rc = devUpload(j, j->dCellStart, cstart.data(), cstart.size())) return rc;
    if (int rc = devUpload(j, j->dCellAtoms, catoms.data(), catoms.size())) return rc;

    j->nbx = (j->g.nx + fgk::kCutBX - 1) / fgk::kCutBX;
    j->nby = (j->g.ny + fgk::kCutBY - 1) / fgk::kCutBY;
    j->nbz = (j->g.nz + fgk::kCutBZ - 1) / fgk::kCutBZ;
    j->nb = j->nbx * j->nby * j->nbz;
    if (int rc = devAlloc(j, j->dCounts, size_t(2) * size_t(j->nb))) return rc;
    if (int rc = devAlloc(j, j->dStart, size_t(j->nb) + 1)) return rc;
    if (int rc = devAlloc(j, j->dHStart, size_t(j->nb) + 1)) return rc;
    // the typed cutoff kernel: up to kMaxTypes receptor types (its LDS holds
    // one type's rows for up to 4 maps)
    j->typed = j->hasGeom && j->M > 0 && j->T >= 1 && j->T <= fgk::kMaxTypes;
    if (j->typed) {
        std::vector<int32_t> ty(size_t(std::max(N, 1)), 0);
        for (int i = 0; i < N; ++i) ty[size_t(i)] = A.type[i];
        if (int rc = devUpload(j, j->dType, ty.data(), ty.size())) return rc;
    }
    // tables
    // bins 0..800 of every row, maps in kernel order, and per entry the
    // saturated E / 100 of the H-bond terms (FG/CalculateGridmaps.cpp:210),
    // formed with the device's operations
    const int Mx = std::max(j->M, 1);
    const size_t pairN = size_t(std::max(d.tables.n_types, 1)) * fgk::kPairBins * size_t(Mx);
    std::vector<double> perm(pairN, 0.0), rsph(pairN, 0.0);
    if (j->hasGeom && d.tables.pair) {
        for (int t = 0; t < d.tables.n_types; ++t)
            for (int b = 0; b < fgk::kPairBins; ++b)
                for (int k = 0; k < j->M; ++k) {
                    const double v = d.tables.pair[(size_t(t) * fgk::kTableLen + size_t(b)) * size_t(j->M) +
                                                   size_t(j->order[size_t(k)])];
                    double q = v / 100;
                    q = q > 0 ? q : 0;
                    q = q < 1 ? q : 1;
                    const size_t o = (size_t(t) * fgk::kPairBins + size_t(b)) * size_t(j->M) + size_t(k);
                    perm[o] = v;
                    rsph[o] = q;
                }
    }
    if (int rc = devUpload(j, j->dPair, perm.data(), perm.size())) return rc;
    if (int rc = devUpload(j, j->dRsph, rsph.data(), rsph.size())) return rc;
    // the compact table for the job's own typed kernel (its rows cover every
    // map the kernel is compiled for; those the job lacks hold 0)
    const size_t rows = size_t(std::max(d.tables.n_types, 1)) * fgk::kPairBins;
    j->hiMM = cutKernelMM(j);
    const int first = fgk::ldsMapsOf(j->hiMM), stride = fgk::hiStrideOf(j->hiMM);
    std::vector<double> compact(rows * size_t(stride), 0.0);
    for (size_t r = 0; r < rows; ++r)
        for (int k = 0; first + k < j->M && k < stride; ++k)
            compact[r * size_t(stride) + size_t(k)] = perm[r * size_t(Mx) + size_t(first + k)];
    if (int rc = devUpload(j, j->dPairHi, compact.data(), compact.size())) return rc;
    std::vector<double> zt(fgk::kTableLen, 0.0);
    if (int rc = devUpload(j, j->dDesolv, (j->hasGeom && d.tables.desolv) ? d.tables.desolv : zt.data(), zt.size()))
        return rc;
    std::vector<double> eps(fgk::kTableLen, 1.0);
    if (d.dist_dep_diel) std::copy(d.tables.eps, d.tables.eps + fgk::kTableLen, eps.begin());
    if (int rc = devUpload(j, j->dEps, eps.data(), eps.size())) return rc;
    if (d.covalent && j->hasGeom) {
        if (int rc = devUploadDirect(j, j->dCovalent, d.covalent, size_t(j->M) * size_t(j->P))) return rc;
    }
    return FG_OK;
}

// Candidate lists: one count pass sizes them (the inputs are on the device),
// then they and the type-sorted copies get an arena of their own.
// FP32 inputs of the plain form: atoms in index order, tables in output map order.
int buildPlain(fg_job *j, const fg_job_desc &d)
{
    const int N = j->N;
    const fg_atoms &A = d.atoms;
    std::vector<float4> q(size_t(std::max(N, 1)));
    std::vector<fgk::PlainAtom> pa(size_t(std::max(N, 1)));
    for (int i = 0; i < N; ++i) {
        q[size_t(i)] = make_float4(float(A.x[i]), float(A.y[i]), float(A.z[i]), float(A.w ? A.w[i] : 0.0));
        fgk::PlainAtom c{};
        c.x = float(A.x[i]);
        c.y = float(A.y[i]);
        c.z = float(A.z[i]);
        if (j->hasGeom) {
            c.vol = float(A.vol[i]);
            c.solparQ = float(A.solpar[i] + d.solpar_q * std::fabs(A.charge[i]));
            for (int k = 0; k < 3; ++k) {
                c.rv[k] = float(d.geom.rvector[3 * i + k]);
                c.rv2[k] = float(d.geom.rvector2[3 * i + k]);
            }
            c.type = A.type[i];
            c.hbond = A.hbond[i];
            c.rexp = d.geom.rexp[i];
            c.disorder = d.geom.disorder[i] ? 1 : 0;
        }
        pa[size_t(i)] = c;
    }
    if (int rc = devUpload(j, j->dPlainQ, q.data(), q.size())) return rc;
    if (int rc = devUpload(j, j->dPlainAtoms, pa.data(), pa.size())) return rc;
    const size_t pairN = size_t(std::max(d.tables.n_types, 1)) * fgk::kPairBins * size_t(std::max(j->M, 1));
    std::vector<float> pair(pairN, 0.f), desolv(fgk::kTableLen, 0.f);
    if (j->hasGeom) {
        for (int t = 0; t < d.tables.n_types && d.tables.pair; ++t)
            for (size_t k = 0; k < size_t(fgk::kPairBins) * size_t(j->M); ++k)
                pair[size_t(t) * fgk::kPairBins * size_t(j->M) + k] =
                    float(d.tables.pair[size_t(t) * fgk::kTableLen * size_t(j->M) + k]);
        for (int k = 0; k < fgk::kTableLen; ++k) desolv[size_t(k)] = float(d.tables.desolv[k]);
        for (int m = 0; m < j->M; ++m) {
            const fg_map_params &p = d.maps[m];
            j->pmp[m] = fgk::PlainMap{p.hbond, p.is_hbonder, p.is_covalent, 0, float(p.solpar_probe), float(p.vol_probe)};
        }
    }
    j->solparQ = float(d.solpar_q);
    if (int rc = devUpload(j, j->dPair32, pair.data(), pair.size())) return rc;
    if (int rc = devUpload(j, j->dDesolv32, desolv.data(), desolv.size())) return rc;
    // the FP32 sum grid starts at zero (cleared once the arena exists)
    return devAlloc(j, j->dAcc32, size_t(j->P));
}

// The plain form: cutoff maps (per-point lists) then the chunked e-map.
int runPlainAffinity(fg_job *j)
{
    fgk::PlainCutArgs a{};
    a.g = j->g;
    a.nbx = j->nbx;
    a.nby = j->nby;
    a.nbz = j->nbz;
    a.M = j->M;
    a.hydrogen = j->hydrogen;
    a.P = j->P;
    a.solparQ = j->solparQ;
    a.atoms = j->dPlainAtoms;
    a.brickStart = j->dStart;
    a.brickList = j->dList;
    a.hStart = j->dHStart;
    a.hList = j->dHList;
    a.pair = j->dPair32;
    a.desolv = j->dDesolv32;
    a.maps = j->dMaps;
    a.covalent = j->dCovalent;
    a.covOn = j->covOn;
    for (int k = 0; k < 3; ++k) a.covPoint[k] = j->covPoint[k];
    a.covHwSqInv = j->covHwSqInv;
    a.covBarrier = j->covBarrier;
    for (int m = 0; m < j->M && m < 16; ++m) a.mp[m] = j->pmp[m];
    const dim3 grid(unsigned((j->P + fgk::kPlainCutThreads - 1) / fgk::kPlainCutThreads)), block(fgk::kPlainCutThreads);
    hipStream_t s = j->ctx->stream;
    if (j->M <= 4) hipLaunchKernelGGL(fgk::k_plain_cutoff<4>, grid, block, 0, s, a);
    else if (j->M <= 8) hipLaunchKernelGGL(fgk::k_plain_cutoff<8>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(fgk::k_plain_cutoff<16>, grid, block, 0, s, a);
    FG_HIP(hipGetLastError());
    return FG_OK;
}

// The dielectric variant the reference's CUDA path would run for this job:
// the one asked for, or its automatic choice (FG/electrostatics/Cuda.cpp:
// 114-124): in place when no axis of the grid padded to its 16 x 8 x (8 z)
// blocks (:43-55) exceeds 160 points, else the texture.
int plainDielKind(const fg_job *j)
{
    if (!j->distDepDiel) return fgk::kPlainConst;
    switch (j->plainDiel) {
    case FG_PLAIN_DIEL_INPLACE: return fgk::kPlainInPlace;
    case FG_PLAIN_DIEL_TEXTURE: return fgk::kPlainTexture;
    case FG_PLAIN_DIEL_AUTO: {
        auto align = [](int v, int m) { return (v + m - 1) / m * m; };
        const int mx = std::max({align(j->g.nx, fgk::kPlainBX), align(j->g.ny, fgk::kPlainBY), align(j->g.nz, fgk::kPlainPZ)});
        return mx > 160 ? fgk::kPlainTexture : fgk::kPlainInPlace;
    }
    default: return fgk::kPlainTable;
    }
}

int runPlainElec(fg_job *j, hipStream_t s)
{
    fgk::PlainElecArgs a{};
    a.g = j->g;
    a.atoms = j->dPlainQ;
    a.eps = j->dEps32;
    a.acc = j->dAcc32;
    const dim3 grid(unsigned((j->g.nx + fgk::kPlainBX - 1) / fgk::kPlainBX),
                    unsigned((j->g.ny + fgk::kPlainBY - 1) / fgk::kPlainBY),
                    unsigned((j->g.nz + fgk::kPlainPZ - 1) / fgk::kPlainPZ));
    const int kind = plainDielKind(j);
    auto k = kind == fgk::kPlainConst     ? fgk::k_plain_elec<fgk::kPlainConst>
             : kind == fgk::kPlainInPlace ? fgk::k_plain_elec<fgk::kPlainInPlace>
             : kind == fgk::kPlainTexture ? fgk::k_plain_elec<fgk::kPlainTexture>
                                          : fgk::k_plain_elec<fgk::kPlainTable>;
    for (int a0 = 0; a0 < j->N; a0 += fgk::kPlainAtomsPerLaunch) {
        a.a0 = a0;
        a.a1 = std::min(j->N, a0 + fgk::kPlainAtomsPerLaunch);
        hipLaunchKernelGGL(k, grid, dim3(fgk::kPlainThreads), 0, s, a);
    }
    hipLaunchKernelGGL(fgk::k_plain_elec_finish, dim3(unsigned((j->P + 255) / 256)), dim3(256), 0, s, j->P, j->dAcc32,
                       j->dMaps + int64_t(j->M) * j->P);
    FG_HIP(hipGetLastError());
    return FG_OK;
}

template <class T>
void forget(std::vector<T *> &v, T *x)
{
    v.erase(std::remove(v.begin(), v.end(), x), v.end());
}

void destroyJob(fg_job *j)
{
    if (!j) return;
    if (j->ctx) forget(j->ctx->jobs, j);
    if (j->runExec) (void)hipGraphExecDestroy(j->runExec);
    if (j->runGraph) (void)hipGraphDestroy(j->runGraph);
    j->owned.clear();  // frees the allocations no other job shares
    delete j;
}

// ------------------------------------------------------------------------
// Autotuner of the electrostatic kernel shape
// ------------------------------------------------------------------------
// Device time of one electrostatic launch of shape (px, pz, tile, bxShift), ms.
int timeShape(fg_job *j, int px, int pz, int tile, int bxShift, bool comp, float &ms)
{
    fg_ctx *ctx = j->ctx;
    // one slot per CU left free, as in a step with the cutoff phase
    auto go = [&]() { return launchElecShape(j, px, pz, tile, bxShift, comp, 1, ctx->stream); };
    if (int rc = go()) return rc;  // warm-up
    FG_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    if (int rc = go()) return rc;
    if (int rc = go()) return rc;
    FG_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    FG_HIP(hipGetLastError());
    FG_HIP(hipEventSynchronize(ctx->ev[3]));
    FG_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]));
    ms *= 0.5f;
    return FG_OK;
}

template <int PX, int PZ>
const void *elecKernelV(bool comp, bool clamp)
{
    if (comp) return clamp ? reinterpret_cast<const void *>(fgk::k_elec<PX, PZ, true, true>)
                           : reinterpret_cast<const void *>(fgk::k_elec<PX, PZ, true, false>);
    return clamp ? reinterpret_cast<const void *>(fgk::k_elec<PX, PZ, false, true>)
                 : reinterpret_cast<const void *>(fgk::k_elec<PX, PZ, false, false>);
}

// Workgroups per CU of every e-kernel variant at the job's table size.
int elecOccupancy(fg_job *j)
{
    const size_t lds = size_t(j->tableLen) * sizeof(float);
    for (int px = 1; px <= 2; ++px)
        for (int z8 = 0; z8 < 2; ++z8)
            for (int comp = 0; comp < 2; ++comp) {
                const void *k = px == 2 ? (z8 ? elecKernelV<2, 8>(comp, j->clamp) : elecKernelV<2, 4>(comp, j->clamp))
                                        : (z8 ? elecKernelV<1, 8>(comp, j->clamp) : elecKernelV<1, 4>(comp, j->clamp));
                FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&j->elecPerCU[px - 1][z8][comp], k,
                                                                     fgk::kElecThreads, lds));
            }
    return FG_OK;
}

// Share of the device's workgroup slots one launch of shape (px, pz, tile,
// bxShift) keeps busy.  A launch of fewer items than slots leaves slots
// idle, so its time measures one item's latency rather than throughput.
int elecFill(const fg_job *j, int px, int pz, int bxShift, bool comp, double &fill)
{
    const fgk::ElecArgs a = elecArgs(j, px, pz, 32, bxShift);
    const void *k = px == 2 ? (pz == 8 ? elecKernelV<2, 8>(comp, j->clamp) : elecKernelV<2, 4>(comp, j->clamp))
                            : (pz == 8 ? elecKernelV<1, 8>(comp, j->clamp) : elecKernelV<1, 4>(comp, j->clamp));
    int perCU = 0;
    FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k, fgk::kElecThreads,
                                                         size_t(j->tableLen) * sizeof(float)));
    // the persistent grid of the timed launch (timeShape: one slot per CU
    // left free, as beside a cutoff phase); its workgroups take items until
    // none are left, so a launch of fewer items than workgroups keeps only
    // that share of them busy, and a larger one all of them
    const double slots = double(std::max(perCU - 1, 1)) * double(j->ctx->numCUs);
    fill = std::min(1.0, double(a.items) / slots);
    return FG_OK;
}

// Padded xy area of a 256-lane brick of width 1 << shift over the grid.
long paddedArea(const fg_job *j, int shift)
{
    const int bxl = 1 << shift, byl = fgk::kElecThreads >> shift;
    return long((j->g.nx + bxl - 1) / bxl) * bxl * long((j->g.ny + byl - 1) / byl) * byl;
}

int ilog2(int64_t v)
{
    int r = 0;
    while (v > 1) {
        v >>= 1;
        ++r;
    }
    return r;
}

int tuneJobs(fg_ctx *ctx, fg_job *const *jobs, int n, bool comp)
{
    // kernel variants (points per lane along x, along z); a batch launches one
    static const int kVar[4][2] = {{1, 4}, {1, 8}, {2, 4}, {2, 8}};
    static const int kTiles[3] = {32, 64, 128}, kShifts[4] = {3, 4, 5, 0};
    constexpr float kNone = 1e30f;
    // size classes: (atom octave, point octave, best-padding brick width)
    struct Cls {
        int key[3];
        std::vector<int> members;
        float t[4][3][4];
    };
    std::vector<Cls> cls;
    for (int i = 0; i < n; ++i) {
        const fg_job *j = jobs[i];
        int best = 4;
        for (int sh : {3, 5})
            if (paddedArea(j, sh) < paddedArea(j, best)) best = sh;
        const int key[3] = {ilog2(std::max(j->N, 1)), ilog2(j->P), best};
        auto it = std::find_if(cls.begin(), cls.end(), [&](const Cls &c) {
            return c.key[0] == key[0] && c.key[1] == key[1] && c.key[2] == key[2];
        });
        if (it == cls.end()) {
            cls.push_back(Cls{{key[0], key[1], key[2]}, {}, {}});
            it = cls.end() - 1;
        }
        it->members.push_back(i);
    }
    double total[4] = {0, 0, 0, 0};
    std::string report;
    for (Cls &c : cls) {
        std::sort(c.members.begin(), c.members.end(), [&](int x, int y) {
            return double(jobs[x]->N) * double(jobs[x]->P) < double(jobs[y]->N) * double(jobs[y]->P);
        });
        fg_job *rep = jobs[c.members[c.members.size() / 2]];
        for (int v = 0; v < 4; ++v)
            for (int b = 0; b < 3; ++b)
                for (int d = 0; d < 4; ++d) {
                    c.t[v][b][d] = kNone;
                    if (int rc = timeShape(rep, kVar[v][0], kVar[v][1], kTiles[b], kShifts[d], comp, c.t[v][b][d]))
                        return rc;
                    if (n > 1) {
                        // jobs tuned together run as a batch, which fills the
                        // device with all of their bricks: compare shapes by
                        // throughput (time x share of the slots kept busy)
                        double fill = 1;
                        if (int rc = elecFill(rep, kVar[v][0], kVar[v][1], kShifts[d], comp, fill)) return rc;
                        c.t[v][b][d] = float(c.t[v][b][d] * fill);
                    }
                }
        const double repWork = std::max(1.0, double(rep->N) * double(rep->P));
        double work = 0;
        for (int i : c.members) work += double(jobs[i]->N) * double(jobs[i]->P);
        for (int v = 0; v < 4; ++v) {
            float best = kNone;
            for (int b = 0; b < 3; ++b)
                for (int d = 0; d < 4; ++d) best = std::min(best, c.t[v][b][d]);
            total[v] += best * work / repWork;
        }
    }
    int vi = 0;
    for (int v = 1; v < 4; ++v)
        if (total[v] < total[vi]) vi = v;
    char line[320];
    std::snprintf(line, sizeof line,
                  "px %d pz %d for all jobs (modelled totals: px1 pz4 %.3f ms, px1 pz8 %.3f ms, px2 pz4 %.3f ms, "
                  "px2 pz8 %.3f ms)%s\n",
                  kVar[vi][0], kVar[vi][1], total[0], total[1], total[2], total[3],
                  n > 1 ? "; times x device fill (throughput of a batch)" : "");
    report += line;
    {
        const fg_job *j0 = jobs[0];
        std::snprintf(line, sizeof line,
                      "e-kernel workgroups per CU (compensated; first job's table of %d entries): px1 pz4 %d, px1 pz8 %d, "
                      "px2 pz4 %d, px2 pz8 %d\n",
                      j0->tableLen, j0->elecPerCU[0][0][1], j0->elecPerCU[0][1][1], j0->elecPerCU[1][0][1],
                      j0->elecPerCU[1][1][1]);
        report += line;
    }
    for (Cls &c : cls) {
        int bb = 0, bd = 0;
        for (int b = 0; b < 3; ++b)
            for (int d = 0; d < 4; ++d)
                if (c.t[vi][b][d] < c.t[vi][bb][bd]) {
                    bb = b;
                    bd = d;
                }
        for (int i : c.members) {
            jobs[i]->px = kVar[vi][0];
            jobs[i]->pz = kVar[vi][1];
            jobs[i]->tile = kTiles[bb];
            jobs[i]->bxShift = kShifts[bd];
        }
        const fg_job *rep = jobs[c.members[c.members.size() / 2]];
        std::snprintf(line, sizeof line,
                      "class atoms~2^%d points~2^%d bx%d: %zu jobs, median %d atoms %dx%dx%d -> tile %d bx %d "
                      "(%.4f ms; default shape %.4f ms)\n",
                      c.key[0], c.key[1], 1 << c.key[2], c.members.size(), rep->N, rep->g.nx, rep->g.ny, rep->g.nz,
                      kTiles[bb], kShifts[bd] == 0 ? fgk::kElecThreads : 1 << kShifts[bd], c.t[vi][bb][bd],
                      c.t[0][1][1]);
        report += line;
    }
    if (n == 1) {
        // A single job's e-kernel runs concurrently with its cutoff kernels,
        // which take the workgroup slots it leaves, so the fastest e-kernel
        // alone need not give the fastest step: among the shapes within 5 %
        // of it (at most 4), keep the one whose whole step (all maps) is
        // fastest.
        const Cls &c = cls[0];
        fg_job *j = jobs[0];
        float fastest = kNone;
        for (int v = 0; v < 4; ++v)
            for (int b = 0; b < 3; ++b)
                for (int d = 0; d < 4; ++d) fastest = std::min(fastest, c.t[v][b][d]);
        struct Cand {
            float t;
            int v, b, d;
        };
        std::vector<Cand> cand;
        for (int v = 0; v < 4; ++v)
            for (int b = 0; b < 3; ++b)
                for (int d = 0; d < 4; ++d)
                    if (c.t[v][b][d] <= 1.05f * fastest) cand.push_back({c.t[v][b][d], v, b, d});
        std::sort(cand.begin(), cand.end(), [](const Cand &x, const Cand &y) { return x.t < y.t; });
        if (cand.size() > 4) cand.resize(4);
        const int prec = comp ? FG_PREC_COMPENSATED : FG_PREC_FP32;
        float bestStep = kNone;
        Cand pick = cand[0];
        for (const Cand &k : cand) {
            j->px = kVar[k.v][0];
            j->pz = kVar[k.v][1];
            j->tile = kTiles[k.b];
            j->bxShift = kShifts[k.d];
            float step = kNone;
            for (int r = 0; r < 3; ++r) {
                if (int rc = fg_job_run(ctx, j, FG_RUN_ALL, prec)) return rc;
                fg_timing tm{};
                if (int rc = fg_last_timing(ctx, &tm)) return rc;
                if (r > 0) step = std::min(step, float(tm.total_ms));  // the first run warms up
            }
            if (step < bestStep) {
                bestStep = step;
                pick = k;
            }
        }
        j->px = kVar[pick.v][0];
        j->pz = kVar[pick.v][1];
        j->tile = kTiles[pick.b];
        j->bxShift = kShifts[pick.d];
        std::snprintf(line, sizeof line,
                      "whole step (all maps) over %zu shapes within 5 %% of the fastest e-kernel: px %d pz %d tile %d "
                      "bx %d, %.4f ms (its e-kernel alone %.4f ms)\n",
                      cand.size(), j->px, j->pz, j->tile, j->bxShift == 0 ? fgk::kElecThreads : 1 << j->bxShift,
                      bestStep, pick.t);
        report += line;
    }
    ctx->tuneReport = report;
    return FG_OK;
}

}  // namespace

// ------------------------------------------------------------------------
// Batches
// ------------------------------------------------------------------------
struct fg_batch {
    fg_ctx *ctx = nullptr;
    std::vector<fg_job *> jobs;  // largest first
    bool aff = false, near = false, elec = false, comp = true, fmap = false;
    int MM = 4, KH = 0, px = 1, pz = 4;
    bool pointAng = false, clamp = false, typed = true;  // typed: every job has type-sorted lists
    int sharedTableLen = 0, maxTableLen = 0;
    int cutItems = 0, elecItems = 0, elecWGs = 0, farItems = 0;
    int cutPrio = 0;  // CutoffArgs::prio of every job of the batch
    size_t cutLds = 0;  // LDS of one workgroup of the batch's cutoff kernel
    fgk::CutoffArgs *dCut = nullptr;
    fgk::ListArgs *dList = nullptr;  // the candidate-list passes of every job
    fgk::ElecArgs *dElec = nullptr;
    int32_t *dCutBase = nullptr, *dElecBase = nullptr, *dFarBase = nullptr, *dCounters = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void destroyBatch(fg_batch *b)
{
    if (!b) return;
    if (b->ctx) {
        (void)hipStreamSynchronize(b->ctx->stream);
        forget(b->ctx->batches, b);
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : {static_cast<void *>(b->dCut), static_cast<void *>(b->dList), static_cast<void *>(b->dElec),
                    static_cast<void *>(b->dCutBase), static_cast<void *>(b->dElecBase),
                    static_cast<void *>(b->dFarBase), static_cast<void *>(b->dCounters)})
        if (p) (void)hipFree(p);
    delete b;
}

template <class T>
int batchUpload(T *&dst, const std::vector<T> &src)
{
    FG_HIP(hipMalloc(&dst, std::max<size_t>(src.size(), 1) * sizeof(T)));
    if (!src.empty()) FG_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return FG_OK;
}

// Workgroups of a persistent grid: the kernel's occupancy per CU, less
// `reserve` slots per CU left to kernels running concurrently, x CUs.
template <class K>
int persistentWGs(fg_batch *b, K kernel, int threads, size_t lds, int items, int &out, int reserve = 0)
{
    int perCU = 0;
    FG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void *>(kernel), threads, lds));
    out = std::max(1, std::min(items, std::max(perCU - reserve, 1) * b->ctx->numCUs));
    return FG_OK;
}

// One workgroup per (job, brick) item; with launch false, nothing to size.
template <int MM, int KH, bool kAff, bool kPA, bool kTY>
int launchCutoffBatch(fg_batch *b, bool launch)
{
    if (!launch) return FG_OK;
    hipLaunchKernelGGL((fgk::k_cutoff_batch<MM, KH, kAff, kPA, kTY>), dim3(unsigned(b->cutItems)),
                       dim3(fgk::kCutThreads), 0, b->ctx->stream, b->dCut, b->dCutBase, int(b->jobs.size()));
    return FG_OK;
}

template <int MM, int KH>
int launchCutoffBatchPA(fg_batch *b, bool launch)
{
    if (b->typed)
        return b->pointAng ? launchCutoffBatch<MM, KH, true, true, true>(b, launch)
                           : launchCutoffBatch<MM, KH, true, false, true>(b, launch);
    return b->pointAng ? launchCutoffBatch<MM, KH, true, true, false>(b, launch)
                       : launchCutoffBatch<MM, KH, true, false, false>(b, launch);
}

template <int MM>
int launchCutoffBatchKH(fg_batch *b, bool launch)
{
    if (b->KH == 0) return launchCutoffBatchPA<MM, 0>(b, launch);
    if (b->KH <= 2) return launchCutoffBatchPA<MM, 2>(b, launch);
    if (MM < 16 || b->KH <= 4) return launchCutoffBatchPA<MM, 4>(b, launch);
    if constexpr (MM == 16) return launchCutoffBatchPA<16, 16>(b, launch);
    return FG_OK;
}

template <int PX, int PZ, bool kComp, bool kClamp>
int launchElecBatch(fg_batch *b, bool launch, hipStream_t s)
{
    auto k = fgk::k_elec_batch<PX, PZ, kComp, kClamp>;
    const size_t lds = size_t(b->sharedTableLen > 0 ? b->sharedTableLen : b->maxTableLen) * sizeof(float);
    // a persistent grid would hold every slot until the queue is empty: keep
    // one per CU, and the LDS of one cutoff workgroup, for the concurrent
    // cutoff kernel
    if (!b->elecWGs) {
        const bool shared = b->aff || b->near || b->fmap;
        if (int rc = persistentWGs(b, k, fgk::kElecThreads, lds, b->elecItems, b->elecWGs, shared ? 1 : 0))
            return rc;
        if (shared && b->cutLds > 0 && b->ctx->ldsPerCU > b->cutLds) {
            const int fit = int((b->ctx->ldsPerCU - b->cutLds) / (lds + 256));
            b->elecWGs = std::max(1, std::min(b->elecWGs, std::max(fit, 1) * b->ctx->numCUs));
        }
    }
    if (!launch) return FG_OK;
    hipLaunchKernelGGL(k, dim3(b->elecWGs), dim3(fgk::kElecThreads), lds, s, b->dElec, b->dElecBase,
                       int(b->jobs.size()), b->dCounters + 1, b->sharedTableLen);
    return FG_OK;
}

template <int PX, int PZ>
int launchElecBatchV(fg_batch *b, bool launch, hipStream_t s)
{
    if (b->comp)
        return b->clamp ? launchElecBatch<PX, PZ, true, true>(b, launch, s)
                        : launchElecBatch<PX, PZ, true, false>(b, launch, s);
    return b->clamp ? launchElecBatch<PX, PZ, false, true>(b, launch, s)
                    : launchElecBatch<PX, PZ, false, false>(b, launch, s);
}

// The batch's launch sequence.  The electrostatic kernel runs on the
// context's side stream, forked from its stream after the counters are
// cleared and joined before the final e = near + far sum, concurrently with
// the cutoff phase (under capture the fork and join become graph edges).
// `mid` (if any) is recorded after the cutoff phase and e0/e1 bracket the
// electrostatic kernel; neither is allowed under capture.  With launch
// false, only sizes the persistent grids.
int enqueueBatch(fg_batch *b, hipEvent_t mid, bool launch = true)
{
    fg_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream, side = ctx->side;
    const int n = int(b->jobs.size());
    if (launch) hipLaunchKernelGGL(fgk::k_fill_i32, dim3(1), dim3(256), 0, s, b->dCounters, 2, 0);
    if (b->elec) {
        if (launch) {
            FG_HIP(hipEventRecord(ctx->fork, s));
            FG_HIP(hipStreamWaitEvent(side, ctx->fork, 0));
            if (mid) FG_HIP(hipEventRecord(ctx->e0, side));
        }
        int rc = b->px == 2 ? (b->pz == 8 ? launchElecBatchV<2, 8>(b, launch, side)
                                          : launchElecBatchV<2, 4>(b, launch, side))
                            : (b->pz == 8 ? launchElecBatchV<1, 8>(b, launch, side)
                                          : launchElecBatchV<1, 4>(b, launch, side));
        if (rc) return rc;
        if (launch) FG_HIP(hipEventRecord(ctx->e1, side));
    }
    if (b->aff || b->near || b->fmap) {
        // every job's candidate lists, then their type-sorted copies (for the
        // typed kernel), before the cutoff phase, as a single run makes them
        if (launch) {
            hipLaunchKernelGGL(fgk::k_brick_lists_batch<true>, dim3(b->cutItems), dim3(fgk::kListThreads), 0, s,
                               b->dList, b->dCutBase, n);
            if (b->typed && b->aff)
                hipLaunchKernelGGL(fgk::k_type_sort_batch, dim3(b->cutItems), dim3(fgk::kListThreads), 0, s, b->dList,
                                   b->dCutBase, n);
        }
        if (b->fmap && launch)
            hipLaunchKernelGGL(fgk::k_fmap_batch, dim3(b->cutItems), dim3(fgk::kCutThreads), 0, s, b->dCut, b->dCutBase,
                               n);
        int rc = FG_OK;
        if (!b->aff && !b->near) rc = FG_OK;
        else if (!b->aff) rc = launchCutoffBatch<1, 0, false, false, false>(b, launch);
        else if (b->MM == 4) rc = launchCutoffBatchKH<4>(b, launch);
        else if (b->MM == 6) rc = launchCutoffBatchKH<6>(b, launch);
        else if (b->MM == 8) rc = launchCutoffBatchKH<8>(b, launch);
        else rc = launchCutoffBatchKH<16>(b, launch);
        if (rc) return rc;
    }
    if (mid && launch) FG_HIP(hipEventRecord(mid, s));
    if (b->elec && launch) {
        FG_HIP(hipStreamWaitEvent(s, ctx->e1, 0));
        if (b->near)  // e = near (k_cutoff_batch) + far (k_elec_batch)
            hipLaunchKernelGGL(fgk::k_add_far_batch, dim3(b->farItems), dim3(256), 0, s, b->dElec, b->dFarBase, n);
    }
    FG_HIP(hipGetLastError());
    return FG_OK;
}

int buildBatch(fg_ctx *ctx, fg_batch *b, fg_job *const *jobs, int n, unsigned what, bool comp, bool useGraph)
{
    b->ctx = ctx;
    b->jobs.assign(jobs, jobs + n);
    // largest jobs first: the items at the end of the queue are the short ones
    std::stable_sort(b->jobs.begin(), b->jobs.end(), [](const fg_job *x, const fg_job *y) { return x->N > y->N; });
    b->comp = comp;
    b->elec = (what & FG_RUN_ELEC) != 0;
    b->near = b->elec && comp;
    b->aff = (what & FG_RUN_AFFINITY) != 0;
    for (fg_job *j : b->jobs) b->fmap = b->fmap || ((what & FG_RUN_FLOATING) && j->floating);
    // one kernel variant (points per lane along x and z) for the batch: the
    // one carrying the most work
    // jobs never given a shape run at the one the autotuner picks for most
    // pockets of a batch: 8 points per lane along z, 256-column bricks,
    // 64-atom tiles (shapes change no bits)
    const bool batchDefault = n > 1;
    auto shapeOf = [&](const fg_job *j, int &px, int &pz, int &tile, int &shift) {
        if (batchDefault && !j->shaped) {
            px = 1;
            pz = 8;
            tile = 64;
            shift = 0;
        } else {
            px = j->px;
            pz = j->pz;
            tile = j->tile;
            shift = j->bxShift;
        }
    };
    double work[4] = {0, 0, 0, 0};
    int maxM = 0;
    for (fg_job *j : b->jobs) {
        int jpx, jpz, jtile, jshift;
        shapeOf(j, jpx, jpz, jtile, jshift);
        work[(jpx == 2 ? 2 : 0) + (jpz == 8 ? 1 : 0)] += double(j->N) * double(j->P);
        maxM = std::max(maxM, j->M);
        b->KH = std::max(b->KH, j->KH);
        b->pointAng = b->pointAng || j->pointAng;
        b->clamp = b->clamp || j->clamp;
        b->maxTableLen = std::max(b->maxTableLen, j->tableLen);
        // a job without type-sorted lists (more than kMaxTypes receptor
        // types) makes the batch visit every job's candidates in list order:
        // the same maps up to the FP64 summation order of the job's own run
        b->typed = b->typed && j->typed;
    }
    {
        int v = 0;
        for (int k = 1; k < 4; ++k)
            if (work[k] > work[v]) v = k;
        b->px = v >= 2 ? 2 : 1;
        b->pz = v & 1 ? 8 : 4;
    }
    // Tables longer than kTableLen entries (jobs whose distances reach past
    // 81.9 A) would cost the persistent e-kernel workgroups per CU, and the
    // longest table sets the LDS of every workgroup: the batch runs the
    // clamped variant on the first kTableLen entries instead (entries past
    // 8191 repeat entry 8191, so the clamped index reads the same value).
    const bool clampAll = b->elec && b->maxTableLen > fgk::kTableLen;
    if (clampAll) {
        b->clamp = true;
        b->maxTableLen = fgk::kTableLen;
    }
    b->MM = maxM <= 4 ? 4 : b->KH > 4 ? 16 : maxM <= 6 ? 6 : maxM <= 8 ? 8 : 16;
    if (b->aff && maxM == 0) b->aff = false;
    b->cutLds = !b->aff ? sizeof(fgk::CutoffLds<1, false, false>)
                : b->MM == 4 ? (b->typed ? sizeof(fgk::CutoffLds<4, true, true>) : sizeof(fgk::CutoffLds<4, true, false>))
                : b->MM == 6 ? (b->typed ? sizeof(fgk::CutoffLds<6, true, true>) : sizeof(fgk::CutoffLds<6, true, false>))
                : b->MM == 8 ? (b->typed ? sizeof(fgk::CutoffLds<8, true, true>) : sizeof(fgk::CutoffLds<8, true, false>))
                             : (b->typed ? sizeof(fgk::CutoffLds<16, true, true>) : sizeof(fgk::CutoffLds<16, true, false>));
    // one LDS table for all jobs when their dielectric tables are bit-equal
    {
        bool same = true;
        for (const fg_job *j : b->jobs) same = same && j->epsId >= 0 && j->epsId == b->jobs[0]->epsId;
        b->sharedTableLen = same ? b->maxTableLen : 0;
    }
    {
        double cand = 0, pairs = 0;
        for (const fg_job *j : b->jobs) {
            cand += double(j->listTotal);
            pairs += double(j->N) * double(j->P);
        }
        b->cutPrio = cutoffPriority(cand, pairs);
    }
    std::vector<fgk::CutoffArgs> ca;
    std::vector<fgk::ListArgs> la;
    std::vector<fgk::ElecArgs> ea;
    std::vector<int32_t> cutBase(1, 0), elecBase(1, 0), farBase(1, 0);
    for (fg_job *j : b->jobs) {
        ca.push_back(cutoffArgs(j, b->near ? 1 : 0));
        ca.back().prio = b->cutPrio;
        if (!(b->typed && b->aff)) ca.back().brickList = j->dList;
        la.push_back(listArgs(j));
        if (!(b->typed && b->aff)) la.back().sorted = nullptr;  // no type sort
        int jpx, jpz, jtile, jshift;
        shapeOf(j, jpx, jpz, jtile, jshift);
        fgk::ElecArgs e = elecArgs(j, b->px, b->pz, jtile, jshift);
        if (clampAll) {
            e.tableLen = std::min(e.tableLen, int32_t(fgk::kTableLen));
            e.clampMax = e.tableLen - 1;
            e.tableCut2 = tableCut2Of(e.tableLen, j->elecEd);
        }
        if (b->sharedTableLen > 0) e.clampMax = b->sharedTableLen - 1;
        ea.push_back(e);
        const int64_t c = int64_t(cutBase.back()) + j->nb, el = int64_t(elecBase.back()) + int64_t(e.items);
        if (c > INT32_MAX || el > INT32_MAX) return fgc::fail(FG_E_ARG, "fg_batch_create: too many work items");
        cutBase.push_back(int32_t(c));
        elecBase.push_back(int32_t(el));
        const int64_t fa = int64_t(farBase.back()) + (j->P + 255) / 256;
        if (fa > INT32_MAX) return fgc::fail(FG_E_ARG, "fg_batch_create: too many work items");
        farBase.push_back(int32_t(fa));
    }
    b->cutItems = cutBase.back();
    b->elecItems = elecBase.back();
    b->farItems = farBase.back();
    if (int rc = batchUpload(b->dCut, ca)) return rc;
    if (int rc = batchUpload(b->dList, la)) return rc;
    if (int rc = batchUpload(b->dElec, ea)) return rc;
    if (int rc = batchUpload(b->dCutBase, cutBase)) return rc;
    if (int rc = batchUpload(b->dElecBase, elecBase)) return rc;
    if (int rc = batchUpload(b->dFarBase, farBase)) return rc;
    if (int rc = batchUpload(b->dCounters, std::vector<int32_t>{0, 0})) return rc;
    // persistent grid sizes (queried outside any capture)
    if (int rc = enqueueBatch(b, nullptr, false)) return rc;
    if (!useGraph) return FG_OK;
    hipStream_t s = ctx->stream;
    FG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueueBatch(b, nullptr);
    if (rc) joinSide(ctx);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fgc::fail(FG_E_DEVICE, std::string("fg_batch_create: capture: ") + hipGetErrorString(e));
    b->graph = g;
    FG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return FG_OK;
}

}  // namespace

extern "C" {

int fg_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int fg_create(int device, fg_ctx **out)
{
    if (!out) return fgc::fail(FG_E_ARG, "fg_create: out is NULL");
    *out = nullptr;
    int n = fg_device_count();
    if (n <= 0) return fgc::fail(FG_E_DEVICE, "fg_create: no GPU is visible (libfgmaps has no CPU fallback)");
    if (device < 0 || device >= n) return fgc::fail(FG_E_ARG, "fg_create: device index out of range");
    FG_HIP(hipSetDevice(device));
    fg_ctx *c = new fg_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
        fg_destroy(c);
        return fgc::fail(FG_E_DEVICE, "fg_create: hipStreamCreate failed");
    }
    for (hipEvent_t *e : {&c->ev[0], &c->ev[1], &c->ev[2], &c->ev[3], &c->fork, &c->e0, &c->e1, &c->join})
        if (hipEventCreate(e) != hipSuccess) {
            fg_destroy(c);
            return fgc::fail(FG_E_DEVICE, "fg_create: hipEventCreate failed");
        }
    if (hipDeviceGetAttribute(&c->numCUs, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        c->numCUs <= 0)
        c->numCUs = 256;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess &&
        lds > 0)
        c->ldsPerCU = size_t(lds);
    if (int rc = scratchSetUp(device)) {
        fg_destroy(c);
        return rc;
    }
    *out = c;
    return FG_OK;
}

void fg_destroy(fg_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (fg_job *j : ctx->jobs) j->ctx = nullptr;
    for (fg_batch *b : ctx->batches) b->ctx = nullptr;
    for (fg_mapset *m : ctx->mapsets) m->ctx = nullptr;
    for (hipEvent_t e : {ctx->ev[0], ctx->ev[1], ctx->ev[2], ctx->ev[3], ctx->fork, ctx->e0, ctx->e1, ctx->join,
                         ctx->pinEv[0], ctx->pinEv[1]})
        if (e) (void)hipEventDestroy(e);
    for (void *p : ctx->pin)
        if (p) (void)hipHostFree(p);
    if (ctx->upPin) (void)hipHostFree(ctx->upPin);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    delete ctx;
}

namespace {

// Host half of job creation: checks the description, builds every input on
// the host and stages it for the job's first arena.  No HIP calls, so many
// jobs can be built on parallel threads.  On success *out is a job that is
// not yet registered with ctx.
int jobBuildHost(fg_ctx *ctx, const fg_job_desc *d, fg_job **out)
{
    *out = nullptr;
    const fg_grid &g = d->grid;
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || g.nx > 1025 || g.ny > 1025 || g.nz > 1025 || !(g.spacing > 0))
        return fgc::fail(FG_E_ARG, "fg_job_create: invalid grid");
    if (d->atoms.n < 0 || (d->atoms.n > 0 && (!d->atoms.x || !d->atoms.y || !d->atoms.z)))
        return fgc::fail(FG_E_ARG, "fg_job_create: invalid atoms");
    if (d->n_maps < 0 || d->n_maps > 14) return fgc::fail(FG_E_ARG, "fg_job_create: n_maps must be 0..14");
    fg_job *j = new fg_job();
    // j->ctx stays null until the host build succeeds: destroying a job that
    // failed here touches nothing shared (builds may run on several threads)
    j->g = fgk::GridGeom{g.nx, g.ny, g.nz, g.div2x, g.div2y, g.div2z, g.spacing};
    j->P = int64_t(g.nx) * g.ny * g.nz;
    j->N = d->atoms.n;
    j->M = d->n_maps;
    j->T = d->tables.n_types;
    // the kernels index the pair table with 32-bit offsets, formed by a
    // 24-bit multiply of (type x 801 + bin) by the map count
    if (j->T < 0 || int64_t(j->T) * fgk::kPairBins * std::max(j->M, 1) >= (int64_t(1) << 31) ||
        int64_t(j->T) * fgk::kPairBins >= (int64_t(1) << 24)) {
        delete j;
        return fgc::fail(FG_E_ARG, "fg_job_create: pair table too large (n_types x 801 x n_maps >= 2^31)");
    }
    j->distDepDiel = d->dist_dep_diel;
    j->hydrogen = d->hydrogen;
    j->covOn = d->covalent ? 0 : d->cov_on;
    for (int k = 0; k < 3; ++k) j->covPoint[k] = d->cov_point[k];
    j->covHwSqInv = d->cov_hw_sq_inv;
    j->covBarrier = d->cov_barrier;
    j->floating = d->floating_grid != 0;
    const fg_atoms &A = d->atoms;
    j->hasGeom = A.charge && A.vol && A.solpar && A.type && A.hbond && d->geom.rvector && d->geom.rvector2 &&
                 d->geom.rexp && d->geom.disorder && d->tables.pair && d->tables.desolv && (j->M == 0 || d->maps);
    if (j->M > 0 && !j->hasGeom) {
        destroyJob(j);
        return fgc::fail(FG_E_ARG, "fg_job_create: affinity maps need charge/vol/solpar/type/hbond, bond geometry and tables");
    }
    if (d->dist_dep_diel && !d->tables.eps) {
        destroyJob(j);
        return fgc::fail(FG_E_ARG, "fg_job_create: distance-dependent dielectric needs tables.eps");
    }
    if (j->hasGeom) {
        for (int i = 0; i < j->N; ++i)
            if (A.type[i] < 0 || A.type[i] >= std::max(j->T, 1) || A.hbond[i] < 0 || A.hbond[i] > 5) {
                destroyJob(j);
                return fgc::fail(FG_E_ARG, "fg_job_create: atom type/hbond out of range of the tables");
            }
        if (j->hydrogen >= std::max(j->T, 1)) {
            destroyJob(j);
            return fgc::fail(FG_E_ARG, "fg_job_create: hydrogen type index out of range");
        }
    }
    // Per map, what it does with each atom H-bond class (FG/CalculateGridmaps.cpp:206-230);
    // maps that keep an H-bond min/max go first in kernel order.
    std::vector<int32_t> actions(size_t(j->M), 0);
    std::vector<bool> keeps(size_t(j->M), false);
    for (int m = 0; m < j->M; ++m) {
        const fg_map_params &p = d->maps[m];
        int32_t acts = 0;
        for (int h = 0; h < 6; ++h) {
            int act = fgk::kActAdd;
            const bool aDon = h == fgk::HB_DS || h == fgk::HB_D1;
            if (p.is_hbonder) {
                if ((p.hbond == fgk::HB_AS || p.hbond == fgk::HB_A2) && aDon) act = fgk::kActDonor;
                else if (p.hbond == fgk::HB_A1 && aDon) act = fgk::kActMinRacc;
                else if ((p.hbond == fgk::HB_DS || p.hbond == fgk::HB_D1) && h >= fgk::HB_AS) act = fgk::kActMinRdon;
            }
            acts |= act << (2 * h);
            if (act == fgk::kActMinRacc || act == fgk::kActMinRdon) keeps[size_t(m)] = !p.is_covalent;
        }
        actions[size_t(m)] = acts;
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int m = 0; m < j->M; ++m)
            if (keeps[size_t(m)] == (pass == 0)) j->order.push_back(m);
    j->KH = int(std::count(keeps.begin(), keeps.end(), true));
    bool anyCovalent = false;
    for (int m = 0; m < j->M; ++m) anyCovalent = anyCovalent || d->maps[m].is_covalent;
    for (int h = 0; h < 6 && !anyCovalent; ++h) {
        bool add = true;
        for (int m = 0; m < j->M; ++m) add = add && ((actions[size_t(m)] >> (2 * h)) & 3) == fgk::kActAdd;
        if (add) j->addOnlyClasses |= 1 << h;
    }
    j->mp.resize(size_t(j->M));
    for (int k = 0; k < j->M; ++k) {
        const int m = j->order[size_t(k)];
        j->mp[size_t(k)] = fgk::MapParam{actions[size_t(m)], d->maps[m].is_covalent, m, 0, d->maps[m].solpar_probe,
                                         d->maps[m].vol_probe};
    }
    double gridMaxAbs = 0;
    const int n[3] = {g.nx, g.ny, g.nz}, dv[3] = {g.div2x, g.div2y, g.div2z};
    for (int k = 0; k < 3; ++k)
        gridMaxAbs = std::max({gridMaxAbs, std::fabs(double(0 - dv[k]) * g.spacing), std::fabs(double(n[k] - 1 - dv[k]) * g.spacing)});
    // default shape.  Small grids: 2D bricks of 16 lanes along x, one point
    // per lane along x and 4 along z.  Grids of at least 1024 such bricks of
    // 8 planes: linear items of 2 x 4 points per lane, tile 128 (what the
    // autotuner picks on 127^3 and 161^3 grids), with 8 planes per lane when
    // that still leaves about 1.25 items per workgroup of the persistent
    // launch (3 per CU).
    {
        const long bricks8 = long((g.nx + 15) / 16) * ((g.ny + 15) / 16) * ((g.nz + 7) / 8);
        if (bricks8 >= 1024) {
            const long items8 = long(g.nx) * g.ny * g.nz / (2 * 8 * fgk::kElecThreads);
            j->px = 2;
            j->bxShift = 0;
            j->tile = 128;
            j->pz = items8 >= 960 ? 8 : 4;
        }
    }
    // inputs and outputs staged for one arena
    int rc = buildCutoff(j, *d);
    if (!rc) rc = buildElecTiles(ctx, j, *d, gridMaxAbs);
    if (!rc) rc = buildPlain(j, *d);
    if (!rc) rc = devAlloc(j, j->dMaps, size_t(j->M + 2) * size_t(j->P));
    if (!rc) rc = devAlloc(j, j->dFar, size_t(j->P));
    if (!rc) rc = devAlloc(j, j->dCounter, size_t(1));
    if (!rc) rc = devAlloc(j, j->dSummary, size_t(2) * size_t(j->M + 2) + 1);
    if (!rc) rc = devAlloc(j, j->dSummaryPart, size_t(2) * size_t(j->M + 2) * fgk::kSumChunks);
    if (!rc) rc = devAlloc(j, j->dMilli, size_t(j->M + 2) * size_t(j->P) + 1);
    if (!rc && A.charge && j->N > 0) rc = devUploadDirect(j, j->dCharge, A.charge, size_t(j->N));
    if (!rc && j->floating) rc = devAlloc(j, j->dFmap, size_t(j->P));
    if (rc) {
        destroyJob(j);
        return rc;
    }
    j->ctx = ctx;
    *out = j;
    return FG_OK;
}

// Places the slots a job recorded at `base` (its arena) and clears them.
void placeArena(fg_job *j, char *base)
{
    for (const fg_job::Slot &sl : j->slots) *sl.dst = base + sl.off;
    j->slots.clear();
    j->arenaBytes = 0;
}

// One device allocation of `bytes`, freed when the last owner lets go.
int sharedBlock(size_t bytes, std::shared_ptr<void> &out)
{
    void *p = nullptr;
    FG_HIP(hipMalloc(&p, std::max<size_t>(bytes, 1)));
    out.reset(p, [](void *q) { (void)hipFree(q); });
    return FG_OK;
}

// Host bytes `src` that go to byte `dst` of the shared block, staged at
// byte `pin` of the context's pinned upload buffer.
struct Upload {
    size_t dst;
    const void *src;
    size_t bytes, pin;
};

// Copies the uploads' bytes into ctx->upPin on up to `threads` threads, in
// pieces of at most 2 MiB.
void stageUploads(fg_ctx *ctx, const std::vector<Upload> &ups, int threads)
{
    struct Piece {
        const unsigned char *src;
        unsigned char *dst;
        size_t bytes;
    };
    std::vector<Piece> pieces;
    size_t total = 0;
    for (const Upload &u : ups)
        for (size_t o = 0; o < u.bytes; o += size_t(2) << 20) {
            const size_t b = std::min(u.bytes - o, size_t(2) << 20);
            pieces.push_back(Piece{static_cast<const unsigned char *>(u.src) + o, ctx->upPin + u.pin + o, b});
            total += b;
        }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < pieces.size();)
            std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
    };
    threads = int(std::min<size_t>(size_t(std::max(threads, 1)), std::max<size_t>(total >> 21, 1)));
    std::vector<std::thread> pool;
    try {
        for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    } catch (const std::system_error &) {
        // fewer threads: the running ones share the pieces
    }
    work();
    for (std::thread &t : pool) t.join();
}

// Device half of creating jobs[0..n) on ctx's stream, with two device
// allocations for all of them: the first arenas (inputs and outputs, cleared,
// then the staged inputs and the caller arrays copied in from pinned memory),
// then, after the count passes and one read-back of every job's list totals,
// the candidate lists' arenas.  Registers the jobs with ctx.  The caller-owned
// arrays of the descriptions are read here (direct uploads).  On failure
// every job is destroyed.
int jobsFinishDevice(fg_ctx *ctx, fg_job *const *jobs, int n, int threads)
{
    auto fail = [&](int rc) {
        (void)hipStreamSynchronize(ctx->stream);
        for (int i = 0; i < n; ++i) destroyJob(jobs[i]);
        return rc;
    };
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    // the first arenas, then two list totals per job
    std::vector<size_t> off(static_cast<size_t>(n));
    std::vector<Upload> ups;
    size_t total = 0, pinBytes = 0;
    for (int i = 0; i < n; ++i) {
        fg_job *j = jobs[i];
        off[size_t(i)] = total;
        auto add = [&](size_t dst, const void *src, size_t bytes) {
            if (!bytes) return;
            ups.push_back(Upload{dst, src, bytes, pinBytes});
            pinBytes = up(pinBytes + bytes);
        };
        add(total, j->stage.data(), j->stage.size());
        for (const fg_job::Slot &sl : j->slots)
            if (sl.direct) add(total + sl.off, sl.direct, sl.bytes);
        total = up(total + j->arenaBytes);
    }
    const size_t totOff = total;
    total += size_t(2) * size_t(n) * sizeof(int32_t);
    // pinned staging: the host copies run on many threads and the DMAs do not
    // block the host; pageable copies when pinned memory cannot be had
    if (pinBytes > ctx->upPinBytes) {
        if (ctx->upPin) (void)hipHostFree(ctx->upPin);
        ctx->upPin = nullptr;
        ctx->upPinBytes = 0;
        void *p = nullptr;
        if (hipHostMalloc(&p, pinBytes, hipHostMallocDefault) == hipSuccess) {
            ctx->upPin = static_cast<unsigned char *>(p);
            ctx->upPinBytes = pinBytes;
        } else {
            (void)hipGetLastError();
        }
    }
    const bool pinned = ctx->upPin != nullptr && pinBytes <= ctx->upPinBytes;
    if (pinned) stageUploads(ctx, ups, threads);
    std::shared_ptr<void> block;
    if (int rc = sharedBlock(total, block)) return fail(rc);
    char *base = static_cast<char *>(block.get());
    for (int i = 0; i < n; ++i) {
        placeArena(jobs[i], base + off[size_t(i)]);
        jobs[i]->owned.push_back(block);
    }
    int32_t *dTot = reinterpret_cast<int32_t *>(base + totOff);
    std::vector<int32_t> tot(size_t(2) * size_t(n));
    hipError_t e = hipMemsetAsync(base, 0, total, ctx->stream);
    for (size_t u = 0; u < ups.size() && e == hipSuccess; ++u)
        e = hipMemcpyAsync(base + ups[u].dst, pinned ? static_cast<const void *>(ctx->upPin + ups[u].pin) : ups[u].src,
                           ups[u].bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail(fgc::fail(FG_E_DEVICE, std::string("fg_job_create: ") + hipGetErrorString(e)));
    for (int i = 0; i < n; ++i) {
        if (int rc = elecOccupancy(jobs[i])) return fail(rc);
        if (int rc = runListCount(jobs[i], listArgs(jobs[i]), dTot + 2 * i)) return fail(rc);
    }
    e = hipMemcpyAsync(tot.data(), dTot, tot.size() * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(fgc::fail(FG_E_DEVICE, std::string("fg_job_create: ") + hipGetErrorString(e)));
    // the candidate lists
    size_t total2 = 0;
    for (int i = 0; i < n; ++i) {
        fg_job *j = jobs[i];
        std::vector<unsigned char>().swap(j->stage);
        off[size_t(i)] = total2;
        const int32_t t0 = tot[size_t(2 * i)], t1 = tot[size_t(2 * i + 1)];
        if (t0 < 0 || t1 < 0) return fail(fgc::fail(FG_E_DEVICE, "fg_job_create: candidate count overflow"));
        j->listTotal = t0;
        devAlloc(j, j->dList, size_t(t0));
        devAlloc(j, j->dHList, size_t(t1));
        if (j->typed) {
            devAlloc(j, j->dSorted, size_t(t0));
            devAlloc(j, j->dTypeStart, size_t(j->nb) * size_t(j->T + 1));
        }
        total2 = up(total2 + j->arenaBytes);
    }
    std::shared_ptr<void> block2;
    if (int rc = sharedBlock(total2, block2)) return fail(rc);
    for (int i = 0; i < n; ++i) {
        placeArena(jobs[i], static_cast<char *>(block2.get()) + off[size_t(i)]);
        jobs[i]->owned.push_back(block2);
    }
    // the lists themselves, once
    for (int i = 0; i < n; ++i)
        if (int rc = fillLists(jobs[i])) return fail(rc);
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(fgc::fail(FG_E_DEVICE, std::string("fg_job_create: ") + hipGetErrorString(e)));
    for (int i = 0; i < n; ++i) ctx->jobs.push_back(jobs[i]);
    return FG_OK;
}

}  // namespace

int fg_job_create(fg_ctx *ctx, const fg_job_desc *d, fg_job **out)
{
    if (!ctx || !d || !out) return fgc::fail(FG_E_ARG, "fg_job_create: NULL argument");
    *out = nullptr;
    FG_HIP(hipSetDevice(ctx->device));
    fg_job *j = nullptr;
    if (int rc = jobBuildHost(ctx, d, &j)) return rc;
    if (int rc = jobsFinishDevice(ctx, &j, 1, 8)) return rc;
    *out = j;
    return FG_OK;
}

int fg_job_create_many(fg_ctx *ctx, const fg_job_desc *descs, int n, int threads, fg_job **out)
{
    if (!ctx || !out || n < 0 || (!descs && n > 0)) return fgc::fail(FG_E_ARG, "fg_job_create_many: NULL argument");
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    FG_HIP(hipSetDevice(ctx->device));
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    const int copyThreads = std::min(threads, 16);
    threads = std::min(threads, std::max(n, 1));
    std::vector<int> rcs(size_t(n), FG_OK);
    std::vector<std::string> errs(static_cast<size_t>(n));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i; (i = next.fetch_add(1)) < n;) {
            rcs[size_t(i)] = jobBuildHost(ctx, &descs[i], &out[i]);
            if (rcs[size_t(i)] != FG_OK) errs[size_t(i)] = fg_last_error();
        }
    };
    std::vector<std::thread> pool;
    try {
        for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    } catch (const std::system_error &) {
        // fewer threads than asked for: the ones running (and this one) share the work
    }
    work();
    for (std::thread &t : pool) t.join();
    int bad = -1, rc = FG_OK;
    std::string msg;
    for (int i = 0; i < n && bad < 0; ++i)
        if (rcs[size_t(i)] != FG_OK) {
            bad = i;
            rc = rcs[size_t(i)];
            msg = errs[size_t(i)];
        }
    if (bad >= 0) {
        for (int i = 0; i < n; ++i) {
            destroyJob(out[i]);  // host halves only: nothing on the device yet
            out[i] = nullptr;
        }
        return fgc::fail(rc, "job " + std::to_string(bad) + ": " + msg);
    }
    // the device halves of all jobs together, on the context's stream
    if ((rc = jobsFinishDevice(ctx, out, n, copyThreads)) != FG_OK) {
        for (int i = 0; i < n; ++i) out[i] = nullptr;  // destroyed by jobsFinishDevice
        return rc;
    }
    return FG_OK;
}

void fg_job_destroy(fg_job *job)
{
    if (job && job->ctx) (void)hipStreamSynchronize(job->ctx->stream);
    destroyJob(job);
}

// The launches of one run on the two streams.  Under capture the timing
// events become event-record nodes of the graph (kGraphTimed) or are left out
// (kGraphUntimed); the fork and join events only order the two streams.
enum RunMode { kStreams, kGraphTimed, kGraphUntimed };
static int enqueueRun(fg_ctx *ctx, fg_job *j, unsigned what, int precision, RunMode mode)
{
    auto rec = [&](hipEvent_t ev, hipStream_t st) {
        if (mode == kGraphUntimed) return hipSuccess;
        return mode == kGraphTimed ? hipEventRecordWithFlags(ev, st, hipEventRecordExternal) : hipEventRecord(ev, st);
    };
    hipStream_t s = ctx->stream;
    const bool elec = what & FG_RUN_ELEC, aff = (what & FG_RUN_AFFINITY) != 0;
    const bool near = elec && precision == FG_PREC_COMPENSATED;
    const bool fmap = (what & FG_RUN_FLOATING) && j->floating;
    FG_HIP(rec(ctx->ev[0], s));
    // the electrostatic kernels run on the side stream, concurrently with the
    // cutoff phase: they write e (or, compensated, the far part) and read
    // nothing the cutoff kernels write.  They are enqueued first: theirs is
    // the longer chain, and its long workgroups must hold the device from
    // the start (the cutoff kernels fill the slots they leave); enqueued after
    // the lists, the cutoff kernels took the slots first and L's step grew
    // from 23.5 to 33.5 ms.
    if (elec) {
        FG_HIP(hipEventRecord(ctx->fork, s));
        FG_HIP(hipStreamWaitEvent(ctx->side, ctx->fork, 0));
        FG_HIP(rec(ctx->e0, ctx->side));
        if (precision == FG_PREC_PLAIN) {
            if (int rc = runPlainElec(j, ctx->side)) return rc;
        } else if (int rc = launchElec(j, precision, (aff || near || fmap) ? 1 : 0, ctx->side)) {
            return rc;
        }
        FG_HIP(rec(ctx->e1, ctx->side));
        FG_HIP(hipEventRecord(ctx->join, ctx->side));
    }
    if (precision == FG_PREC_PLAIN) {
        // the plain cutoff kernel reads the unsorted lists
        if ((aff && j->hasGeom) || fmap)
            if (int rc = fillLists(j, false)) return rc;
        if (aff && j->hasGeom)
            if (int rc = runPlainAffinity(j)) return rc;
        if (fmap)
            if (int rc = launchFmap(j)) return rc;
    } else if (aff || near || fmap) {
        const bool affMaps = aff && j->hasGeom;
        if (int rc = fillLists(j)) return rc;
        if (aff || near)
            if (int rc = launchCutoff(j, affMaps, near ? 1 : 0)) return rc;
        if (fmap)
            if (int rc = launchFmap(j)) return rc;
    }
    FG_HIP(rec(ctx->ev[1], s));
    if (elec) {
        FG_HIP(hipStreamWaitEvent(s, ctx->join, 0));
        if (near)  // e = near (k_cutoff) + far (k_elec)
            hipLaunchKernelGGL(fgk::k_add_far, dim3(unsigned((j->P + 255) / 256)), dim3(256), 0, s, j->P, j->dFar,
                               j->dMaps + int64_t(j->M) * j->P);
        FG_HIP(hipGetLastError());
    }
    FG_HIP(rec(ctx->ev[2], s));
    return FG_OK;
}

int fg_job_run(fg_ctx *ctx, fg_job *j, unsigned what, int precision)
{
    if (!ctx || !j || j->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_job_run: job does not belong to this context");
    if (precision < FG_PREC_PLAIN || precision > FG_PREC_FP32) return fgc::fail(FG_E_ARG, "fg_job_run: bad precision");
    if ((what & FG_RUN_AFFINITY) && j->M > 0 && !j->hasGeom)
        return fgc::fail(FG_E_ARG, "fg_job_run: job has no affinity inputs");
    FG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (precision == FG_PREC_COMPENSATED || precision == FG_PREC_FP32) {
        // One graph per (maps, precision, shape, phase timing), captured at
        // its first run and replayed: launched as one graph without event
        // nodes, the e-kernel and the cutoff phase share the device the same
        // way every time (L: 23.4 ms per run back to back), while launched
        // on the two streams run after run they did not (28.6 ms; 23.4-25 ms
        // with a wait after each run).  Event-record nodes bring back the
        // 28 ms when runs follow each other without a wait, so the phases are
        // timed only when asked for; the run's total is timed around it.
        const bool phases = (what & FG_RUN_PHASE_TIMING) != 0;
        const int key[5] = {precision, j->px, j->pz, j->tile, j->bxShift};
        if (!j->runExec || j->runWhat != what || std::memcmp(key, j->runKey, sizeof key) != 0) {
            // earlier replays of the old graph may still be queued
            if (j->runExec) FG_HIP(hipStreamSynchronize(s));
            if (j->runExec) (void)hipGraphExecDestroy(j->runExec);
            if (j->runGraph) (void)hipGraphDestroy(j->runGraph);
            j->runExec = nullptr;
            j->runGraph = nullptr;
            FG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = enqueueRun(ctx, j, what, precision, phases ? kGraphTimed : kGraphUntimed);
            if (rc) joinSide(ctx);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : fgc::fail(FG_E_DEVICE, std::string("fg_job_run: capture: ") + hipGetErrorString(e));
            }
            j->runGraph = g;
            FG_HIP(hipGraphInstantiate(&j->runExec, g, nullptr, nullptr, 0));
            j->runWhat = what;
            std::memcpy(j->runKey, key, sizeof key);
        }
        if (!phases) FG_HIP(hipEventRecord(ctx->ev[0], s));
        FG_HIP(hipGraphLaunch(j->runExec, s));
        if (!phases) FG_HIP(hipEventRecord(ctx->ev[2], s));
        ctx->splitTiming = phases;
    } else if (int rc = enqueueRun(ctx, j, what, precision, kStreams)) {
        return rc;
    } else {
        ctx->splitTiming = true;
    }
    ctx->haveTiming = true;
    ctx->elecTimed = (what & FG_RUN_ELEC) != 0;
    return FG_OK;
}

int fg_sync(fg_ctx *ctx)
{
    if (!ctx) return fgc::fail(FG_E_ARG, "fg_sync: NULL context");
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return FG_OK;
}

int fg_last_timing(fg_ctx *ctx, fg_timing *t)
{
    if (!ctx || !t) return fgc::fail(FG_E_ARG, "fg_last_timing: NULL argument");
    if (!ctx->haveTiming) return fgc::fail(FG_E_ARG, "fg_last_timing: nothing has run");
    FG_HIP(hipEventSynchronize(ctx->ev[2]));
    float a = -1, b = -1, c = 0;
    if (ctx->splitTiming) {
        // the phases overlap: e0..e1 ran on the side stream within ev[0]..ev[2]
        FG_HIP(hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
        b = 0;
        if (ctx->elecTimed) FG_HIP(hipEventElapsedTime(&b, ctx->e0, ctx->e1));
    }
    FG_HIP(hipEventElapsedTime(&c, ctx->ev[0], ctx->ev[2]));
    t->affinity_ms = a;
    t->elec_ms = b;
    t->total_ms = c;
    return FG_OK;
}

namespace {

// Device-to-host copies go through the context's two pinned staging buffers:
// the DMA of one buffer overlaps the host's copy of the other out to the
// caller's (pageable) memory, on several host threads for large pieces.
// Copying straight to pageable memory runs at a fraction of the link's rate.
constexpr size_t kStageBytes = size_t(32) << 20;

struct HostSeg {
    void *dst;
    const void *src;
    size_t bytes;
};

struct StagePart {
    char *dst;
    size_t off, bytes;  // place in the staging buffer
};

// Copies the parts of one staging buffer out to their destinations: the
// buffer's bytes are split into equal ranges, one per host thread (up to 8,
// at least 2 MiB each), whatever the parts' sizes.
void copyOut(const char *pin, const std::vector<StagePart> &parts)
{
    size_t total = 0;
    for (const StagePart &pt : parts) total = std::max(total, pt.off + pt.bytes);
    constexpr size_t kPerThread = size_t(2) << 20;
    const int n = int(std::max<size_t>(1, std::min<size_t>(8, total / kPerThread)));
    auto range = [&](size_t lo, size_t hi) {
        for (const StagePart &pt : parts) {
            const size_t a = std::max(lo, pt.off), b = std::min(hi, pt.off + pt.bytes);
            if (a < b) std::memcpy(pt.dst + (a - pt.off), pin + a, b - a);
        }
    };
    if (n == 1) {
        range(0, total);
        return;
    }
    const size_t step = (total + size_t(n) - 1) / size_t(n);
    std::vector<std::thread> pool;
    for (int t = 1; t < n; ++t) pool.emplace_back(range, step * size_t(t), std::min(total, step * size_t(t + 1)));
    range(0, std::min(total, step));
    for (auto &th : pool) th.join();
}

// Copies every segment (device src -> host dst) in order and returns when all
// are on the host.  Segments whose sources follow each other share DMAs.
int fetchSegments(fg_ctx *ctx, const std::vector<HostSeg> &segs)
{
    if (!ctx->pin[0]) {
        for (int b = 0; b < 2; ++b) {
            FG_HIP(hipHostMalloc(&ctx->pin[b], kStageBytes, hipHostMallocDefault));
            FG_HIP(hipEventCreateWithFlags(&ctx->pinEv[b], hipEventDisableTiming));
        }
    }
    std::vector<StagePart> parts[2];
    int cur = 0;
    size_t fill = 0;
    const char *runSrc = nullptr;  // the device run being gathered into buffer cur
    size_t runOff = 0, runLen = 0;
    auto flushRun = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (runLen) e = hipMemcpyAsync(static_cast<char *>(ctx->pin[cur]) + runOff, runSrc, runLen, hipMemcpyDeviceToHost,
                                       ctx->stream);
        runLen = 0;
        return e;
    };
    bool pending = false;  // buffer cur ^ 1 holds a DMA whose host copy is due
    auto finish = [&](int b) -> hipError_t {
        const hipError_t e = hipEventSynchronize(ctx->pinEv[b]);
        if (e != hipSuccess) return e;
        copyOut(static_cast<const char *>(ctx->pin[b]), parts[b]);
        parts[b].clear();
        return hipSuccess;
    };
    // closes buffer cur (its DMAs and event), then copies the other one out
    auto rotate = [&]() -> int {
        FG_HIP(flushRun());
        FG_HIP(hipEventRecord(ctx->pinEv[cur], ctx->stream));
        if (pending) FG_HIP(finish(cur ^ 1));
        pending = true;
        cur ^= 1;
        fill = 0;
        return FG_OK;
    };
    for (const HostSeg &sg : segs) {
        size_t done = 0;
        while (done < sg.bytes) {
            if (fill == kStageBytes)
                if (int rc = rotate()) return rc;
            const size_t len = std::min(sg.bytes - done, kStageBytes - fill);
            const char *src = static_cast<const char *>(sg.src) + done;
            if (runLen && runSrc + runLen == src) {
                runLen += len;
            } else {
                FG_HIP(flushRun());
                runSrc = src;
                runOff = fill;
                runLen = len;
            }
            parts[cur].push_back(StagePart{static_cast<char *>(sg.dst) + done, fill, len});
            fill += len;
            done += len;
        }
    }
    if (fill)
        if (int rc = rotate()) return rc;
    if (pending) FG_HIP(finish(cur ^ 1));
    return FG_OK;
}

// The fetch segments of job j: its maps into out[m] (a null entry is
// skipped), the floating grid (floats) into fm.
void jobSegments(const fg_job *j, double *const *out, std::vector<float> &fm, std::vector<HostSeg> &segs)
{
    const size_t bytes = size_t(j->P) * sizeof(double);
    for (int m = 0; m < j->M + 2; ++m)
        if (out[m]) segs.push_back(HostSeg{out[m], j->dMaps + int64_t(m) * j->P, bytes});
    if (j->floating && out[j->M + 2]) {
        fm.resize(size_t(j->P));
        segs.push_back(HostSeg{fm.data(), j->dFmap, fm.size() * sizeof(float)});
    }
}

}  // namespace

int fg_job_fetch_many(fg_ctx *ctx, fg_job *const *jobs, int n, double *const *const *outs)
{
    if (!ctx || (n > 0 && (!jobs || !outs))) return fgc::fail(FG_E_ARG, "fg_job_fetch_many: NULL argument");
    for (int i = 0; i < n; ++i)
        if (!jobs[i] || jobs[i]->ctx != ctx || !outs[i])
            return fgc::fail(FG_E_ARG, "fg_job_fetch_many: job " + std::to_string(i) + " is NULL or not of this context");
    FG_HIP(hipSetDevice(ctx->device));
    std::vector<HostSeg> segs;
    std::vector<std::vector<float>> fms(static_cast<size_t>(std::max(n, 0)));
    for (int i = 0; i < n; ++i) jobSegments(jobs[i], outs[i], fms[size_t(i)], segs);
    if (int rc = fetchSegments(ctx, segs)) return rc;
    for (int i = 0; i < n; ++i)
        for (size_t p = 0; p < fms[size_t(i)].size(); ++p) outs[i][jobs[i]->M + 2][p] = double(fms[size_t(i)][p]);
    return FG_OK;
}

int fg_job_fetch(fg_ctx *ctx, fg_job *j, double *const *out)
{
    if (!ctx || !j || !out) return fgc::fail(FG_E_ARG, "fg_job_fetch: NULL argument");
    fg_job *js[1] = {j};
    double *const *os[1] = {out};
    return fg_job_fetch_many(ctx, js, 1, os);
}

int fg_job_fetch_milli_many(fg_ctx *ctx, fg_job *const *jobs, int n, int32_t *const *const *outs)
{
    if (!ctx || (n > 0 && (!jobs || !outs))) return fgc::fail(FG_E_ARG, "fg_job_fetch_milli_many: NULL argument");
    for (int i = 0; i < n; ++i)
        if (!jobs[i] || jobs[i]->ctx != ctx || !outs[i])
            return fgc::fail(FG_E_ARG,
                             "fg_job_fetch_milli_many: job " + std::to_string(i) + " is NULL or not of this context");
    FG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // round every job's maps on the device (each job's buffer made on first use)
    for (int i = 0; i < n; ++i) {
        fg_job *j = jobs[i];
        int32_t *overflow = j->dMilli + size_t(j->M + 2) * size_t(j->P);
        const int64_t cnt = int64_t(j->M + 2) * j->P;
        hipLaunchKernelGGL(fgk::k_fill_i32, dim3(1), dim3(256), 0, s, overflow, 1, 0);
        hipLaunchKernelGGL(fgk::k_round_milli, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, s, j->dMaps, cnt,
                           j->dMilli, overflow);
    }
    FG_HIP(hipGetLastError());
    std::vector<int32_t> over(static_cast<size_t>(std::max(n, 0)), 0);
    std::vector<HostSeg> segs;
    for (int i = 0; i < n; ++i) {
        const fg_job *j = jobs[i];
        segs.push_back(HostSeg{&over[size_t(i)], j->dMilli + size_t(j->M + 2) * size_t(j->P), sizeof(int32_t)});
        for (int m = 0; m < j->M + 2; ++m)
            if (outs[i][m])
                segs.push_back(HostSeg{outs[i][m], j->dMilli + int64_t(m) * j->P, size_t(j->P) * sizeof(int32_t)});
    }
    if (int rc = fetchSegments(ctx, segs)) return rc;
    for (int i = 0; i < n; ++i)
        if (over[size_t(i)])
            return fgc::fail(FG_E_ARG, "fg_job_fetch_milli_many: job " + std::to_string(i) +
                                           " has map values whose thousandths exceed 32 bits; fetch it as FP64");
    return FG_OK;
}

int fg_job_fetch_milli(fg_ctx *ctx, fg_job *j, int32_t *const *out)
{
    if (!ctx || !j || !out) return fgc::fail(FG_E_ARG, "fg_job_fetch_milli: NULL argument");
    fg_job *js[1] = {j};
    int32_t *const *os[1] = {out};
    return fg_job_fetch_milli_many(ctx, js, 1, os);
}

int fg_job_cutoff_pairs(fg_ctx *ctx, fg_job *j, int64_t *pairs)
{
    if (!ctx || !j || !pairs) return fgc::fail(FG_E_ARG, "fg_job_cutoff_pairs: NULL argument");
    FG_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    fgk::CutoffArgs ca = cutoffArgs(j, 0);
    ca.brickList = j->dList;  // the unsorted lists
    unsigned long long *d = nullptr;
    FG_HIP(hipMalloc(&d, sizeof *d));
    unsigned long long h = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof *d, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fgk::k_count_pairs, dim3(j->nb), dim3(fgk::kCutThreads), 0, s, ca, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return fgc::fail(FG_E_DEVICE, std::string("fg_job_cutoff_pairs: ") + hipGetErrorString(e));
    *pairs = int64_t(h);
    return FG_OK;
}

int fg_job_summary(fg_ctx *ctx, fg_job *j, fg_summary *out)
{
    if (!ctx || !j || !out) return fgc::fail(FG_E_ARG, "fg_job_summary: NULL argument");
    if (j->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_job_summary: job does not belong to this context");
    FG_HIP(hipSetDevice(ctx->device));
    const int nm = j->M + 2;
    if (nm > 256) return fgc::fail(FG_E_ARG, "fg_job_summary: more than 256 maps");
    hipLaunchKernelGGL(fgk::k_job_summary, dim3(unsigned(fgk::kSumChunks), unsigned(nm + 1)), dim3(256), 0,
                       ctx->stream, j->dMaps, j->P, nm, j->dCharge, j->dCharge ? j->N : 0, j->dSummaryPart,
                       j->dSummary);
    FG_HIP(hipGetLastError());
    hipLaunchKernelGGL(fgk::k_job_summary_final, dim3(1), dim3(256), 0, ctx->stream, j->dSummaryPart, nm,
                       j->dSummary);
    FG_HIP(hipGetLastError());
    std::vector<double> h(size_t(2 * nm + 1));
    FG_HIP(hipMemcpyAsync(h.data(), j->dSummary, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    *out = fg_summary{};
    out->n_maps = nm;
    for (int m = 0; m < nm; ++m) {
        out->min[m] = h[size_t(2 * m)];
        out->max[m] = h[size_t(2 * m + 1)];
    }
    out->has_charge = j->dCharge ? 1 : 0;
    const double q = j->dCharge ? h[size_t(2 * nm)] : 0.0;
    out->total_charge = q;
    out->charge_residual = std::fabs(q - std::nearbyint(q));
    out->charge_integral = j->dCharge && out->charge_residual < 0.005 ? 1 : 0;
    return FG_OK;
}

int fg_job_get_shape(const fg_job *job, fg_shape *out)
{
    if (!job || !out) return fgc::fail(FG_E_ARG, "fg_job_get_shape: NULL argument");
    out->px = job->px;
    out->pz = job->pz;
    out->tile = job->tile;
    out->bx = job->bxShift == 0 ? fgk::kElecThreads : 1 << job->bxShift;
    return FG_OK;
}

int fg_job_set_shape(fg_job *job, const fg_shape *sh)
{
    if (!job || !sh) return fgc::fail(FG_E_ARG, "fg_job_set_shape: NULL argument");
    const int shift = sh->bx == 8 ? 3 : sh->bx == 16 ? 4 : sh->bx == 32 ? 5 : sh->bx == 256 ? 0 : -1;
    if ((sh->px != 1 && sh->px != 2) || (sh->pz != 4 && sh->pz != 8) ||
        (sh->tile != 32 && sh->tile != 64 && sh->tile != 128) || shift < 0)
        return fgc::fail(FG_E_ARG, "fg_job_set_shape: px must be 1 or 2, pz 4 or 8, tile 32/64/128, bx 8/16/32/256");
    job->px = sh->px;
    job->pz = sh->pz;
    job->tile = sh->tile;
    job->bxShift = shift;
    job->shaped = true;
    return FG_OK;
}

int fg_job_set_plain_dielectric(fg_job *job, int kind)
{
    if (!job) return fgc::fail(FG_E_ARG, "fg_job_set_plain_dielectric: NULL job");
    if (kind < FG_PLAIN_DIEL_TABLE || kind > FG_PLAIN_DIEL_TEXTURE)
        return fgc::fail(FG_E_ARG, "fg_job_set_plain_dielectric: unknown variant");
    job->plainDiel = kind;
    return FG_OK;
}

int fg_job_plain_dielectric(const fg_job *job)
{
    if (!job) return fgc::fail(FG_E_ARG, "fg_job_plain_dielectric: NULL job");
    switch (plainDielKind(job)) {
    case fgk::kPlainInPlace: return FG_PLAIN_DIEL_INPLACE;
    case fgk::kPlainTexture: return FG_PLAIN_DIEL_TEXTURE;
    default: return FG_PLAIN_DIEL_TABLE;
    }
}

int fg_autotune(fg_ctx *ctx, fg_job *const *jobs, int n, int precision)
{
    if (!ctx || (!jobs && n > 0)) return fgc::fail(FG_E_ARG, "fg_autotune: NULL argument");
    if (precision != FG_PREC_COMPENSATED && precision != FG_PREC_FP32)
        return fgc::fail(FG_E_ARG, "fg_autotune: precision must be FG_PREC_COMPENSATED or FG_PREC_FP32");
    for (int i = 0; i < n; ++i)
        if (!jobs[i] || jobs[i]->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_autotune: job does not belong to this context");
    FG_HIP(hipSetDevice(ctx->device));
    const int rc = tuneJobs(ctx, jobs, n, precision == FG_PREC_COMPENSATED);
    if (!rc)
        for (int i = 0; i < n; ++i) jobs[i]->shaped = true;
    return rc;
}

const char *fg_tune_report(const fg_ctx *ctx) { return ctx ? ctx->tuneReport.c_str() : ""; }

int fg_batch_create(fg_ctx *ctx, fg_job *const *jobs, int n, unsigned what, int precision, unsigned flags,
                    fg_batch **out)
{
    if (!ctx || !out || (!jobs && n > 0)) return fgc::fail(FG_E_ARG, "fg_batch_create: NULL argument");
    *out = nullptr;
    if (precision != FG_PREC_COMPENSATED && precision != FG_PREC_FP32)
        return fgc::fail(FG_E_ARG, "fg_batch_create: precision must be FG_PREC_COMPENSATED or FG_PREC_FP32");
    if (n <= 0) return fgc::fail(FG_E_ARG, "fg_batch_create: no jobs");
    for (int i = 0; i < n; ++i) {
        if (!jobs[i] || jobs[i]->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_batch_create: job does not belong to this context");
        if ((what & FG_RUN_AFFINITY) && jobs[i]->M > 0 && !jobs[i]->hasGeom)
            return fgc::fail(FG_E_ARG, "fg_batch_create: a job has no affinity inputs");
    }
    FG_HIP(hipSetDevice(ctx->device));
    fg_batch *b = new fg_batch();
    int rc = buildBatch(ctx, b, jobs, n, what, precision == FG_PREC_COMPENSATED, (flags & FG_BATCH_NO_GRAPH) == 0);
    if (rc) {
        destroyBatch(b);
        return rc;
    }
    ctx->batches.push_back(b);
    *out = b;
    return FG_OK;
}

int fg_batch_launch(fg_ctx *ctx, fg_batch *b)
{
    if (!ctx || !b || b->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_batch_launch: batch does not belong to this context");
    FG_HIP(hipSetDevice(ctx->device));
    FG_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    if (b->exec) {
        FG_HIP(hipGraphLaunch(b->exec, ctx->stream));
        ctx->splitTiming = false;
    } else {
        if (int rc = enqueueBatch(b, ctx->ev[1])) return rc;
        ctx->splitTiming = true;
        ctx->elecTimed = b->elec;
    }
    FG_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    ctx->haveTiming = true;
    return FG_OK;
}

void fg_batch_destroy(fg_batch *b) { destroyBatch(b); }

int fg_batch_run(fg_ctx *ctx, fg_job *const *jobs, int n, unsigned what, int precision)
{
    fg_batch *b = nullptr;
    if (int rc = fg_batch_create(ctx, jobs, n, what, precision, 0, &b)) return rc;
    int rc = fg_batch_launch(ctx, b);
    if (!rc) rc = fg_sync(ctx);
    fg_batch_destroy(b);
    return rc;
}

// The one-shot entry points run a job: created from the caller's host
// buffers (copied before the call returns), run on `streams` and fetched into
// `outs` by fg_wait.  The asynchronous forms give the job a context of its own
// (streams on the caller's device), so that work the caller enqueues on its
// context meanwhile runs concurrently with it.
struct fg_handle {
    fg_ctx *own = nullptr;  // the handle's context (asynchronous forms), destroyed by fg_wait
    fg_ctx *ctx = nullptr;  // the context the job runs on
    fg_job *job = nullptr;
    std::vector<double *> outs;
};

namespace {

int startOneShot(fg_ctx *ctx, bool async, const fg_job_desc &d, unsigned what, int precision,
                 std::vector<double *> outs, fg_handle **out)
{
    *out = nullptr;
    fg_handle *h = new fg_handle();
    h->ctx = ctx;
    if (async) {
        if (int rc = fg_create(ctx->device, &h->own)) {
            delete h;
            return rc;
        }
        h->ctx = h->own;
    }
    int rc = fg_job_create(h->ctx, &d, &h->job);
    if (!rc) rc = fg_job_run(h->ctx, h->job, what, precision);
    if (rc) {
        if (h->job) fg_job_destroy(h->job);
        if (h->own) fg_destroy(h->own);
        delete h;
        return rc;
    }
    h->outs = std::move(outs);
    *out = h;
    return FG_OK;
}

// An e-map job has no affinity maps: its outputs are [e, d].
fg_job_desc elecDesc(const fg_grid *grid, const fg_atoms *atoms, const double *eps, int ddd)
{
    fg_job_desc d{};
    d.grid = *grid;
    d.atoms.n = atoms->n;
    d.atoms.x = atoms->x;
    d.atoms.y = atoms->y;
    d.atoms.z = atoms->z;
    d.atoms.w = atoms->w;
    d.tables.eps = eps;
    d.dist_dep_diel = ddd;
    d.hydrogen = -1;
    return d;
}

}  // namespace

int fg_wait(fg_handle *h)
{
    if (!h) return FG_OK;
    int rc = fg_job_fetch(h->ctx, h->job, h->outs.data());
    fg_job_destroy(h->job);
    if (h->own) fg_destroy(h->own);
    delete h;
    return rc;
}

int fg_elec_map_async(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, const double *eps, int ddd,
                      int precision, double *e_out, fg_handle **out)
{
    if (!out) return fgc::fail(FG_E_ARG, "fg_elec_map_async: out is NULL");
    *out = nullptr;
    if (!ctx || !grid || !atoms || !e_out) return fgc::fail(FG_E_ARG, "fg_elec_map_async: NULL argument");
    return startOneShot(ctx, true, elecDesc(grid, atoms, eps, ddd), FG_RUN_ELEC, precision, {e_out, nullptr}, out);
}

int fg_elec_map(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, const double *eps, int ddd, int precision,
                double *e_out)
{
    if (!ctx || !grid || !atoms || !e_out) return fgc::fail(FG_E_ARG, "fg_elec_map: NULL argument");
    fg_handle *h = nullptr;
    if (int rc = startOneShot(ctx, false, elecDesc(grid, atoms, eps, ddd), FG_RUN_ELEC, precision, {e_out, nullptr}, &h))
        return rc;
    return fg_wait(h);
}

int fg_covalent_map(fg_ctx *ctx, const fg_grid *grid, const double cov_point[3], double hw_sq_inv, double barrier,
                    double *out)
{
    if (!ctx || !grid || !cov_point || !out) return fgc::fail(FG_E_ARG, "fg_covalent_map: NULL argument");
    const fg_grid &g = *grid;
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0 || !(g.spacing > 0)) return fgc::fail(FG_E_ARG, "fg_covalent_map: invalid grid");
    FG_HIP(hipSetDevice(ctx->device));
    fgk::CutoffArgs a{};
    a.g = fgk::GridGeom{g.nx, g.ny, g.nz, g.div2x, g.div2y, g.div2z, g.spacing};
    a.P = int64_t(g.nx) * g.ny * g.nz;
    a.covOn = 1;
    for (int k = 0; k < 3; ++k) a.covPoint[k] = cov_point[k];
    a.covHwSqInv = hw_sq_inv;
    a.covBarrier = barrier;
    double *d = nullptr;
    FG_HIP(hipMalloc(&d, size_t(a.P) * sizeof(double)));
    hipLaunchKernelGGL(fgk::k_covalent, dim3(unsigned((a.P + 255) / 256)), dim3(256), 0, ctx->stream, a, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, size_t(a.P) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fgc::fail(FG_E_DEVICE, std::string("fg_covalent_map: ") + hipGetErrorString(e));
    return FG_OK;
}

int fg_floating_grid(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, float *out)
{
    if (!ctx || !grid || !atoms || !out) return fgc::fail(FG_E_ARG, "fg_floating_grid: NULL argument");
    fg_job_desc d{};
    d.grid = *grid;
    d.atoms.n = atoms->n;
    d.atoms.x = atoms->x;
    d.atoms.y = atoms->y;
    d.atoms.z = atoms->z;
    d.hydrogen = -1;
    d.floating_grid = 1;
    fg_job *j = nullptr;
    if (int rc = fg_job_create(ctx, &d, &j)) return rc;
    int rc = fg_job_run(ctx, j, FG_RUN_FLOATING, FG_PREC_COMPENSATED);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(out, j->dFmap, size_t(j->P) * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fgc::fail(FG_E_DEVICE, std::string("fg_floating_grid: ") + hipGetErrorString(e));
    }
    fg_job_destroy(j);
    return rc;
}

namespace {

int affinityOneShot(fg_ctx *ctx, bool async, const fg_grid *grid, const fg_atoms *atoms, const fg_hbond_geom *geom,
                    const fg_tables *tables, const fg_map_params *maps, int n_maps, double solpar_q, int hydrogen,
                    double *const *maps_out, double *d_out, fg_handle **out)
{
    const char *who = async ? "fg_affinity_d_maps_async" : "fg_affinity_d_maps";
    if (!ctx || !grid || !atoms || !geom || !tables || (!maps && n_maps > 0) || (!maps_out && n_maps > 0) || !d_out)
        return fgc::fail(FG_E_ARG, std::string(who) + ": NULL argument");
    if (n_maps < 0 || n_maps > 14) return fgc::fail(FG_E_ARG, std::string(who) + ": n_maps must be 0..14");
    fg_job_desc d{};
    d.grid = *grid;
    d.atoms = *atoms;
    d.geom = *geom;
    d.tables = *tables;
    d.n_maps = n_maps;
    d.maps = maps;
    d.solpar_q = solpar_q;
    d.hydrogen = hydrogen;
    d.dist_dep_diel = 0;
    // covalent maps arrive as initial values in maps_out (copied into the job
    // before the call returns)
    const int64_t P = int64_t(grid->nx) * grid->ny * grid->nz;
    std::vector<double> cov;
    bool anyCov = false;
    for (int m = 0; m < n_maps; ++m) anyCov = anyCov || maps[m].is_covalent;
    if (anyCov && P > 0) {
        cov.assign(size_t(n_maps) * size_t(P), 0.0);
        for (int m = 0; m < n_maps; ++m)
            if (maps[m].is_covalent && maps_out[m]) std::copy(maps_out[m], maps_out[m] + P, cov.begin() + int64_t(m) * P);
        d.covalent = cov.data();
    }
    std::vector<double *> outs(size_t(n_maps) + 2, nullptr);
    for (int m = 0; m < n_maps; ++m) outs[size_t(m)] = maps_out[m];
    outs[size_t(n_maps) + 1] = d_out;
    return startOneShot(ctx, async, d, FG_RUN_AFFINITY, FG_PREC_FP64, std::move(outs), out);
}

}  // namespace

int fg_affinity_d_maps(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, const fg_hbond_geom *geom,
                       const fg_tables *tables, const fg_map_params *maps, int n_maps, double solpar_q, int hydrogen,
                       double *const *maps_out, double *d_out)
{
    fg_handle *h = nullptr;
    if (int rc = affinityOneShot(ctx, false, grid, atoms, geom, tables, maps, n_maps, solpar_q, hydrogen, maps_out,
                                 d_out, &h))
        return rc;
    return fg_wait(h);
}

int fg_affinity_d_maps_async(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, const fg_hbond_geom *geom,
                             const fg_tables *tables, const fg_map_params *maps, int n_maps, double solpar_q,
                             int hydrogen, double *const *maps_out, double *d_out, fg_handle **out)
{
    if (!out) return fgc::fail(FG_E_ARG, "fg_affinity_d_maps_async: out is NULL");
    *out = nullptr;
    return affinityOneShot(ctx, true, grid, atoms, geom, tables, maps, n_maps, solpar_q, hydrogen, maps_out, d_out,
                           out);
}

// ------------------------------------------------------------------------
// The maps' consumer: AutoDock map sets and trilinear interpolation
// ------------------------------------------------------------------------
namespace {

int checkGrid(const fg_map_grid *g, const char *who)
{
    for (int k = 0; k < 3; ++k)
        if (g->num_points1[k] <= 1 || g->num_points1[k] > 4097 || g->num_points1[k] != g->num_points[k] + 1)
            return fgc::fail(FG_E_ARG, std::string(who) + ": invalid grid");
    if (!(g->inv_spacing > 0)) return fgc::fail(FG_E_ARG, std::string(who) + ": invalid grid spacing");
    return FG_OK;
}

int newMapset(fg_ctx *ctx, const fg_map_grid *g, int nslots, fg_mapset **out)
{
    fg_mapset *ms = new fg_mapset();
    ms->ctx = ctx;
    ms->g = *g;
    ms->nslots = nslots;
    const size_t n = size_t(g->num_points1[0]) * g->num_points1[1] * g->num_points1[2] * size_t(nslots);
    if (hipMalloc(&ms->d, n * sizeof(float)) != hipSuccess) {
        delete ms;
        return fgc::fail(FG_E_DEVICE, "fg_mapset: out of device memory");
    }
    // the corner-packed copy the interpolation reads (3 x 32 bytes per atom
    // instead of 24 scattered floats)
    if (hipMalloc(&ms->cells, n * 8 * sizeof(float)) != hipSuccess) {
        (void)hipFree(ms->d);
        delete ms;
        return fgc::fail(FG_E_DEVICE, "fg_mapset: out of device memory (the map set and its 8x corner-packed copy)");
    }
    ctx->mapsets.push_back(ms);
    *out = ms;
    return FG_OK;
}

// Fills the corner-packed copy from the plain layout.
hipError_t packCells(fg_ctx *ctx, fg_mapset *ms)
{
    const int64_t n = int64_t(ms->g.num_points1[0]) * ms->g.num_points1[1] * ms->g.num_points1[2] * ms->nslots;
    hipLaunchKernelGGL(fgk::k_pack_cells, dim3(unsigned((n + fgk::kInterpThreads - 1) / fgk::kInterpThreads)),
                       dim3(fgk::kInterpThreads), 0, ctx->stream, ms->d, ms->g.num_points1[0], ms->g.num_points1[1],
                       ms->g.num_points1[2], ms->nslots, ms->cells);
    return hipGetLastError();
}

}  // namespace

int fg_mapset_create(fg_ctx *ctx, const fg_map_grid *g, int nslots, const float *values, fg_mapset **out)
{
    if (!ctx || !g || !values || !out) return fgc::fail(FG_E_ARG, "fg_mapset_create: NULL argument");
    *out = nullptr;
    if (nslots <= 0 || nslots > 64) return fgc::fail(FG_E_ARG, "fg_mapset_create: nslots must be 1..64");
    if (int rc = checkGrid(g, "fg_mapset_create")) return rc;
    FG_HIP(hipSetDevice(ctx->device));
    fg_mapset *ms = nullptr;
    if (int rc = newMapset(ctx, g, nslots, &ms)) return rc;
    const size_t n = size_t(g->num_points1[0]) * g->num_points1[1] * g->num_points1[2] * size_t(nslots);
    hipError_t e = hipMemcpy(ms->d, values, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = packCells(ctx, ms);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        fg_mapset_destroy(ms);
        return fgc::fail(FG_E_DEVICE, std::string("fg_mapset_create: ") + hipGetErrorString(e));
    }
    *out = ms;
    return FG_OK;
}

int fg_mapset_from_job(fg_ctx *ctx, fg_job *j, const fg_map_grid *g, fg_mapset **out)
{
    if (!ctx || !j || !g || !out || j->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_mapset_from_job: bad argument");
    *out = nullptr;
    if (int rc = checkGrid(g, "fg_mapset_from_job")) return rc;
    if (g->num_points1[0] != j->g.nx || g->num_points1[1] != j->g.ny || g->num_points1[2] != j->g.nz)
        return fgc::fail(FG_E_ARG, "fg_mapset_from_job: the grid's points do not match the job's");
    FG_HIP(hipSetDevice(ctx->device));
    fg_mapset *ms = nullptr;
    const int V = j->M + 2;
    if (int rc = newMapset(ctx, g, V, &ms)) return rc;
    hipLaunchKernelGGL(fgk::k_export_mapset, dim3(unsigned((j->P + fgk::kInterpThreads - 1) / fgk::kInterpThreads)),
                       dim3(fgk::kInterpThreads), 0, ctx->stream, j->dMaps, j->P, V, ms->d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = packCells(ctx, ms);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        fg_mapset_destroy(ms);
        return fgc::fail(FG_E_DEVICE, std::string("fg_mapset_from_job: ") + hipGetErrorString(e));
    }
    *out = ms;
    return FG_OK;
}

int fg_mapset_fetch(fg_ctx *ctx, const fg_mapset *ms, float *values)
{
    if (!ctx || !ms || !values) return fgc::fail(FG_E_ARG, "fg_mapset_fetch: NULL argument");
    if (ms->ctx != ctx) return fgc::fail(FG_E_ARG, "fg_mapset_fetch: map set does not belong to this context");
    FG_HIP(hipSetDevice(ctx->device));
    const size_t n = size_t(ms->g.num_points1[0]) * ms->g.num_points1[1] * ms->g.num_points1[2] * size_t(ms->nslots);
    FG_HIP(hipMemcpyAsync(values, ms->d, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    FG_HIP(hipStreamSynchronize(ctx->stream));
    return FG_OK;
}

int fg_mapset_info(const fg_mapset *ms, fg_map_grid *g, int *nslots)
{
    if (!ms) return fgc::fail(FG_E_ARG, "fg_mapset_info: NULL argument");
    if (g) *g = ms->g;
    if (nslots) *nslots = ms->nslots;
    return FG_OK;
}

void fg_mapset_destroy(fg_mapset *ms)
{
    if (!ms) return;
    if (ms->ctx) {
        (void)hipStreamSynchronize(ms->ctx->stream);
        forget(ms->ctx->mapsets, ms);
    }
    if (ms->d) (void)hipFree(ms->d);
    if (ms->cells) (void)hipFree(ms->cells);
    delete ms;
}

namespace {

// Device staging of one trilinterp call, one buffer: the poses and the
// ligand's arrays going in, the per-atom and per-pose outputs coming back.
struct InterpLayout {
    size_t xyz, type, q, aq, ign, elec, emap, et, mt, tot, bytes;
};

InterpLayout interpLayout(int64_t n, int A, int npose, bool elec, bool emap)
{
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    InterpLayout L{};
    size_t off = 0;
    L.xyz = off; off = align(off + size_t(n) * 3 * sizeof(float));
    L.type = off; off = align(off + size_t(A) * sizeof(int32_t));
    L.q = off; off = align(off + size_t(A) * sizeof(float));
    L.aq = off; off = align(off + size_t(A) * sizeof(float));
    L.ign = off; off = align(off + size_t(A) * sizeof(int32_t));
    L.elec = off; off = align(off + (elec ? size_t(n) * sizeof(float) : 0));
    L.emap = off; off = align(off + (emap ? size_t(n) * sizeof(float) : 0));
    L.et = off; off = align(off + size_t(npose) * sizeof(double));
    L.mt = off; off = align(off + size_t(npose) * sizeof(double));
    L.tot = off; off = align(off + size_t(npose) * sizeof(float));
    L.bytes = off;
    return L;
}

int interpCheck(const char *who, fg_ctx *ctx, const fg_mapset *ms, const fg_ligand *lig, int npose, const float *xyz)
{
    const std::string w(who);
    if (!ctx || !ms || !lig || (npose > 0 && !xyz)) return fgc::fail(FG_E_ARG, w + ": NULL argument");
    if (ms->ctx != ctx) return fgc::fail(FG_E_ARG, w + ": map set does not belong to this context");
    const int A = lig->natoms, V = ms->nslots;
    if (A <= 0 || npose < 0 || !lig->type || !lig->charge || !lig->abs_charge)
        return fgc::fail(FG_E_ARG, w + ": invalid ligand");
    if (lig->elec_slot < 0 || lig->elec_slot >= V || lig->desolv_slot < 0 || lig->desolv_slot >= V)
        return fgc::fail(FG_E_ARG, w + ": e/d slot out of range");
    for (int i = 0; i < A; ++i)
        if (lig->type[i] < 0 || lig->type[i] >= V) return fgc::fail(FG_E_ARG, w + ": atom type slot out of range");
    if (int64_t(npose) * A > (int64_t(1) << 31)) return fgc::fail(FG_E_ARG, w + ": too many (pose, atom) pairs");
    return FG_OK;
}

// Uploads the poses and the ligand into the staging buffer b (device memory
// laid out as L), interpolates, and copies the outputs back; returns when
// they are on the host.  Every output the kernels write is written whole, so
// b needs no clearing.
int interpRun(const char *who, fg_ctx *ctx, const fg_mapset *ms, const fg_ligand *lig, int npose, const float *xyz,
              float *elec, float *emap, double *elecTotal, double *emapTotal, float *total, char *b,
              const InterpLayout &L)
{
    const int A = lig->natoms, V = ms->nslots;
    const int64_t n = int64_t(npose) * A;
    hipStream_t s = ctx->stream;
    auto up = [&](size_t o, const void *src, size_t bytes) { return hipMemcpyAsync(b + o, src, bytes, hipMemcpyHostToDevice, s); };
    hipError_t e = up(L.xyz, xyz, size_t(n) * 3 * sizeof(float));
    if (e == hipSuccess) e = up(L.type, lig->type, size_t(A) * sizeof(int32_t));
    if (e == hipSuccess) e = up(L.q, lig->charge, size_t(A) * sizeof(float));
    if (e == hipSuccess) e = up(L.aq, lig->abs_charge, size_t(A) * sizeof(float));
    if (e == hipSuccess && lig->ignore) e = up(L.ign, lig->ignore, size_t(A) * sizeof(int32_t));
    fgk::InterpArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = ms->g.lo[k];
        a.hi[k] = ms->g.hi[k];
        a.center[k] = ms->g.center[k];
    }
    a.inv = ms->g.inv_spacing;
    a.nx = ms->g.num_points1[0];
    a.ny = ms->g.num_points1[1];
    a.nz = ms->g.num_points1[2];
    a.V = V;
    a.natoms = A;
    a.npose = npose;
    a.elecSlot = lig->elec_slot;
    a.desolvSlot = lig->desolv_slot;
    a.cells = ms->cells;
    a.xyz = reinterpret_cast<const float *>(b + L.xyz);
    a.type = reinterpret_cast<const int32_t *>(b + L.type);
    a.charge = reinterpret_cast<const float *>(b + L.q);
    a.absq = reinterpret_cast<const float *>(b + L.aq);
    a.ignore = lig->ignore ? reinterpret_cast<const int32_t *>(b + L.ign) : nullptr;
    a.elec = elec ? reinterpret_cast<float *>(b + L.elec) : nullptr;
    a.emap = emap ? reinterpret_cast<float *>(b + L.emap) : nullptr;
    a.elecTotal = reinterpret_cast<double *>(b + L.et);
    a.emapTotal = reinterpret_cast<double *>(b + L.mt);
    a.total = reinterpret_cast<float *>(b + L.tot);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], s);
    if (e == hipSuccess) {
        if (A <= fgk::kInterpThreads) {
            const int ppw = fgk::kInterpThreads / A;
            const unsigned groups = unsigned((npose + ppw - 1) / ppw);
            hipLaunchKernelGGL(fgk::k_interp_poses, dim3(groups), dim3(fgk::kInterpThreads), 0, s, a);
        } else {
            hipLaunchKernelGGL(fgk::k_interp_big_poses, dim3(unsigned(npose)), dim3(fgk::kInterpThreads), 0, s, a);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[2], s);
    auto down = [&](void *dst, size_t o, size_t bytes) {
        return dst ? hipMemcpyAsync(dst, b + o, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
    };
    if (e == hipSuccess) e = down(elec, L.elec, size_t(n) * sizeof(float));
    if (e == hipSuccess) e = down(emap, L.emap, size_t(n) * sizeof(float));
    if (e == hipSuccess) e = down(elecTotal, L.et, size_t(npose) * sizeof(double));
    if (e == hipSuccess) e = down(emapTotal, L.mt, size_t(npose) * sizeof(double));
    if (e == hipSuccess) e = down(total, L.tot, size_t(npose) * sizeof(float));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fgc::fail(FG_E_DEVICE, std::string(who) + ": " + hipGetErrorString(e));
    ctx->haveTiming = true;
    ctx->splitTiming = false;
    return FG_OK;
}

}  // namespace

int fg_trilinterp(fg_ctx *ctx, const fg_mapset *ms, const fg_ligand *lig, int npose, const float *xyz, float *elec,
                  float *emap, double *elecTotal, double *emapTotal, float *total)
{
    if (int rc = interpCheck("fg_trilinterp", ctx, ms, lig, npose, xyz)) return rc;
    if (npose == 0) return FG_OK;
    FG_HIP(hipSetDevice(ctx->device));
    const InterpLayout L = interpLayout(int64_t(npose) * lig->natoms, lig->natoms, npose, elec, emap);
    char *b = nullptr;
    FG_HIP(hipMalloc(&b, L.bytes));
    const int rc = interpRun("fg_trilinterp", ctx, ms, lig, npose, xyz, elec, emap, elecTotal, emapTotal, total, b, L);
    (void)hipFree(b);
    return rc;
}

int fg_trilinterp_noalloc(fg_ctx *ctx, const fg_mapset *ms, const fg_ligand *lig, int npose, const float *xyz,
                          float *elec, float *emap, double *elecTotal, double *emapTotal, float *total)
{
    if (int rc = interpCheck("fg_trilinterp_noalloc", ctx, ms, lig, npose, xyz)) return rc;
    if (npose == 0) return FG_OK;
    const InterpLayout L = interpLayout(int64_t(npose) * lig->natoms, lig->natoms, npose, elec, emap);
    if (L.bytes > size_t(FG_NOALLOC_SCRATCH_BYTES))
        return fgc::fail(FG_E_ARG, "fg_trilinterp_noalloc: " + std::to_string(npose) + " poses of " +
                                       std::to_string(lig->natoms) + " atoms need " + std::to_string(L.bytes) +
                                       " bytes of staging, more than its fixed " +
                                       std::to_string(FG_NOALLOC_SCRATCH_BYTES) + "; split the call");
    char *b = static_cast<char *>(scratchOf(ctx->device));
    if (!b) return fgc::fail(FG_E_DEVICE, "fg_trilinterp_noalloc: no staging buffer on this device");
    FG_HIP(hipSetDevice(ctx->device));
    return interpRun("fg_trilinterp_noalloc", ctx, ms, lig, npose, xyz, elec, emap, elecTotal, emapTotal, total, b, L);
}

}  // extern "C"
