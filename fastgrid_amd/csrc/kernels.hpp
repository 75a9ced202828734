// CDNA4 (gfx950) kernels of libfgmaps.  Included by device.hip only.
//
// Three map kernels:
//  * k_cutoff: affinity maps, desolvation map and the near-field part of the
//    electrostatic map.  One workgroup per brick of 8x8x4 grid points, one
//    point per lane.  Candidate atoms (those within 8 A of the brick) are
//    streamed in a uniform order, so every receptor-atom branch (H-bond
//    class, disorder, per-map probe class) is wave-uniform.  The affinity
//    and desolvation sums are FP64 in the reference's operation order with
//    exactly its table bins, so they match the reference CPU path
//    (FG/CalculateGridmaps.cpp:84-316); the near-field e term takes 1/r from
//    the hardware FP64 reciprocal square root (2^-24 relative).
//  * k_elec: far-field electrostatics in FP32, persistent: workgroups of 256
//    lanes take work items (2D bricks of 8, 16 or 32 lanes along x, or
//    linear items of 256 lane units) from a counter, each lane PX points
//    along x times PZ along z.  Receptor atoms are spatially sorted into
//    tiles (a tuned multiple of kSubTile atoms); partial sums are FP32 over
//    64 atoms, folded into a compensated two-float accumulator, and the final
//    per-point value is formed in FP64.  The inverse dielectric table lives
//    in LDS.  In the compensated form its bins below kNearBins hold 0: those
//    pairs belong to k_cutoff, which takes exactly the pairs the far field's
//    own FP32 bin drops, so the two kernels partition the (point, atom)
//    pairs exactly.
//  * k_elec_fp64: the whole electrostatic sum in the reference's FP64
//    arithmetic (FG/electrostatics/Cpu.cpp:27-88), used as the on-device
//    FP64 oracle for large workloads.
// plus the candidate-list builders (k_brick_lists: count at job creation,
// fill in every run; k_type_sort) and a scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fgk {

constexpr int kTableLen = 8192;
// The device copies of the pair-energy table (and of its derived tables) hold
// bins 0..800 only: every pair the cutoff kernels look up has r <= 8 A.
constexpr int kPairBins = 801;
constexpr int kSubTile = 32;       // atoms per culling box of the electrostatic tiles
constexpr int kCutBX = 8, kCutBY = 8, kCutBZ = 4;  // k_cutoff brick (points)
constexpr int kCutThreads = kCutBX * kCutBY * kCutBZ;
constexpr int kElecThreads = 256;                  // k_elec lanes per brick

enum { HB_NON = 0, HB_DS = 1, HB_D1 = 2, HB_AS = 3, HB_A1 = 4, HB_A2 = 5 };

// One receptor atom as the cutoff kernel reads it (128 bytes).  Derived terms
// are precomputed on the host in FP64 with the reference's operation order.
struct alignas(16) CutAtom {
    double x, y, z, w;   // centred coordinates; w = q * coeff_estat [* 332/diel]
    double qvol;         // solparQ * vol            (FG/CalculateGridmaps.cpp:302)
    double rv[3];        // bond vector
    double rv2[3];       // second bond vector
    double ang[3];       // racc, rdon, Hramp when they do not depend on the point
    // acts: bits 2k..2k+1 = what map k (kernel order) does with this atom's
    // H-bond class (kAct* below), so the pair loop decodes the actions of
    // every map from one uniform word instead of holding them per map
    int32_t acts, hbond, rexp, flags;  // flags: kFlagDisorder | angular kind << 1 | hbond << 3 | type << 8
};
static_assert(sizeof(CutAtom) == 128, "CutAtom must stay 128 bytes");
constexpr int kFlagDisorder = 1;
// Angular kind of an atom (FG/CalculateGridmaps.cpp:84-197):
//   0  no directional term (racc = rdon = Hramp = 1),
//   1  point-independent (unbonded D1/A1/A2: zero bond vectors), values in ang[],
//   2  evaluated per point.
constexpr int kAngNone = 0, kAngConst = 1, kAngPoint = 2;
// Per atom and map, the two desolvation products of
// FG/CalculateGridmaps.cpp:234-235, formed exactly as the reference forms them:
// ab[2m] = solparProbe_m * vol_a, ab[2m+1] = (solpar_a + solparQ*|q_a|) * volProbe_m.

// What map m does with an atom of H-bond class h (FG/CalculateGridmaps.cpp:206-230):
//   kActAdd     e += E
//   kActDonor   e += E' * Hramp * (racc + (1 - racc) * rsph)   (acceptor probe, donor atom)
//   kActMinRacc keep min/max of E * (racc + (1 - racc) * rsph)  (A1 probe, donor atom)
//   kActMinRdon keep min/max of E * (rdon + (1 - rdon) * rsph)  (donor probe, acceptor atom)
constexpr int kActAdd = 0, kActDonor = 1, kActMinRacc = 2, kActMinRdon = 3;

// Per affinity map, uniform across the kernel, in kernel order: maps that keep
// an H-bond min/max come first.
struct MapParam {
    int32_t actions;  // 2 bits per atom H-bond class (6 classes)
    int32_t isCovalent;
    int32_t out;      // index of the output map
    int32_t pad;
    double solProbe, volProbe;  // the probe's desolvation parameters
};

struct GridGeom {
    int32_t nx, ny, nz, dx2, dy2, dz2;
    double spacing;
};

// Coarse cell grid over the atoms, for building per-brick candidate lists.
struct CellGrid {
    int32_t nc[3];
    double origin[3];  // lower corner of cell (0,0,0)
    double size;       // cell edge (A)
};

struct CutoffArgs {
    GridGeom g;
    int32_t nbx, nby, nbz;  // bricks per axis
    int32_t M, hydrogen, doElecNear;
    // 1: the cutoff waves raise their issue priority above the e-kernel's
    // running beside them (their memory waits leave the VALU to it, and when
    // ready they issue first), for cutoff phases with a large share of the work
    int32_t prio;
    int64_t P;
    const CutAtom *atoms;
    const double *vs;                         // [N][2] vol and solpar + solparQ |q| per atom
    const int32_t *brickStart, *brickList;    // all candidates (type-sorted for the typed kernel)
    const int32_t *hStart, *hList;            // DS/D1 candidates
    const int32_t *typeStart;                 // [nb][T+1] runs of one receptor type in brickList
    int32_t T;                                // receptor types (rows of the pair table)
    int32_t addOnlyClasses;                   // bit h: atoms of H-bond class h add E to every map
                                              // and no map is covalent
    int32_t covMask;                          // bit m: map m (kernel order) is a covalent map
    const double *pair;    // [T][kPairBins][M], maps in kernel order
    // the maps the typed single-job kernel does not stage in LDS (from its
    // LDS count on), [T][kPairBins][hiStrideOf(MM)] for the kernel's MM, each
    // row zero-padded: two maps per 16-byte load from half as many cache lines
    // as rows of all maps (unused by the batch kernel)
    const double *pairHi;
    const double *rsph;    // [T][kPairBins][M]: min(max(pair / 100, 0), 1) (FG/CalculateGridmaps.cpp:210)
    const double *desolv;  // [8192]
    const double *eps;     // [8192]
    double *maps;          // [M+2][P]: affinity maps, e (near part), d
    const double *covalent;  // [M][P] or null
    // covalent maps from their parameters (FG/main.cpp:32-52) when covalent is null
    int32_t covOn;
    int32_t N;                // receptor atoms (the floating grid's full scan)
    float *fmap;              // floating grid [P], or null
    double covPoint[3], covHwSqInv, covBarrier;
    MapParam mp[16];
    // e-map pairs (compensated form): r^2 <= e2Lo are the cutoff kernel's
    // for certain, r^2 > e2Hi the far-field kernel's; in between the far
    // field's FP32 bin decides (kNearBins), recomputed here with spacing100
    double e2Lo, e2Hi;
    float spacing100;
};

struct ElecArgs {
    GridGeom g;
    int32_t nAtoms;          // atoms padded to a multiple of kSubTile
    int32_t tableLen;        // entries of eps32
    int32_t tile;            // atoms per tile (a multiple of kSubTile)
    int32_t bxShift;         // lanes along x = 1 << bxShift (8, 16 or 32); 0: linear items (below)
    int32_t nbx, nby, nbz;   // 2D bricks per axis (bxShift > 0)
    // Linear items (bxShift 0): a unit is PX points 8 apart along x (two
    // adjacent ones in the last tile column) times the PZ planes of one
    // z-slab.  A slab's columns form tiles of 8 PX x 8 columns (narrower in
    // the last tile column, shorter in the last tile row); a tile holds 8 x 8
    // units (lane columns x rows), fewer in the last ones, so that one
    // wave's reads cover a compact patch.  Units are
    // numbered slab after slab, tile row after tile row, tile after tile,
    // row after row in a tile; an item is 256 consecutive units, so items
    // cross tiles and slabs and only the last item of a region has padding
    // lanes.  Region 0 holds the `slabs` full slabs (itemsA items), region 1
    // the remaining nz - slabs * PZ planes, if any.
    int32_t ux;              // lane columns per tile row: 8 (tilesX - 1) + lastLanes
    int32_t tilesX;          // tiles per tile row: ceil(nx / (8 PX))
    int32_t lastLanes;       // lane columns of the last tile: its width w (PX 1), ceil(w / 2) (PX 2)
    int32_t units;           // units per slab: ux * ny
    int32_t slabs;           // full slabs: nz / PZ
    int32_t itemsA;          // items of region 0
    int32_t items;           // work items of a launch (bricks or linear items)
    int32_t clampMax;        // last usable table index
    int32_t floorBitsNeg4;   // -4 * bit pattern of 2^23 (mod 2^32): the table's byte offset for biased bins
    int32_t pad;
    int64_t P;
    const float4 *tiles;     // [nAtoms] (x, y, z, w) scaled by 100 (w: 100*w)
    const float *subBox;     // [nAtoms/kSubTile][6] lo xyz, hi xyz (scaled)
    const float *eps32;      // [tableLen]
    double *e;               // [P] electrostatic map (the FP32 form writes it)
    double *far;             // [P] far-field sums of the compensated form, added to the near part
                             // k_cutoff wrote in e by k_add_far
    float lastCut2;          // sub-tiles farther than this (~58.4 A) lie in the table's constant tail
    float tableCut2;         // sub-tiles closer than this at their farthest stay inside the table
    float spacing100;        // 100 * spacing
};

}  // namespace fgk
