/*
 * fgoracle — FP64 CPU restatement of FastGrid 4.2.1's grid-map computation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (fastgrid_amd/) and the CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links, imports or executes it.
 *
 * Two layers:
 *  1. or_load()/or_precompute()/or_compute()/or_write_outputs(): the whole
 *     fastgrid4 pipeline (GPF + PDBQT parsing, parameter library, tables,
 *     bond vectors, e/d/affinity maps, .map/.fld/.xyz writers) with the
 *     reference's order-dependent semantics.
 *  2. or_elec_map()/or_affinity_maps(): the two hot-path sums on plain arrays,
 *     so tests can feed identical prepared inputs to the oracle and to the
 *     GPU C-ABI.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference root; FG = fastgrid_master/fastgrid).
 *
 * Floating point: compiled with -ffp-contract=off and without fast-math so the
 * FP64 operation order is the reference's on x86-64 (SSE2, no FMA), which is
 * how the reference binary computes.
 */
#ifndef FGORACLE_H
#define FGORACLE_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_DIST 8192          /* FG/autogrid.h:88 */
#define OR_NUM_TYPES 32           /* FG/autogrid.h:93 */
#define OR_MAX_MAPS 16            /* fastgrid_master/autodock/autocomm.h:103 */
#define OR_MAX_ATOM_TYPES 14      /* autocomm.h:102 */
#define OR_MAX_GRID_PTS 1024      /* autocomm.h:94 */
#define OR_NAME 128               /* MAX_CHARS, autocomm.h:76 */

enum { OR_NON = 0, OR_DS, OR_D1, OR_AS, OR_A1, OR_A2 };  /* FG/ParameterLibrary.h:30-38 */

typedef struct {
    char type[8];
    double Rij, epsij, vol, solpar, RijHB, epsijHB;
    int hbond, recIndex, mapIndex, bondIndex;
} or_param;

typedef struct {
    double coeff_vdW, coeff_hbond, coeff_estat, coeff_desolv, coeff_tors;
    int n;                 /* entries in use */
    or_param e[256];
    int slot[65536];       /* 2-character hash -> entry index + 1 (0 = empty) */
} or_lib;

typedef struct {
    char filename[OR_NAME];
    char type[3];
    int atomType, isCovalent, isHBonder, hbond;
    double solparProbe, volProbe, Rij, epsij, RijHB, epsijHB;
    double nbpR[OR_NUM_TYPES], nbpEps[OR_NUM_TYPES];
    int xA[OR_NUM_TYPES], xB[OR_NUM_TYPES], hbonder[OR_NUM_TYPES];
    double *energies;
} or_map;

typedef struct {
    /* files */
    char gpfName[OR_NAME], fld[OR_NAME], xyz[OR_NAME], receptor[OR_NAME];
    char fmap[OR_NAME], paramFile[OR_NAME];
    float *fmapValues;     /* floating grid [P] when fmap is set */
    /* grid */
    int n[3], div2[3], P;
    double spacing, center[3];
    /* receptor */
    char recTypes[OR_NUM_TYPES][3];
    int numRecTypes, N;
    double *x, *y, *z, *w, *charge, *vol, *solpar;
    int *type, *hbond;
    /* bond vectors */
    double *rvec, *rvec2;  /* [N][3] */
    int *rexp;
    unsigned char *disorder;
    /* maps: numMaps = atom maps + e + d */
    int numMaps, numAtomMaps;
    or_map maps[OR_MAX_MAPS + 2];
    /* scalars */
    double eps[OR_MAX_DIST];
    double solparQ, invDielCal, rSmooth, covHalfWidthSqInv, covBarrier, covPoint[3];
    int distDepDiel, disorderH, hydrogen;
    /* tables */
    double *pair;          /* [tableTypes][8192][numAtomMaps] */
    int tableTypes;
    double desolv[OR_MAX_DIST];
    or_lib lib;
    FILE *log;
    char err[512];
} or_input;

/* ---- layer 1: the fastgrid4 pipeline ---- */
/* Returns 0 on success, else the reference's exit code (ExitProgram) and err. */
int or_load(or_input *in, const char *gpf, int v4, int agParamOrder, FILE *log);
int or_precompute(or_input *in);
/* mode bit 0: e-map, bit 1: affinity + d maps, bit 2: floating grid (when the
 * GPF names one).  threads <= 0: OpenMP default. */
void or_compute(or_input *in, int mode, int threads);
int or_write_outputs(or_input *in, const char *programName);
void or_free(or_input *in);

/* ---- layer 2: hot-path sums on plain arrays ---- */
/* FG/electrostatics/Cpu.cpp:27-88.  e[] is accumulated into (the reference
 * starts from the zeroed map) and rounded to 3 dp when round3 != 0. */
void or_elec_map(const int n[3], const int div2[3], double spacing,
                 int N, const double *x, const double *y, const double *z, const double *w,
                 const double *eps, int distDepDiel, double *e, int round3, int threads);

/* FG/CalculateGridmaps.cpp:238-316 for every grid point.  maps[m*P + p] for the
 * numAtomMaps affinity maps, d[p] for desolvation.  Accumulates into the
 * arrays (covalent maps must already hold their values) and rounds when
 * round3 != 0. */
void or_affinity_maps(const int n[3], const int div2[3], double spacing,
                      int N, const double *x, const double *y, const double *z,
                      const double *charge, const double *vol, const double *solpar,
                      const int *type, const int *hbond,
                      const double *rvec, const double *rvec2, const int *rexp,
                      const unsigned char *disorder,
                      int numRecTypes, int numAtomMaps, const double *pair, const double *desolv,
                      const int *mapHbond, const int *mapIsHBonder, const int *mapIsCovalent,
                      const double *mapSolpar, const double *mapVol,
                      double solparQ, int hydrogen, double *maps, double *d, int round3, int threads);

/* FG/main.cpp:54-79: out[p] = float(sqrt(min over atoms of |a - p|^2)),
 * starting from BIG^2 = 1e24 (autodock/autocomm.h:75). */
void or_floating_grid(const int n[3], const int div2[3], double spacing, int N, const double *x, const double *y,
                      const double *z, float *out, int threads);

/* FG/main.cpp:32-52: one covalent map's values. */
void or_covalent_map(const int n[3], const int div2[3], double spacing, const double covPoint[3],
                     double covHalfWidthSqInv, double covBarrier, double *out);

/* Tables (FG/autogrid.h:106-112, FG/PairwiseInteractionEnergies.cpp:59-115,
 * FG/DesolvExpFunc.cpp:27-43). */
void or_ddd_table(double *eps);
void or_desolv_table(double coeffDesolv, double *S);
double or_round3(double a);

/* Counts (point, atom) pairs with r^2 <= 64 (the cutoff pair count). */
long long or_count_cutoff_pairs(const int n[3], const int div2[3], double spacing,
                                int N, const double *x, const double *y, const double *z, int threads);

/* Accessors for the test harness: or_open() = or_load + or_precompute. */
or_input *or_open(const char *gpf, int v4, int agParamOrder, int *rc);
void or_close(or_input *in);
void or_dims(const or_input *in, int out[12]);
void or_scalars(const or_input *in, double out[5]);
const double *or_darray(const or_input *in, const char *name);
const int *or_iarray(const or_input *in, const char *name);
const unsigned char *or_disorder(const or_input *in);
void or_map_params(const or_input *in, int m, int iv[3], double dv[2]);
const char *or_error(const or_input *in);

#ifdef __cplusplus
}
#endif
#endif
