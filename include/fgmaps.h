/*
 * fgmaps.h — C ABI of libfgmaps, the MI355X grid-map library.
 *
 * The reference (FastGrid 4.2.1) has no library API: its map backends are
 * C++ functions inside one executable.  This header is the boundary a
 * maintainer binds in their place.  Each entry point names the reference
 * interface it replaces (paths relative to the reference root;
 * FG = fastgrid_master/fastgrid).
 *
 * Conventions
 *  - Plain C types only; every call returns FG_OK (0) or a negative FG_E*
 *    code, and fg_last_error() returns the message of the calling thread's
 *    last failure.  No exception crosses the ABI.
 *  - Host buffers are caller-owned.  Device buffers, streams and graphs are
 *    owned by an fg_ctx (one context per host thread).
 *  - Map outputs are FP64 and NOT rounded: the caller applies the 3-decimal
 *    rounding (fg_round3, FG/autogrid.h:128-131) so CPU and GPU share one
 *    rounding step.  fg_write_outputs() rounds while writing.
 *  - Grid indexing is x fastest: p = (z*ny + y)*nx + x (FG/autogrid.h:140-161).
 *  - There is no CPU fallback: every compute entry point fails with
 *    FG_E_DEVICE when no GPU is usable.
 */
#ifndef FGMAPS_H
#define FGMAPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FG_OK 0
#define FG_E_ARG -1       /* invalid argument or size */
#define FG_E_DEVICE -2    /* HIP/device failure */
#define FG_E_IO -3        /* file could not be read/written */
#define FG_E_INPUT -4     /* GPF/PDBQT semantic error (see fg_prep_exit_code) */
#define FG_E_NOMEM -5

#define FG_TABLE_LEN 8192 /* MAX_DIST, FG/autogrid.h:88 */
#define FG_MAX_MAPS 16    /* MAX_MAPS, fastgrid_master/autodock/autocomm.h:103 */

/* H-bond classes, FG/ParameterLibrary.h:30-38 */
enum { FG_HB_NON = 0, FG_HB_DS = 1, FG_HB_D1 = 2, FG_HB_AS = 3, FG_HB_A1 = 4, FG_HB_A2 = 5 };

/* Computation forms.  COMPENSATED, FP64 and FP32 compute the affinity and
 * desolvation maps in the reference's FP64 arithmetic and differ only in the
 * electrostatic map; PLAIN is FP32 throughout. */
enum {
    FG_PREC_PLAIN = 0,       /* the reference's GPU method in FP32, one job per launch:
                                every lane sums 8 z-points over all atoms in chunks of
                                4000 per launch (FG/electrostatics/Cuda.cpp:75-199,
                                cuda_internal/Kernels.inl:174-220), and the cutoff maps
                                per point in FP32 (FG/CalculateGridmaps.cpp:199-316) */
    FG_PREC_COMPENSATED = 1, /* FP32 far-field pair math, compensated two-float cross-tile
                                sum, FP64 final sum; pairs within 8 A in FP64 */
    FG_PREC_FP64 = 2,        /* FP64 everywhere, the reference's operation order */
    FG_PREC_FP32 = 3         /* the tiled far-field kernel of COMPENSATED over all pairs
                                with a plain FP32 sum (no compensation, no FP64 near part) */
};

/* Dielectric variants of the plain form's e-kernel: the reference's CUDA
 * kernel variants (--cuda-ddd=, FG/ProgramParameters.cpp:63-112;
 * FG/electrostatics/cuda_internal/Kernels.inl:73-119). */
enum {
    FG_PLAIN_DIEL_TABLE = 0,   /* binned table in global memory (--cuda-ddd=g; also =c) */
    FG_PLAIN_DIEL_AUTO = 1,    /* the reference's own choice (FG/electrostatics/Cuda.cpp:114-124):
                                  INPLACE unless an axis of the padded grid exceeds 160 points,
                                  then TEXTURE */
    FG_PLAIN_DIEL_INPLACE = 2, /* analytic calculateDistDepDielInv in FP32 (--cuda-ddd=i) */
    FG_PLAIN_DIEL_TEXTURE = 3  /* table read at the texel of normalized point sampling
                                  (--cuda-ddd=t) */
};

/* What a job run computes. */
#define FG_RUN_ELEC 1u
#define FG_RUN_AFFINITY 2u /* affinity maps + desolvation map */
#define FG_RUN_FLOATING 4u /* the floating grid (jobs created with floating_grid = 1) */
#define FG_RUN_ALL 7u
/* Also time the phases (fg_timing's elec_ms and affinity_ms) of a compensated
 * or FP32 run.  Those runs replay a graph of their launches; with phase
 * timing it holds the timing events as graph nodes, which cost ~20 % when
 * runs follow each other without a wait (L: 28 ms instead of 23 ms per run)
 * and nothing when each run is waited for. */
#define FG_RUN_PHASE_TIMING 8u

typedef struct fg_ctx fg_ctx;
typedef struct fg_job fg_job;
typedef struct fg_prep fg_prep;

/* Grid: numGridPoints, numGridPointsDiv2, gridSpacing (FG/InputData.h:38-43).
 * Point (x,y,z) sits at ((x-div2x)*spacing, ...) relative to the grid centre. */
typedef struct {
    int32_t nx, ny, nz;
    int32_t div2x, div2y, div2z;
    double spacing;
} fg_grid;

/* Receptor atoms, structure of arrays (FG/InputData.h:50-58).  Coordinates
 * are centred on the grid centre; w = charge * coeff_estat [* 332/diel]. */
typedef struct {
    int32_t n;
    const double *x, *y, *z, *w;
    const double *charge, *vol, *solpar;
    const int32_t *type;  /* receptor type index (atomType) */
    const int32_t *hbond; /* FG_HB_* */
} fg_atoms;

/* Bond geometry for directional H-bonds (FG/BondVectors.h:28-42). */
typedef struct {
    const double *rvector;    /* [n][3] */
    const double *rvector2;   /* [n][3] */
    const int32_t *rexp;      /* [n] */
    const uint8_t *disorder;  /* [n] */
} fg_hbond_geom;

/* Lookup tables (FG/PairwiseInteractionEnergies.h:31-49, FG/DesolvExpFunc.h,
 * FG/InputData.h:60).  pair is [n_types][FG_TABLE_LEN][n_maps]. */
typedef struct {
    int32_t n_types, n_maps;
    const double *pair;
    const double *desolv;  /* [FG_TABLE_LEN] */
    const double *eps;     /* [FG_TABLE_LEN] inverse distance-dependent dielectric */
} fg_tables;

/* Per affinity map (FG/GridMap.h:27-55). */
typedef struct {
    int32_t hbond, is_hbonder, is_covalent;
    double solpar_probe, vol_probe;
} fg_map_params;

/* Everything one map set needs: the reference's InputData + GridMapList +
 * tables + bond vectors after precomputation (FG/main.cpp:150-193). */
typedef struct {
    fg_grid grid;
    fg_atoms atoms;
    fg_hbond_geom geom;
    fg_tables tables;
    int32_t n_maps;                     /* affinity maps */
    const fg_map_params *maps;          /* [n_maps] */
    const double *covalent;             /* [n_maps][P] initial values of covalent maps, or NULL */
    double solpar_q;                    /* solparQ */
    int32_t hydrogen;                   /* receptor type index of "HD", or -1 */
    int32_t dist_dep_diel;              /* 1: binned eps table, 0: constant dielectric */
    /* Covalent maps computed on the device (FG/main.cpp:32-52), used when
     * covalent is NULL and cov_on is 1: every covalent map starts from
     * cov_barrier * (1 - exp(ln(1/2) * max(r, 1e-12))) with
     * r = |cov_point - p|^2 * cov_hw_sq_inv (cov_point grid-centred). */
    int32_t cov_on;
    double cov_point[3];
    double cov_hw_sq_inv, cov_barrier;
    /* 1: the job also computes the floating grid (GPF keyword fmap,
     * FG/main.cpp:54-79): float(sqrt(min over atoms of |atom - p|^2)). */
    int32_t floating_grid;
} fg_job_desc;

/* ---------------------------------------------------------------- */
/* Errors                                                           */
/* ---------------------------------------------------------------- */
const char *fg_last_error(void);
double fg_round3(double a); /* roundOutput, FG/autogrid.h:128-131 */

/* ---------------------------------------------------------------- */
/* Device context                                                   */
/* ---------------------------------------------------------------- */
/* Replaces the device set-up of FG/electrostatics/Cuda.cpp:130-136
 * (cudaSetDevice + cudaStreamCreate) and the --cuda-dev choice. */
int fg_create(int device, fg_ctx **out);
void fg_destroy(fg_ctx *ctx);
/* Number of visible GPUs (0 when none); never fails. */
int fg_device_count(void);

/* ---------------------------------------------------------------- */
/* One-shot entry points on host buffers                            */
/* ---------------------------------------------------------------- */
/* Electrostatic map.  Replaces calculateElectrostaticMapAsync +
 * synchronizeCalculation (FG/electrostatics/Electrostatics.h:32-36,
 * FG/electrostatics/Cuda.cpp:75-199 and Cpu.cpp:27-88).  e_out[P] receives the
 * unrounded sum.  eps may be NULL when dist_dep_diel == 0. */
int fg_elec_map(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, const double *eps,
                int dist_dep_diel, int precision, double *e_out);

/* Affinity maps + desolvation map.  Replaces calculateGridmaps
 * (FG/CalculateGridmaps.h:30-31, FG/CalculateGridmaps.cpp:433-485).
 * maps_out[m] (P doubles each) are in/out: covalent maps must hold their
 * initial values (FG/main.cpp:32-52), the others are overwritten.
 * d_out[P] receives the desolvation map.  Outputs are unrounded. */
int fg_affinity_d_maps(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms,
                       const fg_hbond_geom *geom, const fg_tables *tables,
                       const fg_map_params *maps, int n_maps, double solpar_q, int hydrogen,
                       double *const *maps_out, double *d_out);

/* Asynchronous forms.  The reference starts its e-map on a thread of its own
 * and computes the affinity maps meanwhile (FG/main.cpp:215-222):
 * calculateElectrostaticMapAsync returns a handle and synchronizeCalculation
 * waits for it (FG/electrostatics/Electrostatics.h:28-36,
 * FG/electrostatics/Cuda.cpp:222-249).  fg_elec_map_async and
 * fg_affinity_d_maps_async copy their inputs to the device, enqueue the
 * computation on HIP streams of the handle's own (on ctx's device) and return
 * at once, so work enqueued on ctx meanwhile (for instance fg_affinity_d_maps
 * while an e-map runs) overlaps with it on the GPU.  The input buffers may be
 * reused as soon as the call returns; the output buffers are written by
 * fg_wait, which also frees the handle.  fg_wait(NULL) does nothing, as
 * synchronizeCalculation(0) does.  Every handle must be waited for. */
typedef struct fg_handle fg_handle;
int fg_elec_map_async(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, const double *eps,
                      int dist_dep_diel, int precision, double *e_out, fg_handle **out);
int fg_affinity_d_maps_async(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms,
                             const fg_hbond_geom *geom, const fg_tables *tables,
                             const fg_map_params *maps, int n_maps, double solpar_q, int hydrogen,
                             double *const *maps_out, double *d_out, fg_handle **out);
int fg_wait(fg_handle *handle);

/* Initial values of a covalent map.  Replaces initCovalentMaps
 * (FG/main.cpp:32-52): out[p] = barrier * (1 - exp(ln(1/2) * max(r, 1e-12))),
 * r = |cov_point - p|^2 * hw_sq_inv, cov_point grid-centred. */
int fg_covalent_map(fg_ctx *ctx, const fg_grid *grid, const double cov_point[3], double hw_sq_inv,
                    double barrier, double *out);

/* Floating grid.  Replaces calculateFloatingGrid (FG/main.cpp:54-79):
 * out[p] = float(sqrt(min over atoms of |atom - p|^2)); only atoms->n and
 * atoms->x/y/z are read. */
int fg_floating_grid(fg_ctx *ctx, const fg_grid *grid, const fg_atoms *atoms, float *out);

/* ---------------------------------------------------------------- */
/* Device-resident jobs (no reference counterpart: they split the   */
/* one-shot calls into upload / compute / fetch so inputs and maps  */
/* stay in HBM between runs, for benchmarking and batching)         */
/* ---------------------------------------------------------------- */
int fg_job_create(fg_ctx *ctx, const fg_job_desc *desc, fg_job **out);
/* n jobs at once: the host-side inputs (atom records, tiles, cells, tables)
 * are built on up to `threads` host threads (<= 0: one per core, at most 16),
 * then uploaded and sized on ctx's stream.  The n jobs share two device
 * allocations, freed when the last of them is destroyed.  Uploads are staged
 * in a pinned host buffer that ctx keeps (as large as the largest creation's
 * inputs) until fg_destroy.  On an error no job is kept (every out[i] is
 * NULL) and the message names the first failing job.  No reference
 * counterpart (one receptor per process there). */
int fg_job_create_many(fg_ctx *ctx, const fg_job_desc *descs, int n, int threads, fg_job **out);
void fg_job_destroy(fg_job *job);
/* Enqueue the computation on the context's stream; returns immediately.  The
 * electrostatic kernel runs on the context's second stream, concurrently with
 * the cutoff kernels (the reference likewise overlaps its GPU e-map with the
 * CPU affinity maps); the context's stream waits for it, so work enqueued
 * after this call, fg_sync and fg_job_fetch see the finished maps. */
int fg_job_run(fg_ctx *ctx, fg_job *job, unsigned what, int precision);
/* Wait for everything enqueued on the context. */
int fg_sync(fg_ctx *ctx);
/* Copy maps back: out[m] for the n_maps affinity maps, out[n_maps] = e,
 * out[n_maps+1] = d, and out[n_maps+2] = the floating grid when the job has
 * one (P doubles each, the floating grid's values are floats; NULL entries
 * are skipped). */
int fg_job_fetch(fg_ctx *ctx, fg_job *job, double *const *out);
/* fg_job_fetch of many jobs at once: outs[i] is job i's out array.  The
 * copies of all jobs are pipelined through the context's pinned staging
 * buffers (also for fg_job_fetch). */
int fg_job_fetch_many(fg_ctx *ctx, fg_job *const *jobs, int n, double *const *const *outs);
/* The maps as the writers print them, rounded on the device to integer
 * thousandths (roundOutput's floor(1000 v + 0.5), 0 below 0.0005 in
 * magnitude, FG/autogrid.h:128-131): out[m] for m = 0 .. n_maps + 1 (P
 * int32 each, NULL entries skipped), half the bytes of the FP64 maps.  Fails
 * with FG_E_ARG if a value's thousandths do not fit 32 bits (|v| > 2.1e6;
 * fetch such a job with fg_job_fetch).  Every job holds a device buffer for
 * the rounded maps (4 bytes per value).  No reference counterpart. */
int fg_job_fetch_milli(fg_ctx *ctx, fg_job *job, int32_t *const *out);
int fg_job_fetch_milli_many(fg_ctx *ctx, fg_job *const *jobs, int n, int32_t *const *const *outs);
/* Number of (point, atom) pairs with r <= 8 A (computed on the device). */
int fg_job_cutoff_pairs(fg_ctx *ctx, fg_job *job, int64_t *pairs);

/* Checks of a job's last run, computed on the GPU in FP64 from the maps in
 * HBM and the receptor's charges.  min/max: per map (affinity maps in the
 * job's order, then e, then d) the extremes of the rounded values
 * (roundOutput, FG/autogrid.h:128-131) that the writers print and that the
 * reference's log summary reports (GridMapList::logSummary /
 * calculateEnergyMinMax, FG/GridMap.cpp:71-101, 214-224).  total_charge: the
 * receptor's partial charges summed in atom order, the reference's
 * "Total charge" (FG/InputDataLoader.cpp:325, :380); charge_residual its
 * distance to the nearest integer and charge_integral whether that is below
 * 0.005 e (a receptor prepared with whole-residue charges has an integral
 * net charge; the synthetic receptors are neutral).  has_charge is 0 for
 * jobs made without charges. */
typedef struct {
    int32_t n_maps;                  /* affinity maps + 2 */
    double min[FG_MAX_MAPS + 2], max[FG_MAX_MAPS + 2];
    int32_t has_charge, charge_integral;
    double total_charge, charge_residual;
} fg_summary;
int fg_job_summary(fg_ctx *ctx, fg_job *job, fg_summary *out);

/* ---------------------------------------------------------------- */
/* Electrostatic kernel shapes and the autotuner (no reference      */
/* counterpart: the reference picks its CUDA launch shape by fixed  */
/* rules, FG/electrostatics/Cuda.cpp:75-128)                        */
/* ---------------------------------------------------------------- */
/* Shape of a job's electrostatic kernel: grid points per lane along z (4 or
 * 8), receptor atoms per tile (32, 64 or 128), lanes of the 256-lane brick
 * along x (8, 16 or 32; 256: linear items, 256 consecutive lane units taken
 * tile of 8 x 8 lane columns after tile and z-slab after z-slab, which pad
 * only at the end of the grid) and grid points per lane along x (1 or 2).
 * The shape changes only the speed: every shape gives the same bits. */
typedef struct {
    int32_t pz, tile, bx, px;
} fg_shape;
int fg_job_get_shape(const fg_job *job, fg_shape *out);
int fg_job_set_shape(fg_job *job, const fg_shape *shape);

/* Dielectric variant of the job's plain-form e-kernel (default
 * FG_PLAIN_DIEL_TABLE); fg_job_plain_dielectric returns the variant that
 * runs (AUTO resolved; TABLE for a constant dielectric). */
int fg_job_set_plain_dielectric(fg_job *job, int kind);
int fg_job_plain_dielectric(const fg_job *job);

/* Groups the jobs into size classes (atom count and point count octaves,
 * best-padding brick width), times every shape on the median job of each
 * class with HIP events, and gives every job its class's fastest shape.  With
 * more than one job the jobs are taken to run as one batch, which fills the
 * device with all of their bricks: each time is multiplied by the share of
 * the device's workgroup slots its launch kept busy, so shapes compare by
 * throughput rather than by one small launch's latency.  The
 * points per lane along x and z are common to all the jobs (a batch launches
 * one kernel for all of them): the pair with the least total modelled time
 * wins.  A single job's e-kernel runs concurrently with its cutoff kernels:
 * the (at most 4) shapes within 5 % of its fastest e-kernel are then timed
 * over whole steps of all its maps and the fastest step wins.
 * precision: FG_PREC_COMPENSATED or FG_PREC_FP32.  The
 * jobs' maps are undefined until their next run.  fg_tune_report() returns
 * one text line per class of the last call on ctx. */
int fg_autotune(fg_ctx *ctx, fg_job *const *jobs, int n_jobs, int precision);
const char *fg_tune_report(const fg_ctx *ctx);

/* ---------------------------------------------------------------- */
/* Batches: many jobs in one hipGraph of persistent kernels (no     */
/* reference counterpart: the reference runs one receptor per      */
/* process)                                                         */
/* ---------------------------------------------------------------- */
typedef struct fg_batch fg_batch;
#define FG_BATCH_NO_GRAPH 1u /* launch the kernels directly on each run (split timing) */

/* Prepares one run of jobs[0..n_jobs) (created on ctx; they must outlive the
 * batch): the candidate lists of every job (fill and type sort over all
 * (job, brick) items), then one persistent cutoff kernel and
 * one persistent electrostatic kernel, each taking (job, brick) items from a
 * device counter, largest jobs first.  The sequence is captured once into a
 * hipGraph.  precision: FG_PREC_COMPENSATED or FG_PREC_FP32.  Every job
 * keeps its own tile size and brick width; all must share one pz. */
int fg_batch_create(fg_ctx *ctx, fg_job *const *jobs, int n_jobs, unsigned what, int precision, unsigned flags,
                    fg_batch **out);
/* Enqueues one run of the batch on the context's stream; returns at once. */
int fg_batch_launch(fg_ctx *ctx, fg_batch *batch);
void fg_batch_destroy(fg_batch *batch);
/* fg_batch_create + fg_batch_launch + fg_sync + fg_batch_destroy. */
int fg_batch_run(fg_ctx *ctx, fg_job *const *jobs, int n_jobs, unsigned what, int precision);

/* Device time of the kernels of the last fg_job_run / fg_batch_launch on ctx,
 * measured with HIP events (ms).  The electrostatic kernels run on a second
 * stream of the context, concurrently with the cutoff kernels, so elec_ms +
 * affinity_ms exceeds total_ms when they overlap.  A batch replayed from its
 * graph, and a compensated or FP32 job run without FG_RUN_PHASE_TIMING,
 * report total_ms only (elec_ms = affinity_ms = -1). */
typedef struct {
    double elec_ms;      /* electrostatic kernel(s), on the second stream; 0 if none ran */
    double affinity_ms;  /* cutoff kernel(s), incl. the per-brick atom lists */
    double total_ms;     /* first to last event, incl. the final e = near + far sum */
} fg_timing;
int fg_last_timing(fg_ctx *ctx, fg_timing *t);

/* ---------------------------------------------------------------- */
/* Host front end: the reference's parsing, precompute and writers  */
/* ---------------------------------------------------------------- */
typedef struct {
    int32_t v4;                /* --v4: AutoDock 4.0 default parameters (FG/main.cpp:145-147) */
    int32_t ag421_param_order; /* read parameter_file where the GPF names it (AutoGrid 4.2.1) */
} fg_load_opts;

/* InputDataLoader::load + parameter file + tables + bond vectors
 * (FG/main.cpp:145-193, FG/InputDataLoader.cpp:100-797).  Writes the .xyz file
 * as the reference does.  On FG_E_INPUT, fg_prep_exit_code() gives the
 * reference's exit code for the failure. */
int fg_prep_load(const char *gpf_path, const fg_load_opts *opts, fg_prep **out);
int fg_prep_exit_code(void);
/* Prepares n GPFs on up to `threads` host threads (<= 0: one per core, at
 * most 16) and stores out[i] for gpfs[i]; the files are read and written
 * (.xyz) as fg_prep_load does, paths relative to the process's working
 * directory.  On an error no prep is kept (every out[i] is NULL) and the
 * status, message and exit code are those of the first failing GPF in list
 * order.  No reference
 * counterpart: the reference prepares one receptor per process
 * (FG/main.cpp:161-193); this is the host side of batches of pockets. */
int fg_prep_load_many(const char *const *gpf_paths, int n, const fg_load_opts *opts, int threads, fg_prep **out);
void fg_prep_free(fg_prep *prep);
/* Views into the prepared arrays (valid until fg_prep_free). */
int fg_prep_desc(const fg_prep *prep, fg_job_desc *out);
/* Map bookkeeping for writers and callers. */
int fg_prep_num_maps(const fg_prep *prep);                 /* affinity maps */
const char *fg_prep_map_type(const fg_prep *prep, int m);  /* e.g. "OA" */
int fg_prep_num_atoms(const fg_prep *prep);
int64_t fg_prep_num_points(const fg_prep *prep);
/* The GPF echo ("GPF> " lines) the reference writes into its .glg log. */
const char *fg_prep_gpf_echo(const fg_prep *prep);
/* Minimum / maximum of written map m (0..n_maps+1) after fg_write_outputs
 * (GridMapList::logSummary, FG/GridMap.cpp:71-101). */
int fg_prep_summary(const fg_prep *prep, int m, double *min, double *max);

/* .fld + .map files (FG/GridMap.cpp:127-212, FG/Utils.cpp:40-84).  maps[m]
 * for m < n_maps, maps[n_maps] = e, maps[n_maps+1] = d, unrounded, and
 * maps[n_maps+2] = the floating grid when the GPF names one (fmap; printed
 * with "%.3f" as the reference prints its float values).  Also
 * records each map's min/max in the prep for fg_prep_summary. */
int fg_write_outputs(fg_prep *prep, double *const *maps, const char *program_name, int threads);
/* The same files from the rounded maps in integer thousandths
 * (fg_job_fetch_milli): maps[m] for m < n_maps + 2, and the floating grid's
 * values in fmap (NULL when the GPF names none).  Byte-identical output. */
int fg_write_outputs_milli(fg_prep *prep, const int32_t *const *maps, const double *fmap, const char *program_name,
                           int threads);

/* ---------------------------------------------------------------- */
/* The maps' consumer: AutoDock 4.2.1's map set and its trilinear    */
/* interpolation of ligand atoms (AD = autogrid_original_4_2_1_ref/  */
/* autodock in the reference)                                       */
/* ---------------------------------------------------------------- */
/* The geometry of GridMapSetInfo (AD/grid.h:36-54) as readfield leaves it:
 * lo = center - (num_points / 2) * spacing, hi likewise. */
typedef struct {
    double spacing, inv_spacing;
    int32_t num_points[3];   /* NELEMENTS: points per axis minus 1 (even) */
    int32_t num_points1[3];  /* points per axis */
    double lo[3], hi[3], center[3];
} fg_map_grid;

/* readfield (AD/readfield.cc:55-140): #SPACING, #NELEMENTS (even, > 0),
 * #CENTER. */
int fg_read_fld(const char *fld_path, fg_map_grid *out);
/* readmap (AD/readmap.cc:110-300) of one .map file into slot `slot` of a
 * host map set values[((z * ny1 + y) * nx1 + x) * nslots + slot]: SPACING,
 * NELEMENTS and CENTER must match g (compared as floats, as the reference
 * compares them), then one value per line as sscanf "%f" reads it; fails
 * unless every grid point got a value. */
int fg_read_map(const char *map_path, const fg_map_grid *g, int slot, int nslots, float *values);
/* The geometry AutoDock reads back from the .fld fg_write_outputs writes for
 * this prep (its printed spacing and center, parsed as readfield parses
 * them). */
int fg_prep_map_grid(const fg_prep *prep, fg_map_grid *out);

/* A map set in HBM: floats [z][y][x][slot], AutoDock's layout
 * (AD/map_declare.h) without its padding, plus a corner-packed copy for
 * fg_trilinterp (a cell's 8 corner values of each slot side by side, 8x the
 * floats) built when the set is made. */
typedef struct fg_mapset fg_mapset;
int fg_mapset_create(fg_ctx *ctx, const fg_map_grid *g, int nslots, const float *values, fg_mapset **out);
/* Binary hand-off of a job's maps with no text in between: slots 0..n-1 are
 * the affinity maps in the job's order, then e and d (n_maps + 2 slots), and
 * every value is the float AutoDock reads from the printed map (0 below
 * 0.0005 in magnitude, else floor(1000 v + 0.5) / 1000).  g is the grid the
 * maps' .fld describes (fg_prep_map_grid); its points must match the job's. */
int fg_mapset_from_job(fg_ctx *ctx, fg_job *job, const fg_map_grid *g, fg_mapset **out);
int fg_mapset_fetch(fg_ctx *ctx, const fg_mapset *ms, float *values);
int fg_mapset_info(const fg_mapset *ms, fg_map_grid *g, int *nslots);
void fg_mapset_destroy(fg_mapset *ms);

/* One ligand's atoms as trilinterp reads them (AD/trilinterp.cc:46-66). */
typedef struct {
    int32_t natoms;
    const int32_t *type;        /* map slot of each atom's affinity map */
    const float *charge;        /* partial charges (Real) */
    const float *abs_charge;    /* their magnitudes */
    const int32_t *ignore;      /* NULL, or nonzero to skip an atom (ignore_inter) */
    int32_t elec_slot;          /* slot of the e map (the reference's ElecMap) */
    int32_t desolv_slot;        /* slot of the d map (DesolvMap) */
} fg_ligand;

/* trilinterp (AD/trilinterp.cc:87-185) of npose poses of the ligand at once:
 * xyz[pose][atom][3] (floats, host).  Per pose and atom elec/emap as the
 * reference stores them (floats), per pose the double totals (the
 * reference's running sums, in atom order) and the returned value
 * (float)((float)elec_total + emap_total).  Any output may be NULL.  Atoms
 * outside the grid get 500 x their squared distance from the centre (the
 * reference's some_atoms_outside_grid path, taken for every atom here); a
 * corner one past the grid's last point reads 0, as AutoDock's zero-padded
 * array does.  fg_last_timing() reports the device time (total_ms). */
int fg_trilinterp(fg_ctx *ctx, const fg_mapset *ms, const fg_ligand *lig, int npose, const float *xyz, float *elec,
                  float *emap, double *elec_total, double *emap_total, float *total);

/* fg_trilinterp that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight scoring loops of a docking search, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as fg_trilinterp, bit for
 * bit; the device staging of the poses and outputs lives in a fixed array
 * of FG_NOALLOC_SCRATCH_BYTES per GPU that the library holds until the
 * process ends (its address is taken when a context is created on the GPU).
 * A call whose staging (per pose 20 bytes + 12 bytes per atom, plus 8 per
 * atom for each per-atom output requested, and the ligand's arrays) would not
 * fit fails with FG_E_ARG: split the poses.  All calls on one GPU share the
 * array: calls that overlap on one GPU (from several contexts or threads)
 * race and must be serialised by the caller; calls on one context are. */
#define FG_NOALLOC_SCRATCH_BYTES (32u << 20)
int fg_trilinterp_noalloc(fg_ctx *ctx, const fg_mapset *ms, const fg_ligand *lig, int npose, const float *xyz,
                          float *elec, float *emap, double *elec_total, double *emap_total, float *total);

/* ---------------------------------------------------------------- */
/* Synthetic receptors (seeded; deterministic across machines)      */
/* ---------------------------------------------------------------- */
typedef struct {
    uint64_t seed;
    int32_t n_atoms;
    int32_t shape;              /* 0 sphere (radius a), 1 box (a x b x c, centred) */
    double a, b, c;
    double min_sep;             /* minimum pair separation (A), e.g. 1.2 */
    int32_t npts[3];            /* points per axis (odd), e.g. 41 41 41 */
    double spacing;             /* e.g. 0.375 */
    const char *ligand_types;   /* e.g. "A C HD N OA" */
    const char *stem;           /* file stem for gridfld/map names */
} fg_synth_spec;
/* Writes <dir>/<stem>.pdbqt and <dir>/<stem>.gpf (paths inside the GPF are
 * relative to dir). */
int fg_synth_write(const fg_synth_spec *spec, const char *dir);

#ifdef __cplusplus
}
#endif
#endif
