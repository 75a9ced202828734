"""Benchmark of the grid-map hot path on one MI355X.

Headline workload (L): the 50,000-atom receptor (random sequential packing in
a 70 x 70 x 50 A slab, min separation 1.2 A, protein-like type/charge mix,
seed 3) on a 127^3 grid at 0.375 A with 8 affinity maps (A C HD N NA OA SA S)
plus the electrostatic and desolvation maps.  Inputs are synthetic and
device-resident before timing.  One step = one full map set: per-brick
candidate lists, the cutoff kernel (affinity + desolvation + near-field
electrostatics, FP64) and the far-field electrostatic kernel (FP32 pair math,
compensated two-float cross-tile sums, FP64 final sum), with the kernel shape
picked by the autotuner before timing.

Metric: Gpair-evals/s = (N*P + P_cut) / t_step, P_cut = (point, atom) pairs
with r <= 8 A.  Reported beside it:
  * the plain form (the reference's GPU method in FP32, one job per launch)
    and the tiled FP32 form without compensation, on the same inputs;
  * per-map max / RMS error of every form against the FP64 form (the
    reference's arithmetic on the GPU);
  * the roofline of the electrostatic kernel (VALU-bound) with its HBM
    traffic from the committed rocprofv3 counter summary;
  * end-to-end seconds of the fastgrid4 command line (parse to files);
  * the batch: 256 synthetic pockets (atoms log-uniform in 600-12000, odd
    grid edges uniform in 21-101 per axis, 6 affinity maps) as one hipGraph of
    persistent kernels, against 256 sequential plain launches;
  * AutoDock's consumer of the maps: trilinear interpolation of 2^19
    synthetic 32-atom ligand poses in the L maps handed off in HBM
    (trilinterp), checked bit for bit against the oracle on a sample;
  * the small (1,500 atoms, 41^3) and medium (8,000 atoms, 81^3)
    configurations: every form's per-map error against the CPU oracle on the
    same inputs, and the oracle's time for the whole map set;
  * the stress size (250,000 atoms, 161^3): every form's time and error,
    roofline fraction and HBM bytes of the electrostatic kernel;
  * the FP64 CPU oracle timed on a bounded sample of L (cpu_baseline).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu] [--no-batch] [--no-interp] [--no-stress]
                    [--no-small] [--no-real] [--no-e2e]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 vector peak
PLAIN_VARIANTS = {0: "table (global memory)", 2: "in place (analytic)", 3: "texture (point-sampled table)"}
HBM_PEAK_GBS = 8000.0
FLOP_PER_ELEC_PAIR = 15       # algorithmic FLOPs of one (point, atom) Coulomb term
WORKLOAD_L = dict(tag="L", seed=3, n_atoms=50000, shape="box", dims=(70.0, 70.0, 50.0), npts=(127, 127, 127),
                  ligand_types="A C HD N NA OA SA S")
WORKLOAD_X = dict(tag="X", seed=5, n_atoms=250000, shape="box", dims=(130.0, 130.0, 90.0), npts=(161, 161, 161),
                  ligand_types="A C HD N NA OA SA S")
BATCH = dict(pockets=256, seed=7, atoms=(600, 12000), edges=(21, 101), ligand_types="A C HD N OA SA", density=0.03)
# AutoDock's consumer of the maps: trilinear interpolation of ligand poses
INTERP = dict(poses=1 << 19, atoms=32, seed=11, site=20.0, spread=5.0, check_poses=4096, cpu_poses=1 << 17)
BYTES_PER_INTERP_ATOM = (24 + 3) * 4   # 8 corners x 3 map floats read and xyz read per atom
# HBM bytes per launch of the electrostatic kernel on L, from rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by a device copy of known size
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_elec_{}_hbm_traffic.json")  # per workload tag


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def r3(a):
    return np.where(np.abs(a) < 0.0005, 0.0, np.floor(a * 1000 + 0.5) * 0.001)


def make_inputs(fg, workdir, w):
    fg.synth_write(workdir, w["tag"], w["seed"], w["n_atoms"], w["shape"], w["dims"], w["npts"], w["ligand_types"])
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        prep = fg.Prepared(w["tag"] + ".gpf")
        arrays = prep.arrays()
    finally:
        os.chdir(cwd)
    return prep, arrays


def time_steps(fg, ctx, job, steps, warmup, precision):
    for _ in range(warmup):
        job.run(fg.RUN_ALL, precision)
    ctx.sync()
    elec, aff = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        # with the phases timed (no cost when each step is waited for)
        job.run(fg.RUN_ALL | fg.RUN_PHASE_TIMING, precision)
        tm = ctx.timing()  # waits for this step's events
        elec.append(tm.elec_ms)
        aff.append(tm.affinity_ms)
    ctx.sync()
    t = (time.perf_counter() - t0) / steps
    return t, float(np.mean(elec)), float(np.mean(aff))


def map_errors(names, out, ref):
    res = {}
    for m, name in enumerate(names):
        d = np.abs(out[m] - ref[m])
        res[name] = {"max_abs": float(d.max()), "rms": float(np.sqrt((d * d).mean())),
                     "max_abs_rounded": float(np.abs(r3(out[m]) - r3(ref[m])).max())}
    return res


def cpu_baseline(arrays, planes=12):
    """FP64 oracle (OpenMP, all host cores) on `planes` central z-planes."""
    from oracle import oracle as orc
    import copy
    a = copy.copy(arrays)
    nz0 = arrays.n[2]
    z0 = (nz0 - planes) // 2
    a.n = (arrays.n[0], arrays.n[1], planes)
    a.div2 = (arrays.div2[0], arrays.div2[1], arrays.div2[2] - z0)
    a.covalent = None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t0 = time.perf_counter()
    orc.elec_map(a, threads=threads)
    t1 = time.perf_counter()
    orc.affinity_maps(a, threads=threads)
    t2 = time.perf_counter()
    pcut = orc.cutoff_pairs(a, threads=threads)
    pairs = a.n_atoms * a.P + pcut
    return dict(value=round(pairs / (t2 - t0) / 1e9, 3), unit="Gpair-evals/s", cores=threads, kind="port",
                sample=f"FP64 oracle (the reference's CPU algorithm restated in C, OpenMP), {planes} central z-planes "
                       f"of the 127^3 grid ({a.P} points, all {a.n_atoms} atoms): e-map {t1 - t0:.2f} s, "
                       f"affinity+d maps {t2 - t1:.2f} s",
                seconds_per_map_set_extrapolated=round((t2 - t0) * nz0 / planes, 2))


def end_to_end(fg, workdir, tag):
    """fastgrid4 from GPF to written .map/.fld files (wall clock, second run)."""
    exe = os.path.join(fg.BIN_DIR, "fastgrid4")
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        p = subprocess.run([exe, "-p", tag + ".gpf", "-l", tag + ".glg"], cwd=workdir, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=600)
        times.append(time.perf_counter() - t0)
        if p.returncode != 0:
            raise RuntimeError(p.stderr)
    return round(times[-1], 3)


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of nbytes (HIP's blit kernel), the reference
    dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


# the autotuner's candidate shapes (pz, tile, bx, px): variant (px, pz) x tile
# x brick width (256: linear items)
ELEC_SHAPES = [(pz, tile, bx, px) for px, pz in ((1, 4), (1, 8), (2, 4), (2, 8)) for tile in (32, 64, 128)
               for bx in (8, 16, 32, 256)]


def elec_shapes(fg, shapes, w=WORKLOAD_L):
    """Launches the compensated electrostatic kernel on the workload at each
    listed shape (pz, tile, bx, px), twice, in order, and prints the HIP-event
    times: the dispatches a rocprofv3 counter run of this mode measures."""
    with tempfile.TemporaryDirectory() as wd:
        prep, a = make_inputs(fg, wd, w)
    ctx = fg.Context(0)
    job = fg.Job(ctx, a)
    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    ctx.sync()
    for sh in shapes:
        job.shape = sh
        ms = []
        for _ in range(2):
            job.run(fg.RUN_ELEC | fg.RUN_PHASE_TIMING, fg.PREC_COMPENSATED)
            ctx.sync()
            ms.append(round(ctx.timing().elec_ms, 3))
        print(json.dumps({"shape": list(sh), "elec_ms": ms}), flush=True)
    job.close()
    ctx.close()


def traffic_entry(shape, tag="L"):
    """HBM bytes per electrostatic launch on the workload from the committed
    counter summary, for the kernel shape (pz, tile, bx, px) the run used."""
    try:
        t = json.load(open(TRAFFIC_FILE.format(tag)))
    except (OSError, ValueError):
        return None
    for e in t.get("shapes", []):
        if list(e["shape"]) == list(shape):
            return e
    return None


def pocket_specs(n):
    rng = np.random.default_rng(BATCH["seed"])
    lo, hi = BATCH["atoms"]
    edges = np.arange(BATCH["edges"][0], BATCH["edges"][1] + 1, 2)
    specs = []
    for i in range(n):
        atoms = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        npts = tuple(int(v) for v in rng.choice(edges, 3))
        specs.append((1000 + i, atoms, npts))
    return specs


def batch_graph_only(fg, ctx, n_pockets, launches=3):
    """The 256-pocket batch made, tuned and replayed `launches` times, and
    nothing else (for rocprofv3 counter runs of the batch kernels)."""
    with tempfile.TemporaryDirectory() as wd:
        gpfs = []
        for i, (seed, atoms, npts) in enumerate(pocket_specs(n_pockets)):
            radius = (3.0 * atoms / (4.0 * math.pi * BATCH["density"])) ** (1.0 / 3.0)
            fg.synth_write(wd, f"p{i}", seed, atoms, "sphere", radius, npts, BATCH["ligand_types"])
            gpfs.append(f"p{i}.gpf")
        cwd = os.getcwd()
        os.chdir(wd)
        try:
            preps = fg.prepare_many(gpfs, threads=min(16, os.cpu_count() or 1))
            jobs = fg.create_jobs(ctx, preps, threads=min(16, os.cpu_count() or 1))
        finally:
            os.chdir(cwd)
    fg.autotune(ctx, jobs)
    batch = fg.Batch(ctx, jobs, fg.RUN_ALL, fg.PREC_COMPENSATED)
    ms = []
    for _ in range(launches):
        ctx.sync()
        t0 = time.perf_counter()
        batch.launch()
        ctx.sync()
        ms.append(round((time.perf_counter() - t0) * 1e3, 3))
    print(json.dumps({"batch_ms": ms}), flush=True)
    batch.close()
    for j in jobs:
        j.close()


def run_batch(fg, ctx, steps, warmup, n_pockets):
    """The 256-pocket batch: one hipGraph of persistent kernels against one
    plain launch per pocket (and one compensated launch per pocket), plus the
    end-to-end time from the GPF files to the maps in HBM and on the host."""
    with tempfile.TemporaryDirectory() as wd:
        gpfs = []
        for i, (seed, atoms, npts) in enumerate(pocket_specs(n_pockets)):
            radius = (3.0 * atoms / (4.0 * math.pi * BATCH["density"])) ** (1.0 / 3.0)
            fg.synth_write(wd, f"p{i}", seed, atoms, "sphere", radius, npts, BATCH["ligand_types"])
            gpfs.append(f"p{i}.gpf")
        cwd = os.getcwd()
        os.chdir(wd)
        try:
            t0 = time.perf_counter()
            fg.prepare_many(gpfs, threads=1)
            threads = min(16, os.cpu_count() or 1)
            t_prep1 = time.perf_counter() - t0
            # end to end with the default kernel shapes: parse + precompute on
            # all host cores, upload, one graph of the whole batch, maps fetched
            t0 = time.perf_counter()
            preps = fg.prepare_many(gpfs, threads=threads)
            t1 = time.perf_counter()
            jobs = fg.create_jobs(ctx, preps, threads=threads)
            ctx.sync()
            t2 = time.perf_counter()
            first = fg.Batch(ctx, jobs, fg.RUN_ALL, fg.PREC_COMPENSATED)
            t3 = time.perf_counter()
            first.launch()
            ctx.sync()
            t4 = time.perf_counter()
            # the maps as the files print them: thousandths rounded on the GPU
            fg.fetch_many(ctx, jobs, milli=True)
            t5 = time.perf_counter()
            fg.fetch_many(ctx, jobs)
            t_fp64 = time.perf_counter() - t5
            first.close()
        finally:
            os.chdir(cwd)
    e2e = {"seconds": round(t5 - t0, 3), "prepare_s": round(t1 - t0, 3), "prepare_one_thread_s": round(t_prep1, 3),
           "host_threads": threads, "create_jobs_s": round(t2 - t1, 3), "build_graph_s": round(t3 - t2, 3),
           "first_launch_s": round(t4 - t3, 3), "fetch_maps_s": round(t5 - t4, 3),
           "fetch_maps_fp64_s": round(t_fp64, 3),
           "form": "GPF files -> fg_prep_load_many (host threads) -> fg_job_create_many (host threads) -> "
                   "fg_batch_create (graph) -> one launch -> fg_job_fetch_milli_many (every map as the int32 "
                   "thousandths the files print, rounded on the GPU), batch default kernel shapes (no tuning); "
                   "fetch_maps_fp64_s: the same maps as unrounded FP64 (fg_job_fetch_many), not in seconds"}
    log(f"batch end to end: {e2e}")
    pairs = sum(p.n_atoms * p.P for p in preps)
    pcut = sum(j.cutoff_pairs() for j in jobs)
    log(f"batch: {n_pockets} pockets, sum N*P = {pairs:.4e}, sum P_cut = {pcut:.4e}")
    t0 = time.perf_counter()
    report = fg.autotune(ctx, jobs)
    t_tune = time.perf_counter() - t0
    log(report.splitlines()[0], f"({t_tune:.1f} s)")

    def timed(fn):
        for _ in range(warmup):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) / steps

    batch = fg.Batch(ctx, jobs, fg.RUN_ALL, fg.PREC_COMPENSATED)
    t_batch = timed(batch.launch)
    log(f"batched compensated: {t_batch * 1e3:.2f} ms per batch")

    def seq(prec):
        def go():
            for j in jobs:
                j.run(fg.RUN_ALL, prec)
        return go

    t_plain = timed(seq(fg.PREC_PLAIN))
    log(f"sequential plain: {t_plain * 1e3:.2f} ms")
    # the reference's own dielectric choice per pocket (all in place here:
    # no pocket grid pads past 160 points)
    for j in jobs:
        j.plain_dielectric = fg.PLAIN_DIEL_AUTO
    auto_kinds = sorted({PLAIN_VARIANTS[j.plain_dielectric] for j in jobs})
    t_plain_auto = timed(seq(fg.PREC_PLAIN))
    log(f"sequential plain, reference's dielectric choice: {t_plain_auto * 1e3:.2f} ms")
    t_plain_best = min(t_plain, t_plain_auto)
    # one launch per pocket is latency-bound: tune each size class's median
    # job alone and give its shape to the class, then restore the batch's
    batch_shapes = [j.shape for j in jobs]
    classes = {}
    for j, p in zip(jobs, preps):
        classes.setdefault((p.n_atoms.bit_length(), p.P.bit_length()), []).append((p.n_atoms * p.P, j))
    for members in classes.values():
        members.sort(key=lambda m: m[0])
        rep = members[len(members) // 2][1]
        fg.autotune(ctx, [rep])
        for _, m in members:
            m.shape = rep.shape
    t_seq = timed(seq(fg.PREC_COMPENSATED))
    log(f"sequential compensated: {t_seq * 1e3:.2f} ms")
    for j, sh in zip(jobs, batch_shapes):
        j.shape = sh

    # errors against the FP64 form, per map kind over all pockets
    batch.launch()
    ctx.sync()
    names = None
    worst = {"compensated": {}, "plain": {}}
    for j in jobs:
        comp = j.fetch()
        j.run(fg.RUN_ALL, fg.PREC_FP64)
        ref = j.fetch()
        j.run(fg.RUN_ALL, fg.PREC_PLAIN)  # the reference's dielectric choice
        plain = j.fetch()
        names = BATCH["ligand_types"].split() + ["e", "d"]
        for form, out in (("compensated", comp), ("plain", plain)):
            for name, e in map_errors(names, out, ref).items():
                w = worst[form].setdefault(name, {"max_abs": 0.0, "max_abs_rounded": 0.0})
                w["max_abs"] = max(w["max_abs"], e["max_abs"])
                w["max_abs_rounded"] = max(w["max_abs_rounded"], e["max_abs_rounded"])
    batch.close()
    for j in jobs:
        j.close()
    total = pairs + pcut
    return {
        "workload": f"{n_pockets} synthetic pockets: atoms log-uniform in {BATCH['atoms']}, odd grid edges uniform "
                    f"in {BATCH['edges']} per axis at 0.375 A, sphere of density {BATCH['density']} A^-3, "
                    f"6 affinity maps ({BATCH['ligand_types']}) + e + d, seed {BATCH['seed']}",
        "elec_pairs": pairs, "cutoff_pairs": pcut,
        "batched": {"ms_per_batch": round(t_batch * 1e3, 3), "pockets_per_s": round(n_pockets / t_batch, 1),
                    "gpair_evals_per_s": round(total / t_batch / 1e9, 2),
                    "form": "compensated, one hipGraph of persistent kernels, autotuned shapes"},
        "sequential_plain": {"ms_per_batch": round(t_plain * 1e3, 3), "pockets_per_s": round(n_pockets / t_plain, 1),
                             "gpair_evals_per_s": round(total / t_plain / 1e9, 2),
                             "dielectric": "table (global memory)"},
        "sequential_plain_auto": {"ms_per_batch": round(t_plain_auto * 1e3, 3),
                                  "pockets_per_s": round(n_pockets / t_plain_auto, 1),
                                  "gpair_evals_per_s": round(total / t_plain_auto / 1e9, 2),
                                  "dielectric": auto_kinds},
        "sequential_compensated": {"ms_per_batch": round(t_seq * 1e3, 3),
                                   "pockets_per_s": round(n_pockets / t_seq, 1),
                                   "gpair_evals_per_s": round(total / t_seq / 1e9, 2),
                                   "form": "one launch per pocket, shapes tuned for one launch per size class"},
        "speedup_vs_sequential_plain": round(t_plain_best / t_batch, 2),
        "speedup_vs_sequential_plain_basis": "the faster of sequential_plain and sequential_plain_auto",
        "speedup_vs_sequential_compensated": round(t_seq / t_batch, 2),
        "autotune_s": round(t_tune, 2),
        "end_to_end": e2e,
        "autotune_report": report.splitlines(),
        "max_error_vs_fp64": worst,
    }


def run_interp(fg, ctx, job, prep, steps, warmup, cpu):
    """AutoDock's trilinear interpolation (autodock/trilinterp.cc) of 2^19
    synthetic 32-atom ligand poses in the job's maps, handed off in HBM as
    AutoDock reads them from the printed files (no text).  Poses as a docking
    search makes them: centres uniform in a 20 A box at the grid centre, atoms
    within +-5 A of the centre.  Reports atom evaluations/s of the device
    kernels computing the per-pose totals (a scoring loop's output), a
    bit-for-bit check of every output on a sample against the oracle, and
    the oracle on the host cores as the CPU baseline."""
    from oracle import oracle as orc
    g = prep.map_grid()
    ms = fg.MapSet.from_job(ctx, job, g)
    M = job.arrays.n_maps
    n, A = INTERP["poses"], INTERP["atoms"]
    rng = np.random.default_rng(INTERP["seed"])
    types = rng.integers(0, M, A).astype(np.int32)
    q = rng.uniform(-0.7, 0.7, A).astype(np.float32)
    aq = np.abs(q)
    centre = np.array(list(g.center))
    site, spread = INTERP["site"], INTERP["spread"]
    centres = centre + rng.uniform(-site / 2, site / 2, (n, 1, 3))
    xyz = (centres + rng.uniform(-spread, spread, (n, A, 3))).astype(np.float32)
    for _ in range(warmup):
        fg.trilinterp(ctx, ms, xyz, types, q, aq, M, M + 1, per_atom=False)
    kms, walls = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        fg.trilinterp(ctx, ms, xyz, types, q, aq, M, M + 1, per_atom=False)
        walls.append(time.perf_counter() - t0)
        kms.append(ctx.timing().total_ms)
    r = fg.trilinterp(ctx, ms, xyz[:INTERP["check_poses"]], types, q, aq, M, M + 1)
    k = float(np.mean(kms))
    evals = n * A
    # bit-for-bit check of the first poses against the oracle on the same maps
    c = INTERP["check_poses"]
    og = orc.OiGrid()
    for f in ("spacing", "inv_spacing"):
        setattr(og, f, getattr(g, f))
    for f in ("num_points", "num_points1", "lo", "hi", "center"):
        for i in range(3):
            getattr(og, f)[i] = getattr(g, f)[i]
    maps = ms.fetch()
    oe, om, oet, omt, otot = orc.trilinterp(og, maps, M, M + 1, xyz[:c], q, aq, types)
    exact = bool(np.array_equal(r["elec"][:c], oe) and np.array_equal(r["emap"][:c], om)
                 and np.array_equal(r["elec_total"][:c], oet) and np.array_equal(r["emap_total"][:c], omt)
                 and np.array_equal(r["total"][:c], otot))
    out = {"workload": f"{n} poses x {A} atoms (random types of the {M} affinity maps, charges in +-0.7), pose "
                       f"centres uniform in a {site:g} A box at the grid centre, atoms within +-{spread:g} A, in "
                       f"the L map set ({M} affinity maps + e + d, {g.num_points1[0]}x{g.num_points1[1]}x"
                       f"{g.num_points1[2]}) handed off from the job in HBM; per-pose totals",
           "atom_evals_per_launch": evals, "kernel_ms": round(k, 4),
           "gevals_per_s": round(evals / (k * 1e-3) / 1e9, 2),
           "algorithmic_gbs": round(evals * BYTES_PER_INTERP_ATOM / (k * 1e-3) / 1e9, 1),
           "seconds_per_call_with_pcie": round(float(np.mean(walls)), 4),
           "bit_exact_vs_oracle_poses": c if exact else 0}
    log(f"trilinterp: {out}")
    if not exact:
        raise SystemExit("trilinterp differs from the oracle")
    if cpu:
        s = INTERP["cpu_poses"]
        threads = os.cpu_count() or 1
        t0 = time.perf_counter()
        orc.trilinterp(og, maps, M, M + 1, xyz[:s], q, aq, types, threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(s * A / dt / 1e9, 4), "unit": "Gatom-evals/s", "cores": threads,
                               "kind": "port", "sample": f"oracle restatement of trilinterp, first {s} poses"}
    ms.close()
    return out


# A real receptor: half of collagen 1T60 (25,437 atoms, every polar H and O
# bonded), the fixture of the reference's own tests, on a 127^3 grid
REAL = dict(pdbqt=os.path.join(ROOT, "tests", "golden", "colagen2", "macromol_1t60_half_prot.pdbqt.gz"),
            npts=126, spacing=0.375, center=(64.0, -16.0, 37.0), receptor_types="A C HD N OA SA",
            ligand_types="A C HD N NA OA SA S")


# The small and medium configurations: the same generator, 1,500 atoms in a
# 24 A sphere on 41^3 (5 maps) and 8,000 atoms in a 40 A sphere on 81^3
# (8 maps)
WORKLOAD_S = dict(tag="S", seed=1, n_atoms=1500, shape="sphere", dims=24.0, npts=(41, 41, 41),
                  ligand_types="A C HD N OA")
WORKLOAD_M = dict(tag="M", seed=2, n_atoms=8000, shape="sphere", dims=40.0, npts=(81, 81, 81),
                  ligand_types="A C HD N NA OA SA S")


def run_small_medium(fg, ctx, steps, warmup, cpu=True):
    """S and M: every GPU form's per-map max |error| against the CPU oracle's
    FP64 maps of the same input files (the reference's CPU algorithm), the
    GPU times per map set, and the oracle's own time for the whole map set on
    the host cores."""
    from oracle import oracle as orc
    out = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    for w in (WORKLOAD_S, WORKLOAD_M):
        with tempfile.TemporaryDirectory() as wd:
            prep, a = make_inputs(fg, wd, w)
        names = prep.map_types()[:a.n_maps] + ["e", "d"]
        t0 = time.perf_counter()
        oe = orc.elec_map(a, threads=threads)
        om, od = orc.affinity_maps(a, threads=threads)
        t_cpu = time.perf_counter() - t0
        ref = list(om) + [oe, od]
        job = fg.Job(ctx, a)
        pcut = job.cutoff_pairs()
        res = {"workload": f"{w['tag']}: {a.n_atoms} atoms in a {w['dims']:g} A sphere, {w['npts'][0]}^3 grid @ "
                           f"0.375 A, {a.n_maps} affinity maps + e + d, seed {w['seed']}",
               "elec_pairs": a.n_atoms * a.P, "cutoff_pairs": pcut}
        if cpu:
            res["cpu_oracle"] = {"seconds_per_map_set": round(t_cpu, 3), "cores": threads, "kind": "port",
                                 "gpair_evals_per_s": round((a.n_atoms * a.P + pcut) / t_cpu / 1e9, 3)}
        job.plain_dielectric = fg.PLAIN_DIEL_AUTO
        auto = PLAIN_VARIANTS[job.plain_dielectric]
        for key, prec, diel in (("compensated", fg.PREC_COMPENSATED, None), ("plain", fg.PREC_PLAIN, fg.PLAIN_DIEL_TABLE),
                                ("plain_auto", fg.PREC_PLAIN, fg.PLAIN_DIEL_AUTO), ("fp64", fg.PREC_FP64, None)):
            if diel is not None:
                job.plain_dielectric = diel
            t, _, _ = time_steps(fg, ctx, job, steps, warmup, prec)
            err = map_errors(names, job.fetch(), ref)
            res[key] = {"ms_per_step": round(t * 1e3, 3),
                        "max_abs_error_vs_cpu_oracle": {n: e["max_abs"] for n, e in err.items()},
                        "max_abs_rounded_error_vs_cpu_oracle": {n: e["max_abs_rounded"] for n, e in err.items()}}
            if key == "plain_auto":
                res[key]["dielectric"] = auto
        job.close()
        log(f"{w['tag']}: " + ", ".join(f"{k} {res[k]['ms_per_step']} ms" for k in ("compensated", "plain", "plain_auto",
                                                                                      "fp64")))
        out[w["tag"]] = res
    return out


def run_real(fg, ctx, steps, warmup):
    """The real receptor: FP64 form as the error reference, autotuned
    compensated and plain forms timed."""
    import gzip
    w = REAL
    with tempfile.TemporaryDirectory() as wd:
        with gzip.open(w["pdbqt"], "rb") as g:
            open(os.path.join(wd, "rec.pdbqt"), "wb").write(g.read())
        lig = w["ligand_types"].split()
        n = w["npts"]
        gpf = [f"npts {n} {n} {n}", f"gridfld rec.maps.fld", f"spacing {w['spacing']}",
               f"receptor_types {w['receptor_types']}", f"ligand_types {w['ligand_types']}",
               "receptor rec.pdbqt", "gridcenter {:.3f} {:.3f} {:.3f}".format(*w["center"]), "smooth 0.5"]
        gpf += [f"map rec.{t}.map" for t in lig] + ["elecmap rec.e.map", "dsolvmap rec.d.map",
                                                    "dielectric -0.1465"]
        open(os.path.join(wd, "rec.gpf"), "w").write("\n".join(gpf) + "\n")
        cwd = os.getcwd()
        os.chdir(wd)
        try:
            prep = fg.Prepared("rec.gpf")
            a = prep.arrays()
            names = prep.map_types()[:a.n_maps] + ["e", "d"]
        finally:
            os.chdir(cwd)
    rv, rv2 = a.rvector.reshape(-1, 3), a.rvector2.reshape(-1, 3)
    bonded = ~((rv == 0).all(1) & (rv2 == 0).all(1)) & np.isin(a.hbond, (2, 4, 5))
    job = fg.Job(ctx, a)
    pcut = job.cutoff_pairs()
    tot = a.n_atoms * a.P + pcut
    job.run(fg.RUN_ALL, fg.PREC_FP64)
    ctx.sync()
    ref = job.fetch()
    fg.autotune(ctx, [job])
    res = {"workload": f"half of collagen 1T60 ({a.n_atoms} atoms, {int(bonded.sum())} bonded H-bond donors/"
                       f"acceptors), {n + 1}^3 grid @ {w['spacing']} A, {a.n_maps} affinity maps + e + d",
           "cutoff_pairs": pcut, "shape": job.shape}
    for key, prec in (("compensated", fg.PREC_COMPENSATED), ("plain", fg.PREC_PLAIN)):
        t, el, af = time_steps(fg, ctx, job, steps, warmup, prec)
        err = map_errors(names, job.fetch(), ref)
        res[key] = {"value": round(tot / t / 1e9, 2), "ms_per_step": round(t * 1e3, 3), "elec_ms": round(el, 3),
                    "cutoff_ms": round(af, 3), "e_max_abs_vs_fp64": err["e"]["max_abs"],
                    "affinity_d_max_abs_vs_fp64": max(v["max_abs"] for k, v in err.items() if k != "e")}
        log(f"real receptor {key}: {t * 1e3:.2f} ms/step (elec {el:.2f}, cutoff+lists {af:.2f})")
    job.close()
    return res


def run_workload(fg, ctx, w, steps, warmup, tune=True, interp=False, cpu=True, shape=None, forms=None):
    """One large job: FP64 form (error reference), plain, FP32 and
    compensated forms.  Returns a dict of results and the job arrays."""
    with tempfile.TemporaryDirectory() as wd:
        log("generating", w)
        prep, a = make_inputs(fg, wd, w)
    N, P, M = a.n_atoms, a.P, a.n_maps
    t0 = time.perf_counter()
    job = fg.Job(ctx, a)
    t_create = time.perf_counter() - t0
    pcut = job.cutoff_pairs()
    elec_pairs = N * P
    log(f"N={N} P={P} maps={M}+2 N*P={elec_pairs:.4e} P_cut={pcut:.4e}")
    job.run(fg.RUN_ALL, fg.PREC_FP64)
    ctx.sync()
    ref = job.fetch()
    t64 = ctx.timing()
    log(f"fp64 form: elec {t64.elec_ms:.1f} ms")
    if shape:
        job.shape = shape
    report = fg.autotune(ctx, [job]) if tune and not shape else ""
    shape = job.shape
    log("tuned shape (pz, tile, bx, px):", shape)
    for line in report.splitlines()[1:]:
        log("  " + line)
    names = prep.map_types()[:M] + ["e", "d"]
    res = {"N": N, "P": P, "M": M, "elec_pairs": elec_pairs, "cutoff_pairs": pcut, "fp64_elec_ms": t64.elec_ms,
           "shape": shape, "tune_report": report, "job_create_s": t_create}
    # the plain form twice: with the dielectric table in global memory, and
    # with the dielectric variant the reference's CUDA path picks for this
    # grid (in place up to 160 padded points per axis, texture beyond)
    job.plain_dielectric = fg.PLAIN_DIEL_AUTO
    res["plain_auto_variant"] = PLAIN_VARIANTS[job.plain_dielectric]
    for key, prec, diel in (("plain", fg.PREC_PLAIN, fg.PLAIN_DIEL_TABLE), ("plain_auto", fg.PREC_PLAIN, fg.PLAIN_DIEL_AUTO),
                            ("fp32", fg.PREC_FP32, None), ("compensated", fg.PREC_COMPENSATED, None)):
        if forms and key not in forms:
            continue
        if diel is not None:
            job.plain_dielectric = diel
        t, el, af = time_steps(fg, ctx, job, steps, warmup, prec)
        res[key] = {"t": t, "elec_ms": el, "cutoff_ms": af, "errors": map_errors(names, job.fetch(), ref)}
        log(f"{key}: {t * 1e3:.2f} ms/step (elec {el:.2f}, cutoff phase {af:.2f})")
    # the cutoff kernel alone (affinity and desolvation maps, no electrostatic
    # work beside it; the candidate lists are made when the job is created)
    for _ in range(warmup):
        job.run(fg.RUN_AFFINITY, fg.PREC_COMPENSATED)
    ms = []
    for _ in range(steps):
        job.run(fg.RUN_AFFINITY | fg.RUN_PHASE_TIMING, fg.PREC_COMPENSATED)
        ms.append(ctx.timing().affinity_ms)
    res["cutoff_alone_ms"] = round(float(np.mean(ms)), 3)
    log(f"cutoff kernel alone: {res['cutoff_alone_ms']} ms")
    if interp:
        res["trilinterp"] = run_interp(fg, ctx, job, prep, steps, warmup, cpu)
    job.close()
    return res, a


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-pocket batch")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end command-line timing")
    ap.add_argument("--no-interp", action="store_true", help="skip the trilinear-interpolation line")
    ap.add_argument("--no-real", action="store_true", help="skip the real-receptor line")
    ap.add_argument("--pockets", type=int, default=BATCH["pockets"])
    ap.add_argument("--no-stress", action="store_true", help="skip the 250k-atom / 161^3 stress size")
    ap.add_argument("--no-small", action="store_true", help="skip the small and medium configurations")
    ap.add_argument("--elec-shapes", default="",
                    help="only launch the electrostatic kernel on L at these shapes, e.g. '4,32,16,1;4,32,8,2' "
                         "(pz,tile,bx,px), or 'all' for the autotuner's 48 shapes (for rocprofv3 counter runs)")
    ap.add_argument("--elec-workload", default="L", choices=["L", "X"],
                    help="workload of --elec-shapes: L (50k atoms, 127^3) or X (the 250k-atom / 161^3 stress size)")
    ap.add_argument("--batch-graph-only", action="store_true",
                    help="only make, tune and replay the 256-pocket batch three times (for rocprofv3 counter runs)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the 256-pocket batch and print its results (for rocprofv3 counter runs)")
    ap.add_argument("--shape", default="",
                    help="electrostatic kernel shape 'pz,tile,bx,px' for L instead of the autotuner's pick, with only "
                         "the compensated form timed and nothing else run (for rocprofv3 counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (reference dispatch for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.elec_shapes:
        import fastgrid_amd as fg
        elec_shapes(fg, ELEC_SHAPES if args.elec_shapes == "all" else
                    [tuple(int(v) for v in sh.split(",")) for sh in args.elec_shapes.split(";")],
                    WORKLOAD_X if args.elec_workload == "X" else WORKLOAD_L)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the grid would be split into z-slabs, one per GPU)", file=sys.stderr)
        sys.exit(2)

    import fastgrid_amd as fg
    ctx = fg.Context(0)
    if args.batch_graph_only:
        batch_graph_only(fg, ctx, args.pockets)
        return
    if args.batch_only:
        print(json.dumps({"batch": run_batch(fg, ctx, max(1, args.steps // 2), 1, args.pockets)}), flush=True)
        return
    if args.shape:
        sh = tuple(int(v) for v in args.shape.split(","))
        L, _ = run_workload(fg, ctx, WORKLOAD_L, args.steps, args.warmup, shape=sh, forms=("compensated",))
        c = L["compensated"]
        print(json.dumps({"workload": "L", "shape": sh, "ms_per_step": round(c["t"] * 1e3, 3),
                          "elec_ms": round(c["elec_ms"], 3), "cutoff_ms": round(c["cutoff_ms"], 3),
                          "cutoff_alone_ms": L.get("cutoff_alone_ms")}), flush=True)
        return
    L, a = run_workload(fg, ctx, WORKLOAD_L, args.steps, args.warmup, interp=not args.no_interp, cpu=not args.no_cpu)
    comp, plain, plain_auto, fp32 = L["compensated"], L["plain"], L["plain_auto"], L["fp32"]
    fastest_plain = min(plain["t"], plain_auto["t"])
    tot_pairs = L["elec_pairs"] + L["cutoff_pairs"]
    value = tot_pairs / comp["t"] / 1e9
    far_pairs = L["elec_pairs"] - L["cutoff_pairs"]
    achieved = FLOP_PER_ELEC_PAIR * far_pairs / (comp["elec_ms"] * 1e-3) / 1e12
    tr = traffic_entry(L["shape"])
    line = {
        "metric": "Gpair-evals/s + s per full map set (127^3 grid, 50k atoms) vs plain GPU & CPU",
        "value": round(value, 2),
        "unit": "Gpair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(comp["t"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded random-sequential-packing receptor, AD4.1 default parameters)",
        "config": {
            "workload": "L: 50,000-atom 70x70x50 A slab, 127^3 grid @ 0.375 A, 8 affinity maps + e + d",
            "n_atoms": L["N"], "grid": list(WORKLOAD_L["npts"]), "affinity_maps": L["M"],
            "precision": "compensated (FP32 far-field pairs, two-float cross-tile sums, FP64 final sum; "
                         "pairs within 8 A and the affinity/desolvation maps in FP64)",
            "elec_pairs": L["elec_pairs"], "cutoff_pairs": L["cutoff_pairs"],
            "elec_kernel_shape": {"pz": L["shape"][0], "tile": L["shape"][1], "bx": L["shape"][2],
                                  "px": L["shape"][3]},
            "autotune": L["tune_report"].strip().splitlines(),
        },
        "seconds_per_map_set": round(comp["t"], 5),
        "plain_gpu": {"value": round(tot_pairs / plain["t"] / 1e9, 2), "unit": "Gpair-evals/s",
                      "ms_per_step": round(plain["t"] * 1e3, 3), "elec_ms": round(plain["elec_ms"], 3),
                      "cutoff_ms": round(plain["cutoff_ms"], 3),
                      "form": "the reference's GPU method in FP32 (8 z-points per lane, 4000 atoms per launch, "
                              "dielectric table in global memory, --cuda-ddd=g) and a per-point FP32 cutoff kernel"},
        "plain_gpu_auto": {"value": round(tot_pairs / plain_auto["t"] / 1e9, 2), "unit": "Gpair-evals/s",
                           "ms_per_step": round(plain_auto["t"] * 1e3, 3), "elec_ms": round(plain_auto["elec_ms"], 3),
                           "cutoff_ms": round(plain_auto["cutoff_ms"], 3),
                           "dielectric": L["plain_auto_variant"],
                           "form": "the same with the dielectric variant the reference's CUDA path picks for this "
                                   "grid (FG/electrostatics/Cuda.cpp:114-124)"},
        "speedup_vs_plain_gpu": round(fastest_plain / comp["t"], 2),
        "speedup_vs_plain_gpu_basis": "the faster of plain_gpu and plain_gpu_auto",
        "fp32_tiled": {"value": round(tot_pairs / fp32["t"] / 1e9, 2), "ms_per_step": round(fp32["t"] * 1e3, 3),
                       "elec_ms": round(fp32["elec_ms"], 3)},
        "kernel_ms": {"elec_far": round(comp["elec_ms"], 3), "cutoff_phase": round(comp["cutoff_ms"], 3),
                      "cutoff_alone": L["cutoff_alone_ms"],
                      "elec_fp64_form": round(L["fp64_elec_ms"], 3)},
        "error_vs_fp64": {"compensated": comp["errors"], "fp32_tiled": fp32["errors"], "plain": plain["errors"],
                          "plain_auto": plain_auto["errors"]},
        "job_create_s": round(L["job_create_s"], 4),
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": tr["bytes_per_launch"] if tr else None,
                     "kernel": "k_elec (compensated far field)",
                     "flops_per_launch": FLOP_PER_ELEC_PAIR * far_pairs,
                     "hbm_gbs": round(tr["bytes_per_launch"] / (comp["elec_ms"] * 1e-3) / 1e9, 1) if tr else None,
                     "traffic_source": os.path.relpath(TRAFFIC_FILE.format("L"), ROOT) if tr else None},
    }
    if "trilinterp" in L:
        line["trilinterp"] = L["trilinterp"]
    if not args.no_e2e:
        with tempfile.TemporaryDirectory() as wd:
            w = WORKLOAD_L
            fg.synth_write(wd, w["tag"], w["seed"], w["n_atoms"], w["shape"], w["dims"], w["npts"], w["ligand_types"])
            line["seconds_per_map_set_end_to_end"] = end_to_end(fg, wd, w["tag"])
            log("end to end:", line["seconds_per_map_set_end_to_end"], "s")
    if not args.no_real:
        line["real_receptor"] = run_real(fg, ctx, args.steps, args.warmup)
    if not args.no_batch:
        line["batch"] = run_batch(fg, ctx, max(1, args.steps // 2), 1, args.pockets)
    if not args.no_small:
        line["small_medium"] = run_small_medium(fg, ctx, args.steps, args.warmup, cpu=not args.no_cpu)
    if not args.no_stress:
        X, _ = run_workload(fg, ctx, WORKLOAD_X, max(1, args.steps // 5), 1)
        xt = X["elec_pairs"] + X["cutoff_pairs"]
        line["stress"] = {
            "workload": "X: 250,000-atom 130x130x90 A box, 161^3 grid @ 0.375 A, 8 affinity maps + e + d",
            "compensated": {"value": round(xt / X["compensated"]["t"] / 1e9, 2),
                            "ms_per_step": round(X["compensated"]["t"] * 1e3, 2),
                            "elec_ms": round(X["compensated"]["elec_ms"], 2),
                            "roofline_frac": round(FLOP_PER_ELEC_PAIR * (X["elec_pairs"] - X["cutoff_pairs"]) /
                                                   (X["compensated"]["elec_ms"] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)},
            "plain": {"value": round(xt / X["plain"]["t"] / 1e9, 2), "ms_per_step": round(X["plain"]["t"] * 1e3, 2)},
            "plain_auto": {"value": round(xt / X["plain_auto"]["t"] / 1e9, 2),
                           "ms_per_step": round(X["plain_auto"]["t"] * 1e3, 2), "dielectric": X["plain_auto_variant"]},
            "speedup_vs_plain_gpu": round(min(X["plain"]["t"], X["plain_auto"]["t"]) / X["compensated"]["t"], 2),
            "job_create_s": round(X["job_create_s"], 3),
            "fp32_tiled": {"value": round(xt / X["fp32"]["t"] / 1e9, 2)},
            "error_vs_fp64": {k: X[k]["errors"] for k in ("compensated", "fp32", "plain", "plain_auto")},
            "shape": X["shape"],
        }
        xt_tr = traffic_entry(X["shape"], "X")
        if xt_tr:
            line["stress"]["hbm"] = {
                "bytes_per_launch": xt_tr["bytes_per_launch"],
                "gbs": round(xt_tr["bytes_per_launch"] / (X["compensated"]["elec_ms"] * 1e-3) / 1e9, 1),
                "source": os.path.relpath(TRAFFIC_FILE.format("X"), ROOT)}
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(a)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
