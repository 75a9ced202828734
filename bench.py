"""Benchmark of the grid-map hot path on one MI355X.

Workload: the 50,000-atom receptor (random sequential packing in a
70 x 70 x 50 A slab, min separation 1.2 A, protein-like type/charge mix,
seed 3) on a 127^3 grid at 0.375 A with 8 affinity maps (A C HD N NA OA SA S)
plus the electrostatic and desolvation maps.  Inputs are synthetic and
device-resident before timing.  One step = one full map set: per-brick
candidate lists, the cutoff kernel (affinity + desolvation + near-field
electrostatics, FP64) and the far-field electrostatic kernel (FP32 pair math,
compensated two-float cross-tile sums, FP64 final sum).

Metric: Gpair-evals/s = (N*P + P_cut) / t_step, P_cut = (point, atom) pairs
with r <= 8 A.  Also reported: seconds per full map set, the plain FP32
one-job-per-launch form, the per-map error of both forms against the FP64
form, the roofline of the electrostatic kernel, and the FP64 CPU oracle timed
on a bounded sample of the same workload (cpu_baseline).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 vector peak
FLOP_PER_ELEC_PAIR = 15       # algorithmic FLOPs of one (point, atom) Coulomb term
WORKLOAD = dict(tag="L", seed=3, n_atoms=50000, shape="box", dims=(70.0, 70.0, 50.0), npts=(127, 127, 127),
                ligand_types="A C HD N NA OA SA S")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def r3(a):
    return np.where(np.abs(a) < 0.0005, 0.0, np.floor(a * 1000 + 0.5) * 0.001)


def make_inputs(fg, workdir):
    w = WORKLOAD
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
        job.run(fg.RUN_ALL, precision)
        tm = ctx.timing()  # waits for this step's events
        elec.append(tm.elec_ms)
        aff.append(tm.affinity_ms)
    ctx.sync()
    t = (time.perf_counter() - t0) / steps
    return t, float(np.mean(elec)), float(np.mean(aff))


def cpu_baseline(arrays, planes=12):
    """FP64 oracle (OpenMP, all host cores) on the first `planes` z-planes."""
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
    return dict(value=pairs / (t2 - t0) / 1e9, unit="Gpair-evals/s", cores=threads, kind="port",
                sample=f"FP64 oracle, {planes} central z-planes of the 127^3 grid ({a.P} points, all {a.n_atoms} atoms): "
                       f"e-map {t1 - t0:.2f} s, affinity+d maps {t2 - t1:.2f} s",
                seconds_per_map_set_extrapolated=(t2 - t0) * nz0 / planes)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (reference dispatch for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU split is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import fastgrid_amd as fg
    with tempfile.TemporaryDirectory() as wd:
        log("generating", WORKLOAD)
        prep, a = make_inputs(fg, wd)
    N, P, M = a.n_atoms, a.P, a.n_maps
    ctx = fg.Context(0)
    job = fg.Job(ctx, a)
    pcut = job.cutoff_pairs()
    elec_pairs = N * P
    log(f"N={N} P={P} maps={M}+2 N*P={elec_pairs:.4e} P_cut={pcut:.4e}")

    # FP64 form: the reference's arithmetic on the GPU (error reference)
    job.run(fg.RUN_ALL, fg.PREC_FP64)
    ctx.sync()
    ref = job.fetch()
    t64 = ctx.timing()
    log(f"fp64 form: elec {t64.elec_ms:.1f} ms")

    t_plain, elec_plain, aff_plain = time_steps(fg, ctx, job, args.steps, args.warmup, fg.PREC_PLAIN)
    plain = job.fetch()
    log(f"plain: {t_plain * 1e3:.2f} ms/step (elec {elec_plain:.2f}, cutoff {aff_plain:.2f})")
    t_comp, elec_comp, aff_comp = time_steps(fg, ctx, job, args.steps, args.warmup, fg.PREC_COMPENSATED)
    comp = job.fetch()
    log(f"compensated: {t_comp * 1e3:.2f} ms/step (elec {elec_comp:.2f}, cutoff {aff_comp:.2f})")

    def errs(out):
        res = {}
        names = prep.map_types()[:M] + ["e", "d"]
        for m, name in enumerate(names):
            d = np.abs(out[m] - ref[m])
            res[name] = {"max_abs": float(d.max()), "rms": float(np.sqrt((d * d).mean())),
                         "max_abs_rounded": float(np.abs(r3(out[m]) - r3(ref[m])).max())}
        return res

    err_comp, err_plain = errs(comp), errs(plain)
    value = (elec_pairs + pcut) / t_comp / 1e9
    far_pairs = elec_pairs - pcut
    achieved = FLOP_PER_ELEC_PAIR * far_pairs / (elec_comp * 1e-3) / 1e12
    line = {
        "metric": "Gpair-evals/s + s per full map set (127^3 grid, 50k atoms) vs plain GPU & CPU",
        "value": round(value, 2),
        "unit": "Gpair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_comp * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded random-sequential-packing receptor, AD4.1 default parameters)",
        "config": {
            "workload": "L: 50,000-atom 70x70x50 A slab, 127^3 grid @ 0.375 A, 8 affinity maps + e + d",
            "n_atoms": N, "grid": list(a.n), "affinity_maps": M, "precision": "compensated",
            "elec_pairs": elec_pairs, "cutoff_pairs": pcut,
        },
        "seconds_per_map_set": round(t_comp, 5),
        "plain_gpu": {"value": round((elec_pairs + pcut) / t_plain / 1e9, 2), "unit": "Gpair-evals/s",
                      "ms_per_step": round(t_plain * 1e3, 3), "elec_kernel_ms": round(elec_plain, 3)},
        "kernel_ms": {"elec_far": round(elec_comp, 3), "cutoff_and_lists": round(aff_comp, 3),
                      "elec_fp64_form": round(t64.elec_ms, 3)},
        "error_vs_fp64": {"compensated": err_comp, "plain": err_plain},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "k_elec (compensated far field)",
                     "flops_per_launch": FLOP_PER_ELEC_PAIR * far_pairs},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(a)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
