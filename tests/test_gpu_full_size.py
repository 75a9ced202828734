"""The full-size workloads on an MI355X, against the FP64 oracle where it can
follow and through size-independent properties where it cannot.

- L (50,000 atoms, 127^3, 8 maps): every point of the compensated form
  against the GPU FP64 form (within 3e-4 kcal/mol; at this size each
  workgroup of the persistent electrostatic kernel takes several bricks, so a
  dropped or misplaced brick shows), and the first, middle and last z-planes
  against the CPU oracle.
- X (250,000 atoms, 161^3): the same on two sampled planes, and charge
  negation negating e bit for bit.
- B (the 256-pocket batch of the benchmark, seed 7): every pocket's maps from
  the one-graph batch equal its own run bit for bit, and 8 pockets spanning
  the size range match the oracle."""
import copy

import numpy as np
import pytest

from tests import fixtures as fx

pytestmark = pytest.mark.gpu
COMP_TOL = 3e-4


def _workload(tmp, w):
    import fastgrid_amd as fg
    fg.synth_write(tmp, w["tag"], w["seed"], w["n_atoms"], w["shape"], w["dims"], w["npts"], w["ligand_types"])
    return fx.load_arrays(tmp, w["tag"] + ".gpf")


def _check_planes(a, comp, fp64, zs):
    from oracle import oracle as orc
    M = a.n_maps
    for z0 in zs:
        s = fx.z_slab(a, z0, 1)
        oe = orc.elec_map(s)
        om, od = orc.affinity_maps(s)
        pl = lambda v: fx.slab_of(v, a, z0, 1)
        assert np.array_equal(pl(fp64[M]), oe), z0
        assert np.abs(pl(comp[M]) - oe).max() < COMP_TOL, z0
        for m in range(M):
            assert np.array_equal(fx.milli(pl(comp[m])), fx.milli(om[m])), (z0, m)
            assert np.allclose(pl(comp[m]), om[m], rtol=1e-12, atol=1e-9), (z0, m)
        assert np.array_equal(fx.milli(pl(comp[M + 1])), fx.milli(od)), z0


def _full_grid(ctx, a, zs):
    import fastgrid_amd as fg
    job = fg.Job(ctx, a)
    # compensated first, on a fresh job: its far-field buffer was never written
    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    comp = job.fetch()
    job.run(fg.RUN_ALL, fg.PREC_FP64)
    fp64 = job.fetch()
    M = a.n_maps
    err = np.abs(comp[M] - fp64[M])
    print(f"e: max |compensated - FP64| {err.max():.3e}, RMS {np.sqrt((err * err).mean()):.3e}")
    assert err.max() < COMP_TOL
    for m in list(range(M)) + [M + 1]:
        assert np.array_equal(comp[m], fp64[m]), m
    _check_planes(a, comp, fp64, zs)
    return job, comp


def test_L_full_grid_and_sampled_planes(ctx, tmp_path):
    import bench
    import fastgrid_amd as fg
    a = _workload(str(tmp_path), bench.WORKLOAD_L)
    job, comp = _full_grid(ctx, a, (0, 63, 126))
    # untuned, a grid this large runs linear items of 2 x 4 points per lane
    # (about 1.3 items per workgroup at 8 planes would leave it underfilled)
    assert job.shape == (4, 128, 256, 2)
    # other kernel shapes give the same bits at full size: 2D bricks with
    # 16 points per lane, and linear items that cross rows and z-slabs
    for shape in [(8, 64, 16, 2), (8, 128, 256, 2), (4, 32, 256, 1)]:
        job.shape = shape
        job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
        got = job.fetch()
        for m in range(a.n_maps + 2):
            assert np.array_equal(got[m], comp[m]), (shape, m)
    job.close()


def test_X_sampled_planes_and_sign_symmetry(ctx, tmp_path):
    import bench
    import fastgrid_amd as fg
    a = _workload(str(tmp_path), bench.WORKLOAD_X)
    assert a.n_atoms == 250000 and tuple(a.n) == (161, 161, 161)
    job, comp = _full_grid(ctx, a, (0, 97))
    assert job.shape == (8, 128, 256, 2)  # untuned: linear items, 8 planes per lane
    job.close()
    b = copy.copy(a)
    b.w, b.charge = -a.w, -a.charge
    jb = fg.Job(ctx, b)
    jb.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    neg = jb.fetch()
    jb.close()
    M = a.n_maps
    assert np.array_equal(neg[M], -comp[M])
    for m in list(range(M)) + [M + 1]:
        assert np.array_equal(neg[m], comp[m]), m


def test_B_256_pocket_batch(ctx, tmp_path):
    """The benchmark's batch exactly (bench.pocket_specs, seed 7)."""
    import math
    import os
    import bench
    import fastgrid_amd as fg
    from oracle import oracle as orc
    d = str(tmp_path)
    B = bench.BATCH
    gpfs = []
    for i, (seed, atoms, npts) in enumerate(bench.pocket_specs(B["pockets"])):
        radius = (3.0 * atoms / (4.0 * math.pi * B["density"])) ** (1.0 / 3.0)
        fg.synth_write(d, f"p{i}", seed, atoms, "sphere", radius, npts, B["ligand_types"])
        gpfs.append(f"p{i}.gpf")
    cwd = os.getcwd()
    os.chdir(d)
    try:
        preps = fg.prepare_many(gpfs, threads=16)
        arrays = [p.arrays() for p in preps]
    finally:
        os.chdir(cwd)
    assert len(arrays) == 256
    jobs = fg.create_jobs(ctx, preps, threads=16)
    fg.autotune(ctx, jobs)
    want = []
    for j in jobs:
        j.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
        want.append(j.fetch())
    for j in jobs:  # poison the maps: the batch must write every point
        j.run(fg.RUN_ALL, fg.PREC_PLAIN)
    b = fg.Batch(ctx, jobs, fg.RUN_ALL, fg.PREC_COMPENSATED)
    b.launch()
    ctx.sync()
    got = [j.fetch() for j in jobs]
    for i in range(len(jobs)):
        assert all(np.array_equal(x, y) for x, y in zip(got[i], want[i])), i
    b.close()
    # 8 pockets spanning the work range (N * P), within the oracle's reach
    work = np.array([a.n_atoms * a.P for a in arrays])
    order = [int(i) for i in np.argsort(work) if work[i] <= 3e9]
    picks = sorted({order[int(round(k * (len(order) - 1) / 7))] for k in range(8)})
    assert len(picks) == 8
    for i in picks:
        a, out = arrays[i], got[i]
        M = a.n_maps
        om, od = orc.affinity_maps(a)
        for m in range(M):
            assert np.array_equal(fx.milli(out[m]), fx.milli(om[m])), (i, m)
        assert np.array_equal(fx.milli(out[M + 1]), fx.milli(od)), i
        assert np.abs(out[M] - orc.elec_map(a)).max() < COMP_TOL, i
    for j in jobs:
        j.close()
