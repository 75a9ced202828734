"""Batched form and autotuner on an MI355X.

A batch runs many jobs as one hipGraph of persistent kernels.  Each grid
point's arithmetic depends neither on the kernel shape (tile size, brick,
points per lane) nor on the order in which workgroups take bricks, so a batch
must reproduce every job's own run bit for bit; against the FP64 oracle the
compensated bound of the one-job tests applies.  Jobs made together on host
threads (fg_job_create_many) equal jobs made one by one."""
import numpy as np
import pytest

from tests import fixtures as fx

pytestmark = pytest.mark.gpu

# (seed, atoms, points per axis, ligand types): small to large, odd and
# lopsided grids, one job with 8 maps among 6-map jobs
POCKETS = [
    (11, 600, (21, 23, 25), "A C HD N OA SA"),
    (12, 1500, (41, 21, 61), "A C HD N OA SA"),
    (13, 3000, (33, 55, 21), "A C HD N NA OA SA S"),
    (14, 5000, (61, 61, 61), "A C HD N OA SA"),
    (15, 800, (101, 21, 21), "A C HD N OA SA"),
    (16, 9000, (45, 47, 43), "A C HD N OA SA"),
]


@pytest.fixture(scope="module")
def pockets(tmp_path_factory):
    import dataclasses
    tmp = str(tmp_path_factory.mktemp("pockets"))
    arrays = [fx.pocket(tmp, f"p{i}", seed, n, npts, lig) for i, (seed, n, npts, lig) in enumerate(POCKETS)]
    # two pockets also compute their floating grid
    return [dataclasses.replace(a, floating_grid=1) if i in (1, 3) else a for i, a in enumerate(arrays)]


def _jobs(ctx, arrays):
    import fastgrid_amd as fg
    return [fg.Job(ctx, a) for a in arrays]


def _single(jobs, prec, what=None):
    import fastgrid_amd as fg
    out = []
    for j in jobs:
        j.run(fg.RUN_ALL if what is None else what, prec)
        out.append(j.fetch())
    return out


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("graph", [True, False])
def test_batch_reproduces_single_runs(ctx, pockets, graph):
    import fastgrid_amd as fg
    jobs = _jobs(ctx, pockets)
    for prec in (fg.PREC_COMPENSATED, fg.PREC_FP32):
        want = _single(jobs, prec)
        for j in jobs:  # poison the maps: the batch must write every point
            j.run(fg.RUN_ALL, fg.PREC_PLAIN)
        b = fg.Batch(ctx, jobs, fg.RUN_ALL, prec, graph=graph)
        b.launch()
        ctx.sync()
        for i, j in enumerate(jobs):
            assert _same(j.fetch(), want[i]), (prec, i)
        b.close()
    for j in jobs:
        j.close()


@pytest.mark.parametrize("what", [1, 2])
def test_batch_partial_runs(ctx, pockets, what):
    """Electrostatic map only (with its FP64 near part) / affinity maps only."""
    import fastgrid_amd as fg
    jobs = _jobs(ctx, pockets[:4])
    want = _single(jobs, fg.PREC_COMPENSATED, what)
    b = fg.Batch(ctx, jobs, what, fg.PREC_COMPENSATED)
    b.launch()
    ctx.sync()
    for i, j in enumerate(jobs):
        got = j.fetch()
        M = pockets[i].n_maps
        idx = [M] if what == fg.RUN_ELEC else list(range(M)) + [M + 1]
        for m in idx:
            assert np.array_equal(got[m], want[i][m]), (i, m)
    b.close()


def test_graph_replays_agree(ctx, pockets):
    """Replays of one graph, with other work and synchronisations between
    them, give the eager result every time (the work counters are cleared by
    a kernel inside the graph)."""
    import fastgrid_amd as fg
    jobs = _jobs(ctx, pockets[:3])
    other = fg.Job(ctx, pockets[3])
    want = _single(jobs, fg.PREC_COMPENSATED)
    b = fg.Batch(ctx, jobs)
    for rep in range(4):
        b.launch()
        ctx.sync()
        for i, j in enumerate(jobs):
            assert _same(j.fetch(), want[i]), (rep, i)
        other.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
        ctx.sync()
    b.close()


def test_batch_against_oracle(ctx, pockets):
    import fastgrid_amd as fg
    from oracle import oracle as orc
    jobs = _jobs(ctx, pockets)
    fg.batch_run(ctx, jobs)
    for a, j in zip(pockets, jobs):
        out = j.fetch()
        M = a.n_maps
        om, od = orc.affinity_maps(a)
        oe = orc.elec_map(a)
        for m in range(M):
            assert np.array_equal(fx.milli(out[m]), fx.milli(om[m])), m
        assert np.array_equal(fx.milli(out[M + 1]), fx.milli(od))
        assert np.abs(out[M] - oe).max() < 3e-4
        if a.floating_grid:
            assert np.array_equal(out[M + 2].astype(np.float32), orc.floating_grid(a))
    for j in jobs:
        j.close()


def test_batch_with_tiny_jobs(ctx, pockets, tmp_path):
    """Jobs of 3 to 5 atoms on grids of 3 points along some axes, batched
    with two ordinary pockets: every job's maps equal its own run bit for
    bit, and the tiny jobs match the oracle."""
    import fastgrid_amd as fg
    from oracle import oracle as orc
    tiny = [fx.pocket(str(tmp_path), f"t{i}", 31 + i, n, npts)
            for i, (n, npts) in enumerate([(3, (3, 3, 3)), (5, (3, 41, 3)), (4, (41, 3, 5))])]
    arrays = [pockets[3], tiny[0], tiny[1], pockets[0], tiny[2]]
    jobs = _jobs(ctx, arrays)
    want = _single(jobs, fg.PREC_COMPENSATED)
    fg.batch_run(ctx, jobs)
    for i, j in enumerate(jobs):
        assert _same(j.fetch(), want[i]), i
    for k, a in zip((1, 2, 4), tiny):
        out = jobs[k].fetch()
        om, od = orc.affinity_maps(a)
        M = a.n_maps
        for m in range(M):
            assert np.array_equal(fx.milli(out[m]), fx.milli(om[m])), (k, m)
        assert np.array_equal(fx.milli(out[M + 1]), fx.milli(od)), k
        assert np.abs(out[M] - orc.elec_map(a)).max() < 3e-4, k


def test_shapes_and_autotuner(ctx, pockets):
    """Every shape gives the same bits, within the compensated bound of the
    oracle (the FP32 partials fold after every 64 atoms whatever the tile);
    the autotuner gives every job a valid shape and one common (px, pz)."""
    import fastgrid_amd as fg
    from oracle import oracle as orc
    a = pockets[1]
    job = fg.Job(ctx, a)
    oe = orc.elec_map(a)
    M = a.n_maps
    first = None
    for px in (1, 2):
        for pz in (4, 8):
            for tile in (32, 64, 128):
                for bx in (8, 16, 32, 256):
                    job.shape = (pz, tile, bx, px)
                    assert job.shape == (pz, tile, bx, px)
                    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
                    e = job.fetch()[M]
                    assert np.abs(e - oe).max() < 3e-4, (px, pz, tile, bx)
                    if first is None:
                        first = e
                    else:
                        assert np.array_equal(e, first), (px, pz, tile, bx)
    with pytest.raises(fg.FgError):
        job.shape = (6, 64, 16)
    jobs = _jobs(ctx, pockets)
    report = fg.autotune(ctx, jobs)
    print(report)
    assert report.startswith("px ")
    shapes = [j.shape for j in jobs]
    assert len({(s[0], s[3]) for s in shapes}) == 1
    for pz, tile, bx, px in shapes:
        assert pz in (4, 8) and tile in (32, 64, 128) and bx in (8, 16, 32, 256) and px in (1, 2)
    # one job: the e-kernel's best shapes are compared over whole steps too,
    # and the map bits do not depend on the pick
    before = job.fetch()
    line = fg.autotune(ctx, [job])
    print(line)
    assert "whole step (all maps)" in line
    pz, tile, bx, px = job.shape
    assert pz in (4, 8) and tile in (32, 64, 128) and bx in (8, 16, 32, 256) and px in (1, 2)
    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    assert _same(job.fetch(), before)
    # tuned jobs still reproduce through the batch
    want = _single(jobs, fg.PREC_COMPENSATED)
    fg.batch_run(ctx, jobs)
    for i, j in enumerate(jobs):
        assert _same(j.fetch(), want[i]), i


def test_create_jobs_equals_one_by_one(ctx, pockets):
    """fg_job_create_many builds the host-side inputs on threads: its jobs
    compute the same maps as jobs made one by one, and an invalid job in the
    list fails the call without keeping any job."""
    import dataclasses
    import fastgrid_amd as fg
    one = _jobs(ctx, pockets)
    many = fg.create_jobs(ctx, pockets, threads=4)
    want = _single(one, fg.PREC_COMPENSATED)
    got = _single(many, fg.PREC_COMPENSATED)
    for i in range(len(pockets)):
        assert _same(got[i], want[i]), i
    for j in one + many:
        j.close()
    bad = dataclasses.replace(pockets[0], n=(0, 21, 21))
    with pytest.raises(fg.FgError) as ei:
        fg.create_jobs(ctx, [pockets[1], bad, pockets[2]], threads=2)
    assert "job 1" in str(ei.value) and "invalid grid" in str(ei.value)
