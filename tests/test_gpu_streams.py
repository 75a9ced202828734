"""The two streams of a context on an MI355X: the electrostatic kernel runs on
the second stream, concurrently with the lists and cutoff kernels, and the
compensated e-map is summed afterwards (e = near + far).  Runs of parts of the
maps, runs of other precisions in between and repeated runs must give the
same bits as the same run alone, and the timing events must bracket each
other."""
import numpy as np
import pytest

from tests import fixtures as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def job_arrays(tmp_path_factory):
    run, gpf, _ = fx.synth(str(tmp_path_factory.mktemp("streams")), "S")
    return fx.load_arrays(run, gpf)


def test_partial_and_interleaved_runs_match_a_full_run(ctx, job_arrays):
    import fastgrid_amd as fg
    a = job_arrays
    M = a.n_maps
    job = fg.Job(ctx, a)
    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    full = job.fetch()
    # the e-map alone: near part (the cutoff kernel without the affinity
    # maps, over candidate lists in atom order rather than sorted by type, so
    # its FP64 sums run in another order) and far part on the two streams,
    # then their sum
    job.run(fg.RUN_ELEC, fg.PREC_COMPENSATED)
    e = job.fetch()[M]
    d = np.abs(e - full[M]).max()
    print("RUN_ELEC vs RUN_ALL e-map: max |diff|", d)
    assert d <= 1e-9
    job.run(fg.RUN_ELEC, fg.PREC_COMPENSATED)
    assert np.array_equal(job.fetch()[M], e)
    # the affinity and desolvation maps alone
    job.run(fg.RUN_AFFINITY, fg.PREC_COMPENSATED)
    aff = job.fetch()
    for m in list(range(M)) + [M + 1]:
        assert np.array_equal(aff[m], full[m]), m
    # other forms in between leave nothing behind (the far part has its own
    # buffer; the FP32 and FP64 forms write e directly)
    for prec in (fg.PREC_FP32, fg.PREC_FP64, fg.PREC_PLAIN):
        job.run(fg.RUN_ALL, prec)
    for _ in range(3):
        job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    again = job.fetch()
    for m in range(M + 2):
        assert np.array_equal(again[m], full[m]), m
    job.close()


@pytest.mark.parametrize("prec", ["PREC_COMPENSATED", "PREC_FP32", "PREC_PLAIN"])
def test_timing_events_bracket_the_concurrent_phases(ctx, job_arrays, prec):
    import fastgrid_amd as fg
    job = fg.Job(ctx, job_arrays)
    job.run(fg.RUN_ALL | fg.RUN_PHASE_TIMING, getattr(fg, prec))
    t = ctx.timing()
    # ev0 .. ev2 encloses both the cutoff phase (ev0 .. ev1) and the
    # electrostatic kernel on the second stream (e0 .. e1, e0 after ev0, e1
    # before ev2)
    assert t.elec_ms > 0 and t.affinity_ms > 0
    assert t.total_ms + 1e-3 >= t.elec_ms and t.total_ms + 1e-3 >= t.affinity_ms
    # a run without the electrostatic map reports none
    job.run(fg.RUN_AFFINITY | fg.RUN_PHASE_TIMING, getattr(fg, prec))
    assert ctx.timing().elec_ms == 0
    # without phase timing a graph-replayed run (compensated, FP32) reports
    # its total only; the plain form always times its phases
    job.run(fg.RUN_ALL, getattr(fg, prec))
    t = ctx.timing()
    assert t.total_ms > 0
    if prec == "PREC_PLAIN":
        assert t.elec_ms > 0 and t.affinity_ms > 0
    else:
        assert t.elec_ms == -1 and t.affinity_ms == -1
    job.close()


def test_back_to_back_runs_and_graph_rebuilds(ctx, job_arrays):
    """Runs enqueued without waits between them, with shape changes (a new
    graph each) and phase timing switched on and off in between, give the
    bits of a lone run."""
    import fastgrid_amd as fg
    job = fg.Job(ctx, job_arrays)
    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    want = job.fetch()
    for shape in [(4, 64, 8, 2), (8, 32, 16, 1), (4, 128, 256, 1)]:
        job.shape = shape
        for flags in (0, fg.RUN_PHASE_TIMING, 0):
            for _ in range(3):
                job.run(fg.RUN_ALL | flags, fg.PREC_COMPENSATED)
        got = job.fetch()
        for m in range(len(want)):
            assert np.array_equal(got[m], want[m]), (shape, m)
    job.close()
