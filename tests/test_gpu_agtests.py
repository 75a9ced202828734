"""AutoGrid 4.2.1's own test cases on an MI355X, through the command line,
the job API and the batch.

x1hpv_CASE0 has 11 affinity maps, so it runs the 16-map cutoff kernels (job,
batch and plain form); x1hpv_31 is the same receptor on a 31^3 grid (128
cutoff bricks); x1hpv_hb4 and x1hpv_hb7 keep an H-bond min/max on 4 and 7
maps (repeated HD, HS and NA ligand types), the 4-slot and the 16-slot
kernels.  Goldens: the reference binary's maps (tests/golden/agtests/); the
hsg1_sm GPFs against PyAutoDock's maps, which the reference reproduces
exactly."""
import os

import numpy as np
import pytest

from tests import fixtures as fx
from tests.test_gpu_parity import _kernel_check, compare_files, fastgrid4

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", fx.X1HPV)
def test_cli_x1hpv_goldens(tmp_path, name):
    run, _ = fx.agtests(str(tmp_path))
    fastgrid4(run, f"{name}.gpf")
    compare_files(run, "x1hpv", fx.goldens("agtests", f"{name}.npz"), digests=fx.text_digests("agtests", f"{name}.npz"))


@pytest.mark.parametrize("name", fx.X1HPV)
def test_cli_x1hpv_fp64_form_is_exact(tmp_path, name):
    run, _ = fx.agtests(str(tmp_path))
    fastgrid4(run, f"{name}.gpf", "--precision=fp64")
    compare_files(run, "x1hpv", fx.goldens("agtests", f"{name}.npz"), exact_e=True)


@pytest.mark.parametrize("gpf", fx.HSG1)
def test_cli_hsg1_matches_pyautodock(tmp_path, gpf):
    run, _ = fx.agtests(str(tmp_path))
    fastgrid4(run, f"{gpf}.gpf")
    py = fx.goldens("agtests", "py_hsg1_sm.npz")
    for n in fx.HSG1_MAPS:
        got = fx.read_map(os.path.join(run, f"hsg1_sm.{n}.map"))[1]
        assert fx.hsg1_close(got, py[n]), n
        if n == "e":
            assert np.abs(got - py[n]).max() <= 1
        else:
            assert np.array_equal(got, py[n]), n


@pytest.mark.parametrize("name", ["x1hpv_31", "x1hpv_hb4", "x1hpv_hb7"])
def test_kernels_vs_oracle_many_maps(ctx, tmp_path, name):
    run, _ = fx.agtests(str(tmp_path))
    a = fx.load_arrays(run, f"{name}.gpf")
    assert a.n_maps > 8
    _kernel_check(ctx, a)


def test_batch_with_many_map_jobs(ctx, tmp_path):
    """Jobs of 11, 9 and 12 maps (7 with a min/max) batched with a 6-map
    pocket: one 16-map cutoff kernel with 16 min/max slots serves all, and
    every job's maps equal its own run bit for bit."""
    import fastgrid_amd as fg
    run, _ = fx.agtests(str(tmp_path))
    arrays = [fx.load_arrays(run, f"{n}.gpf") for n in ("x1hpv_31", "x1hpv_hb4", "x1hpv_hb7")]
    arrays.append(fx.pocket(str(tmp_path), "p", 21, 2500, (33, 29, 35)))
    jobs = [fg.Job(ctx, a) for a in arrays]
    for prec in (fg.PREC_COMPENSATED, fg.PREC_FP32):
        want = []
        for j in jobs:
            j.run(fg.RUN_ALL, prec)
            want.append(j.fetch())
        for j in jobs:  # poison the maps: the batch must write every point
            j.run(fg.RUN_ALL, fg.PREC_PLAIN)
        fg.batch_run(ctx, jobs, fg.RUN_ALL, prec)
        for i, j in enumerate(jobs):
            got = j.fetch()
            assert all(np.array_equal(x, y) for x, y in zip(got, want[i])), (prec, i)
    for j in jobs:
        j.close()
