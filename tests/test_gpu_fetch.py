"""Fetching maps: the pipelined copies of many jobs (fg_job_fetch_many), the
device-rounded integer thousandths (fg_job_fetch_milli[_many]) and the
writer that takes them (fg_write_outputs_milli) against the FP64 fetch and
writer, bit for bit and byte for byte; a job whose values pass 32-bit
thousandths is refused by the rounded fetch and written from FP64 by the
command line."""
import os

import numpy as np
import pytest

from tests import fixtures as fx
from tests.test_gpu_batch import POCKETS
from tests.test_gpu_parity import fastgrid4

pytestmark = pytest.mark.gpu


def test_fetch_many_and_milli(ctx, tmp_path):
    import dataclasses
    import fastgrid_amd as fg
    arrays = [fx.pocket(str(tmp_path), f"p{i}", seed, n, npts, lig) for i, (seed, n, npts, lig) in enumerate(POCKETS)]
    arrays[2] = dataclasses.replace(arrays[2], floating_grid=1)
    jobs = [fg.Job(ctx, a) for a in arrays]
    fg.batch_run(ctx, jobs)
    one = [j.fetch() for j in jobs]
    many = fg.fetch_many(ctx, jobs)
    milli = fg.fetch_many(ctx, jobs, milli=True)
    for i, j in enumerate(jobs):
        assert len(many[i]) == len(one[i])
        assert all(np.array_equal(x, y) for x, y in zip(many[i], one[i])), i
        M = arrays[i].n_maps
        assert len(milli[i]) == M + 2
        for m in range(M + 2):
            assert np.array_equal(milli[i][m], fx.milli(one[i][m])), (i, m)
            assert np.array_equal(j.fetch_milli()[m], milli[i][m]), (i, m)
    for j in jobs:
        j.close()


@pytest.mark.parametrize("case", ["colagen2", "floating_grid"])
def test_milli_writer_is_byte_identical(ctx, tmp_path, case):
    import fastgrid_amd as fg
    if case == "colagen2":
        run, gpf = fx.colagen2(str(tmp_path))
        args = ["--v4"]
    else:
        run, gpf, _, args = fx.variant(str(tmp_path), case)
    cwd = os.getcwd()
    os.chdir(run)
    try:
        prep = fg.Prepared(gpf, v4="--v4" in args)
        job = fg.Job(ctx, prep.arrays())
        job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
        full = job.fetch()
        M = prep.n_maps
        names = [f for f in sorted(os.listdir(".")) if f.endswith(".map")]
        prep.write_outputs(full)
        a = {f: open(f, "rb").read() for f in os.listdir(".") if f.endswith((".map", ".fld"))}
        for f in a:
            os.remove(f)
        prep.write_outputs_milli(job.fetch_milli(), fmap=full[M + 2] if len(full) > M + 2 else None)
        b = {f: open(f, "rb").read() for f in os.listdir(".") if f.endswith((".map", ".fld"))}
        assert a == b and len(a) >= M + 3, names
        job.close()
    finally:
        os.chdir(cwd)


def test_values_past_32bit_thousandths(ctx, tmp_path):
    """30 atoms on one spot: the grid point on it sums 30 clamped pair
    energies of 1e5, 3e9 thousandths.  The rounded fetch refuses the job; the
    command line writes it from FP64, equal to the oracle."""
    import fastgrid_amd as fg
    from oracle import oracle as orc
    lines = [f"ATOM  {i + 1:5d}  C   RES A   1       0.000   0.000   0.000  1.00  0.00    {0.1 * (-1) ** i:6.3f} C \n"
             for i in range(30)]
    open(os.path.join(tmp_path, "r.pdbqt"), "w").write("".join(lines))
    open(os.path.join(tmp_path, "r.gpf"), "w").write(
        "npts 4 4 4\ngridfld r.maps.fld\nspacing 0.375\nreceptor_types C\nligand_types C\nreceptor r.pdbqt\n"
        "gridcenter 0.0 0.0 0.0\nsmooth 0.5\nmap r.C.map\nelecmap r.e.map\ndsolvmap r.d.map\ndielectric -0.1465\n")
    a = fx.load_arrays(str(tmp_path), "r.gpf")
    job = fg.Job(ctx, a)
    job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
    assert job.fetch()[0].max() >= 3e6
    with pytest.raises(fg.FgError) as ei:
        job.fetch_milli()
    assert "32 bits" in str(ei.value)
    job.close()
    fastgrid4(str(tmp_path), "r.gpf")
    om, od = orc.affinity_maps(a)
    _, got = fx.read_map(os.path.join(tmp_path, "r.C.map"))
    assert np.array_equal(got, fx.milli(om[0]))
