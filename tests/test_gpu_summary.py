"""The GPU-side checks of a run (fg_job_summary): per map the minimum and
maximum of the rounded values, exactly the extremes of the maps the writers
print (GridMapList::calculateEnergyMinMax, FG/GridMap.cpp:214-224), and the
receptor's total charge summed in atom order as the reference sums it
(FG/InputDataLoader.cpp:325), bit for bit; the command line's log reports
them."""
import os
import re

import numpy as np
import pytest

from tests import fixtures as fx
from tests.test_gpu_parity import fastgrid4

pytestmark = pytest.mark.gpu


def _arrays(tmp, case):
    if case == "S":
        run, gpf, _ = fx.synth(tmp, "S")
        return run, gpf, fx.load_arrays(run, gpf), []
    if case == "colagen2":
        run, gpf = fx.colagen2(tmp)
        return run, gpf, fx.load_arrays(run, gpf, v4=True), ["--v4"]
    run, _ = fx.agtests(tmp)
    return run, f"{case}.gpf", fx.load_arrays(run, f"{case}.gpf"), []


@pytest.mark.parametrize("case", ["S", "colagen2", "x1hpv_hb7"])
def test_summary_matches_the_fetched_maps(ctx, tmp_path, case):
    import fastgrid_amd as fg
    run, gpf, a, args = _arrays(str(tmp_path), case)
    job = fg.Job(ctx, a)
    for prec in (fg.PREC_COMPENSATED, fg.PREC_PLAIN):
        job.run(fg.RUN_ALL, prec)
        out = job.fetch()
        sm = job.summary()
        assert len(sm["min"]) == a.n_maps + 2
        for m in range(a.n_maps + 2):
            r = fx.r3(out[m])
            assert sm["min"][m] == r.min() and sm["max"][m] == r.max(), (prec, m)
    q = 0.0
    for c in a.charge:  # the reference's loop, in atom order
        q += float(c)
    assert sm["has_charge"] and sm["total_charge"] == q
    assert sm["charge_residual"] == abs(q - round(q))
    assert sm["charge_integral"] == (abs(q - round(q)) < 0.005)
    if case == "S":  # the synthetic receptors are made neutral
        assert sm["charge_integral"]
    job.close()


def test_cli_log_reports_the_summary(tmp_path):
    run, gpf, a, args = _arrays(str(tmp_path), "colagen2")
    fastgrid4(run, gpf, *args)
    log = open(os.path.join(run, "fg.glg")).read()
    q = 0.0
    for c in a.charge:
        q += float(c)
    assert f"Total charge :\t\t\t{q:.2f} e" in log
    names = ["C", "HD", "OA", "e", "d"]
    suffix = {"e": "\tElectrostatic Potential", "d": "\tDesolvation Potential"}
    for m, n in enumerate(names):
        v = fx.read_map(os.path.join(run, f"colagen2.{n}.map"))[1] / 1000.0
        row = re.search(rf"^ {m + 1}\t {n}\t\s+(\S+)\t\s*(\S+){suffix.get(n, '')}$", log, re.M)
        assert row, n
        assert row.group(1) == f"{v.min():6.2f}".strip() and row.group(2) == f"{v.max():9.2e}".strip(), n
