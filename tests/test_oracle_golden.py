"""The FP64 CPU oracle against the reference's own golden outputs.

Pins: the shipped AutoGrid 4.00 maps of the reference's end-to-end test
(tests/original_win32, run with --v4), AutoGrid 4.2.1's shipped 1pgp maps,
maps written by the reference FastGrid binary (parameter_file semantics,
synthetic receptors), and the dielectric / smoothed pair-energy tables the
golden log prints (tests/original_win32/log.glg)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests import fixtures as fx


def run_oracle(run, gpf, names, stem, **kw):
    assert orc.run_cli(gpf, run, **kw) == 0
    return {n: fx.read_map(os.path.join(run, f"{stem}.{n}.map"))[1] for n in names}


def test_colagen2_reproduces_shipped_goldens(tmp_path):
    run, gpf = fx.colagen2(str(tmp_path))
    gold = fx.goldens("colagen2", "goldens_autogrid400.npz")
    got = run_oracle(run, gpf, list(gold), "colagen2", v4=True)
    dig = fx.text_digests("colagen2", "goldens_autogrid400.npz")
    for n in gold:
        assert np.array_equal(got[n], gold[n]), n
        assert fx.value_digest(os.path.join(run, f"colagen2.{n}.map")) == dig[n], n
    assert open(os.path.join(run, "colagen2.maps.xyz")).read() == \
        open(os.path.join(fx.GOLDEN, "colagen2", "colagen2.maps.xyz")).read()


def test_pgp_parameter_file_fastgrid_order(tmp_path):
    """FastGrid reads parameter_file after the GPF (FG/main.cpp:170-171)."""
    run, gpf = fx.pgp(str(tmp_path))
    gold = fx.goldens("pgp", "goldens_fastgrid_v4.npz")
    got = run_oracle(run, gpf, list(gold), "1pgp_rec", v4=True)
    for n in gold:
        assert np.array_equal(got[n], gold[n]), n


def test_pgp_autogrid421_order(tmp_path):
    """Read where the GPF names it, the oracle reproduces AutoGrid 4.2.1's
    shipped maps; the affinity maps differ only where AutoGrid prints a value
    after a cast to float (values above 16384)."""
    run, gpf = fx.pgp(str(tmp_path))
    fast = fx.goldens("pgp", "goldens_fastgrid_v4.npz")
    delta = fx.goldens("pgp", "goldens_autogrid421_minus_fastgrid.npz")
    got = run_oracle(run, gpf, list(fast), "1pgp_rec", v4=True, ag421_param_order=True)
    for n in fast:
        ag = fast[n] + delta[n]
        diff = got[n] != ag
        if n in ("e", "d"):
            assert not diff.any(), n
        else:
            v = got[n][diff] / 1000.0
            assert np.all(np.abs(v) > 16384), n
            assert np.all(np.abs(np.float32(v).astype(np.float64) * 1000 - ag[diff]) <= 0.5 + 1e-6), n
    # and FastGrid's own order really differs on the H-bond maps
    assert not np.array_equal(fast["HD"], fast["HD"] + delta["HD"])


@pytest.mark.parametrize("name", ["micro", "S"])
def test_synthetic_receptors_match_reference_binary(tmp_path, name):
    run, gpf, names = fx.synth(str(tmp_path), name)
    gold = fx.goldens("synth", f"{name}_goldens_fastgrid.npz")
    got = run_oracle(run, gpf, names, name)
    dig = fx.text_digests("synth", f"{name}_goldens_fastgrid.npz")
    for n in names:
        assert np.array_equal(got[n], gold[n]), n
        assert fx.value_digest(os.path.join(tmp_path, f"{name}.{n}.map")) == dig[n], n


def test_dielectric_table_matches_golden_log():
    """332 / eps[i] printed at 0.1 A steps (entry 0 is AutoGrid's 1.00)."""
    rows = json.load(open(os.path.join(fx.GOLDEN, "colagen2", "log_tables.json")))["dielectric"]
    eps = orc.ddd_table()
    for r, printed in rows[1:]:
        i = int(round(r * 100))
        assert abs(round(332.0 / eps[i], 2) - printed) < 1e-9, (r, 332.0 / eps[i], printed)
    # FastGrid's own entry 0 is the closed form at r = 0, 332*8.7839/11.82819525
    assert abs(eps[0] - 332.0 * 8.7839 / 11.82819525) < 1e-9


def test_pair_tables_match_golden_log(tmp_path):
    """Smoothed pair energies of the C, HD, OA maps against each receptor type
    (AD4.0 parameters), printed to 2 decimals in the golden log."""
    run, gpf = fx.colagen2(str(tmp_path))
    cwd = os.getcwd()
    os.chdir(run)
    try:
        o = orc.Loaded(gpf, v4=True)
    finally:
        os.chdir(cwd)
    blocks = json.load(open(os.path.join(fx.GOLDEN, "colagen2", "log_tables.json")))["pair"]
    pair = o.darray("pair").reshape(o.tableTypes, orc.TABLE_LEN, o.M)
    for m, blk in enumerate(blocks):
        for row in blk["rows"]:
            j = int(round(row[0] * 100))
            for t, printed in enumerate(row[1:]):
                v = pair[t, j, m]
                if v < 100000:
                    assert abs(round(v, 2) - printed) < 1e-6, (m, t, j, v, printed)
                else:
                    assert printed == 100000.0


@pytest.mark.parametrize("name", fx.VARIANTS)
def test_gpf_variants_match_reference_binary(tmp_path, name):
    """Constant dielectric, disorder_h, covalent map + qasp, odd grid with
    AD4.1 defaults, floating grid: the reference binary's maps and .fld."""
    run, gpf, names, args = fx.variant(str(tmp_path), name)
    got = run_oracle(run, gpf, names, "colagen2", v4="--v4" in args)
    gold = fx.goldens("variants", f"{name}.npz")
    dig = fx.text_digests("variants", f"{name}.npz")
    for n in names:
        assert np.array_equal(got[n], gold[n]), n
        assert fx.value_digest(os.path.join(run, f"colagen2.{n}.map")) == dig[n], n
    assert fx.fld_lines(os.path.join(run, "colagen2.maps.fld")) == fx.variant_fld(name)
