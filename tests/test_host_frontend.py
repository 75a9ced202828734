"""The host front end of libfgmaps (GPF/PDBQT parsing, tables, bond vectors)
against the oracle's own parse and precompute of the same files, bit for bit,
and the multi-threaded preparation of many GPFs.  Host code only: no GPU."""
import os
import shutil

import numpy as np
import pytest

from oracle import oracle as orc
from tests import fixtures as fx

FIELDS_D = ("x", "y", "z", "w", "charge", "vol", "solpar", "rvector", "rvector2", "eps", "desolv")
FIELDS_I = ("type", "hbond", "rexp")


def compare_with_oracle(run, gpf, **kw):
    import fastgrid_amd as fg
    cwd = os.getcwd()
    os.chdir(run)
    try:
        a = fg.Prepared(gpf, **kw).arrays()
        o = orc.Loaded(gpf, **kw)
    finally:
        os.chdir(cwd)
    assert a.n_atoms == o.N and a.n_maps == o.M and tuple(a.n) == o.n and tuple(a.div2) == o.div2
    assert a.spacing == o.spacing and a.solpar_q == o.solpar_q and a.hydrogen == o.hydrogen
    for f in FIELDS_D:
        got = np.asarray(getattr(a, f), dtype=np.float64)
        want = o.darray(f)
        assert got.tobytes() == want.tobytes(), f
    for f in FIELDS_I:
        assert np.array_equal(np.asarray(getattr(a, f)), o.iarray(f)), f
    assert np.array_equal(a.disorder, o.disorder())
    assert a.n_types == o.tableTypes
    assert np.asarray(a.pair, dtype=np.float64).tobytes() == o.darray("pair").tobytes()
    return a


def test_colagen2_matches_oracle(tmp_path):
    run, gpf = fx.colagen2(str(tmp_path))
    compare_with_oracle(run, gpf, v4=True)


@pytest.mark.parametrize("order", [False, True])
def test_pgp_parameter_file_matches_oracle(tmp_path, order):
    run, gpf = fx.pgp(str(tmp_path))
    compare_with_oracle(run, gpf, v4=True, ag421_param_order=order)


@pytest.mark.parametrize("name", ["micro", "S"])
def test_synthetic_matches_oracle(tmp_path, name):
    run, gpf, _ = fx.synth(str(tmp_path), name)
    compare_with_oracle(run, gpf)


@pytest.mark.parametrize("name", fx.VARIANTS)
def test_gpf_variants_match_oracle(tmp_path, name):
    run, gpf, _, args = fx.variant(str(tmp_path), name)
    compare_with_oracle(run, gpf, v4="--v4" in args)


def test_pair_table_smoothing_is_a_running_minimum(tmp_path):
    """Each smoothed row is the minimum of the unsmoothed row over
    [k - 25, k + 25) (smooth 0.5 A), checked by brute force with the
    unsmoothed row recomputed from the closed form of its own coefficients
    (FG/PairwiseInteractionEnergies.cpp:93-114)."""
    run, gpf, _ = fx.synth(str(tmp_path), "micro")
    a = fx.load_arrays(run, gpf)
    pair = np.asarray(a.pair).reshape(a.n_types, orc.TABLE_LEN, a.n_maps)
    r = np.arange(orc.TABLE_LEN) / 100.0
    # one carbon-carbon row: AD4.1 C (R 4.00, eps 0.150, 12-6) with the vdW weight 0.1662
    m = 1  # ligand map C (synthetic maps: A C HD N OA)
    t = 1  # receptor type C
    rij, eij = 4.0, np.sqrt(0.150 * 0.150) * 0.1662
    cA, cB = eij / 6 * rij ** 12 * 6, eij / 6 * rij ** 6 * 12
    with np.errstate(divide="ignore", invalid="ignore"):
        raw = np.minimum(1e5, cA / r ** 12 - cB / r ** 6)
    raw[0], raw[-1] = 1e5, 0.0
    want = raw.copy()
    for k in range(1, orc.TABLE_LEN):
        want[k] = min(1e5, raw[max(0, k - 25):min(orc.TABLE_LEN, k + 25)].min())
    assert np.allclose(pair[t, :, m], want, rtol=1e-12, atol=1e-9)


def test_prepare_many_equals_one_by_one(tmp_path):
    """Eight synthetic receptors prepared on 4 threads give the same arrays as
    one-by-one preparation, in list order."""
    import fastgrid_amd as fg
    d = str(tmp_path)
    names = []
    for i in range(8):
        fg.synth_write(d, f"p{i}", 100 + i, 300 + 40 * i, "sphere", 12.0, (21 + 2 * i, 23, 25), "A C HD N OA")
        names.append(f"p{i}.gpf")
    cwd = os.getcwd()
    os.chdir(d)
    try:
        many = fg.prepare_many(names, threads=4)
        one = [fg.Prepared(n) for n in names]
        for p, q in zip(many, one):
            a, b = p.arrays(), q.arrays()
            assert a.n_atoms == b.n_atoms and tuple(a.n) == tuple(b.n)
            for f in FIELDS_D + ("pair",):
                assert np.asarray(getattr(a, f)).tobytes() == np.asarray(getattr(b, f)).tobytes(), f
        # a missing GPF fails the whole call with that GPF's error and exit code
        with pytest.raises(fg.FgError) as ei:
            fg.prepare_many(names[:3] + ["missing.gpf"] + names[3:], threads=3)
        assert "GPF 3" in str(ei.value) and ei.value.exit_code == 911
    finally:
        os.chdir(cwd)


def test_milli_writer_equals_fp64_writer(tmp_path):
    """The writer of integer thousandths (the command line's path, values
    rounded on the device) prints the same files as the FP64 writer: maps of
    random values at several scales with exact zeros, values that round to
    zero from either side, rounding ties and values near the 32-bit limit of
    thousandths, written from FP64, read back as thousandths and written
    again."""
    import fastgrid_amd as fg
    d = str(tmp_path)
    fg.synth_write(d, "w", 3, 200, "sphere", 8.0, (23, 21, 25), "A C HD N OA")
    cwd = os.getcwd()
    os.chdir(d)
    try:
        prep = fg.Prepared("w.gpf")
        M, P = prep.n_maps, prep.arrays().P
        rng = np.random.default_rng(5)
        edge = np.array([0.0, -0.0, 0.0004, -0.0004, 0.0005, -0.0005, 0.0015, -0.0025, 1.0005, -2.9995,
                         2147483.0, -2147483.0, 2147482.9996, 99999.9994])
        maps = []
        for m in range(M + 2):
            v = rng.standard_normal(P) * 10.0 ** rng.integers(-3, 6, P)
            v[rng.random(P) < 0.2] = 0.0
            v[:edge.size] = edge
            maps.append(v)
        prep.write_outputs(maps)
        files = sorted(f for f in os.listdir(".") if f.endswith((".map", ".fld")))
        a = {f: open(f, "rb").read() for f in files}
        fld = [f for f in files if f.endswith(".fld")][0]
        names = [ln.split("file=")[1].split()[0] for ln in a[fld].decode().splitlines() if ln.startswith("variable")]
        assert len(names) == M + 2
        milli = []
        for m in range(M + 2):
            _, k = fx.read_map(names[m])
            assert k.size == P and np.abs(k).max() < 2 ** 31
            milli.append(k.astype(np.int32))
        for f in files:
            os.remove(f)
        prep.write_outputs_milli(milli)
        b = {f: open(f, "rb").read() for f in files}
        assert a == b and len(a) == M + 3
    finally:
        os.chdir(cwd)
