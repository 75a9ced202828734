"""The backend drop-in boundary on an MI355X: the one-shot entry points that
replace the reference's two map backends, called through the C ABI on the
caller's host buffers.

- fg_elec_map / fg_elec_map_async replace calculateElectrostaticMapAsync +
  synchronizeCalculation (FG/electrostatics/Electrostatics.h:28-36);
- fg_affinity_d_maps / fg_affinity_d_maps_async replace calculateGridmaps
  (FG/CalculateGridmaps.h:30-31), covalent maps arriving as initial values in
  the output buffers as FG/main.cpp:32-52 leaves them.

Against the FP64 oracle: affinity and desolvation maps print identically (and
agree to 1e-12 relative unrounded), the FP64 e-map is bit-exact, the
compensated one within 3e-4 kcal/mol.  The asynchronous forms overlap with
work on the caller's context (FG/main.cpp:215-222) and give the same bits as
the synchronous ones."""
import numpy as np
import pytest

from tests import fixtures as fx

pytestmark = pytest.mark.gpu


def _arrays(tmp, case):
    if case == "S":
        run, gpf, _ = fx.synth(tmp, "S")
        return fx.load_arrays(run, gpf)
    if case == "colagen2":
        run, gpf = fx.colagen2(tmp)
        return fx.load_arrays(run, gpf, v4=True)
    run, gpf, _, args = fx.variant(tmp, case)
    return fx.load_arrays(run, gpf, v4="--v4" in args)


@pytest.mark.parametrize("case", ["S", "colagen2", "constant_dielectric", "covalent_qasp"])
def test_one_shot_entry_points_vs_oracle(ctx, tmp_path, case):
    import fastgrid_amd as fg
    from oracle import oracle as orc
    a = _arrays(str(tmp_path), case)
    oe = orc.elec_map(a)
    om, od = orc.affinity_maps(a)
    assert np.array_equal(ctx.elec_map(a, fg.PREC_FP64), oe)
    e = ctx.elec_map(a, fg.PREC_COMPENSATED)
    assert np.abs(e - oe).max() < 3e-4
    assert np.abs(fx.milli(e) - fx.milli(oe)).max() <= 1
    if case == "covalent_qasp":
        # the initial values come from fg_covalent_map through the output buffers
        assert a.covalent is None and a.cov_on and a.map_is_covalent.any()
    maps, d = ctx.affinity_d_maps(a)
    for m in range(a.n_maps):
        assert np.array_equal(fx.milli(maps[m]), fx.milli(om[m])), m
        assert np.allclose(maps[m], om[m], rtol=1e-12, atol=1e-9), m
    assert np.array_equal(fx.milli(d), fx.milli(od))
    assert np.allclose(d, od, rtol=1e-12, atol=1e-12)


def test_async_forms_overlap_and_agree(ctx, tmp_path):
    """The e-map started asynchronously, the affinity maps computed meanwhile
    on the caller's context (the reference's order, FG/main.cpp:215-222), two
    handles outstanding at once, a handle waited for after the context's
    other work: every result equals the synchronous call's bits."""
    import fastgrid_amd as fg
    a = _arrays(str(tmp_path), "colagen2")
    e_sync = ctx.elec_map(a, fg.PREC_COMPENSATED)
    maps_sync, d_sync = ctx.affinity_d_maps(a)
    pe = ctx.elec_map_async(a, fg.PREC_COMPENSATED)
    maps, d = ctx.affinity_d_maps(a)
    assert np.array_equal(pe.wait(), e_sync)
    assert all(np.array_equal(x, y) for x, y in zip(maps, maps_sync)) and np.array_equal(d, d_sync)
    pa = ctx.affinity_d_maps(a, asynchronous=True)
    pe = ctx.elec_map_async(a, fg.PREC_FP64)
    e64 = pe.wait()
    maps, d = pa.wait()
    assert all(np.array_equal(x, y) for x, y in zip(maps, maps_sync)) and np.array_equal(d, d_sync)
    assert np.abs(e64 - e_sync).max() < 3e-4
    assert fg.LIB.fg_wait(None) == 0
