"""The maps' consumer on an MI355X: AutoDock's trilinear interpolation of
ligand atoms in a map set held in HBM (fg_trilinterp) against the oracle's
restatement (pinned to the reference's compiled trilinterp by
tests/test_interp_oracle.py), bit for bit; and the text-free hand-off of a
job's maps (fg_mapset_from_job) against AutoDock's reading of the .fld/.map
files the writers print for the same job."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests import fixtures as fx
from tests.test_interp_oracle import GOLD, linear_case, write_fld, write_map

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.fixture(scope="module")
def pgp(tmp_path_factory):
    """AutoDock's 1pgp map set read by the product reader, and the oracle's
    reading of the same files."""
    import fastgrid_amd as fg
    z = np.load(GOLD)
    gold = {k: z[k] for k in z.files}
    d = str(tmp_path_factory.mktemp("pgp"))
    sp, npts, c = float(gold["fld_spacing"]), tuple(gold["fld_num_points"]), tuple(gold["fld_center"])
    write_fld(os.path.join(d, "rec.maps.fld"), sp, npts, c)
    paths = []
    for slot in range(gold["maps_milli"].shape[3]):
        paths.append(os.path.join(d, f"rec.{slot}.map"))
        write_map(paths[-1], sp, npts, c, gold["maps_milli"][..., slot])
    g = fg.read_fld(os.path.join(d, "rec.maps.fld"))
    og = orc.read_fld(os.path.join(d, "rec.maps.fld"))
    return gold, g, og, fg.read_mapset(g, paths)


@pytest.mark.parametrize("tag", ["a", "b"])
def test_trilinterp_matches_reference_vectors(ctx, pgp, tag):
    """AutoDock's 1pgp maps and ligand, 240 / 40 poses (some atoms outside the
    grid, two atoms ignored in the second set): every output equals the
    reference's golden vectors and the oracle's doubles."""
    import fastgrid_amd as fg
    gold, g, og, maps = pgp
    ms = fg.MapSet(ctx, g, maps)
    assert same(ms.fetch(), maps)
    xyz, ign = gold[f"{tag}_xyz"], gold[f"{tag}_ignore"]
    print(f"trilinterp set {tag}: {xyz.shape[0]} poses x {xyz.shape[1]} atoms", flush=True)
    r = fg.trilinterp(ctx, ms, xyz, gold["types"], gold["charge"], gold["abs_charge"], 4, 5, ignore=ign)
    oe, om, oet, omt, otot = orc.trilinterp(og, maps, 4, 5, xyz, gold["charge"], gold["abs_charge"], gold["types"],
                                            ignore=ign)
    assert same(r["elec"], gold[f"{tag}_elec"]) and same(r["emap"], gold[f"{tag}_emap"])
    assert same(r["total"], gold[f"{tag}_total"])
    assert same(r["elec_total"], oet) and same(r["emap_total"], omt)
    assert same(r["elec_total"].astype(np.float32), gold[f"{tag}_elec_total"])
    ms.close()


def test_trilinterp_linear_field_and_many_poses(ctx):
    """Maps linear in the grid indices, 20,000 random poses of 24 atoms over
    the whole box and beyond it, 4,099 poses of 17 atoms, and 64 poses of 300
    atoms (more than a workgroup's lanes) with ignored atoms: bit for bit with
    the oracle."""
    import fastgrid_amd as fg
    g0, maps = linear_case()
    g = fg.FgMapGrid()
    for f in ("spacing", "inv_spacing"):
        setattr(g, f, getattr(g0, f))
    for f in ("num_points", "num_points1", "lo", "hi", "center"):
        for k in range(3):
            getattr(g, f)[k] = getattr(g0, f)[k]
    rng = np.random.default_rng(17)
    lo, hi = np.array(list(g0.lo)), np.array(list(g0.hi))
    xyz = (lo - 0.5 + rng.uniform(0, 1, (20000, 24, 3)) * (hi - lo + 1.0)).astype(np.float32)
    types = rng.integers(0, 3, 24).astype(np.int32)
    q = rng.uniform(-0.7, 0.7, 24).astype(np.float32)
    ms = fg.MapSet(ctx, g, maps)
    r = fg.trilinterp(ctx, ms, xyz, types, q, np.abs(q), 1, 2)
    oe, om, oet, omt, otot = orc.trilinterp(g0, maps, 1, 2, xyz, q, np.abs(q), types)
    assert same(r["elec"], oe) and same(r["emap"], om)
    assert same(r["elec_total"], oet) and same(r["emap_total"], omt) and same(r["total"], otot)
    with pytest.raises(fg.FgError):
        fg.trilinterp(ctx, ms, xyz[:2], np.full(24, 3, np.int32), q, np.abs(q), 1, 2)
    # a ragged tail: 4,099 poses of 17 atoms
    # are 274 workgroups of 15 poses, dealt to 8 XCDs in stretches of 35
    odd = (lo - 0.5 + rng.uniform(0, 1, (4099, 17, 3)) * (hi - lo + 1.0)).astype(np.float32)
    ot, oq = types[:17], q[:17]
    r = fg.trilinterp(ctx, ms, odd, ot, oq, np.abs(oq), 1, 2)
    oe, om, oet, omt, otot = orc.trilinterp(g0, maps, 1, 2, odd, oq, np.abs(oq), ot)
    assert same(r["elec"], oe) and same(r["emap"], om)
    assert same(r["elec_total"], oet) and same(r["emap_total"], omt) and same(r["total"], otot)
    # poses of more atoms than a workgroup has lanes (chunked in-order sums)
    big = (lo - 0.5 + rng.uniform(0, 1, (64, 300, 3)) * (hi - lo + 1.0)).astype(np.float32)
    bt = rng.integers(0, 3, 300).astype(np.int32)
    bq = rng.uniform(-0.7, 0.7, 300).astype(np.float32)
    ign = (rng.uniform(0, 1, 300) < 0.05).astype(np.int32)
    r = fg.trilinterp(ctx, ms, big, bt, bq, np.abs(bq), 1, 2, ignore=ign)
    oe, om, oet, omt, otot = orc.trilinterp(g0, maps, 1, 2, big, bq, np.abs(bq), bt, ignore=ign)
    assert same(r["elec"], oe) and same(r["emap"], om)
    assert same(r["elec_total"], oet) and same(r["emap_total"], omt) and same(r["total"], otot)
    # totals alone (no per-atom outputs) are the same
    r2 = fg.trilinterp(ctx, ms, xyz, types, q, np.abs(q), 1, 2, per_atom=False)
    assert r2["elec"] is None and same(r2["total"], fg.trilinterp(ctx, ms, xyz, types, q, np.abs(q), 1, 2)["total"])
    ms.close()


def test_mapset_from_job_equals_reading_the_written_files(ctx, tmp_path):
    """The device hand-off of a job's maps equals AutoDock's reading of the
    .map files the writers print from the same maps, slot for slot; the
    grid equals readfield's reading of the written .fld."""
    import fastgrid_amd as fg
    run, gpf, names = fx.synth(str(tmp_path), "micro")
    cwd = os.getcwd()
    os.chdir(run)
    try:
        prep = fg.Prepared(gpf)
        job = fg.Job(ctx, prep)
        job.run(fg.RUN_ALL, fg.PREC_COMPENSATED)
        out = job.fetch()
        prep.write_outputs(out)
        fld = os.path.join(run, prep_fld(gpf))
        g = fg.read_fld(fld)
        want = fg.read_mapset(g, [f"micro.{n}.map" for n in names])
    finally:
        os.chdir(cwd)
    pg = prep.map_grid()
    for f in ("spacing", "inv_spacing"):
        assert getattr(pg, f) == getattr(g, f)
    for f in ("num_points", "num_points1", "lo", "hi", "center"):
        assert list(getattr(pg, f)) == list(getattr(g, f)), f
    ms = fg.MapSet.from_job(ctx, job, pg)
    got = ms.fetch()
    assert got.shape == want.shape
    assert same(got, want)
    # and the interpolation of the hand-off equals the oracle's on the files
    rng = np.random.default_rng(3)
    lo, hi = np.array(list(g.lo)), np.array(list(g.hi))
    xyz = (lo + rng.uniform(0.05, 0.95, (500, 16, 3)) * (hi - lo)).astype(np.float32)
    types = rng.integers(0, len(names) - 2, 16).astype(np.int32)
    q = rng.uniform(-0.5, 0.5, 16).astype(np.float32)
    e_slot, d_slot = len(names) - 2, len(names) - 1
    r = fg.trilinterp(ctx, ms, xyz, types, q, np.abs(q), e_slot, d_slot)
    og = orc.read_fld(fld)
    oe, om, oet, omt, otot = orc.trilinterp(og, want, e_slot, d_slot, xyz, q, np.abs(q), types)
    assert same(r["total"], otot) and same(r["elec_total"], oet) and same(r["emap_total"], omt)
    ms.close()
    job.close()


def prep_fld(gpf):
    """The gridfld the synthetic GPF names."""
    for line in open(gpf):
        if line.startswith("gridfld"):
            return line.split()[1]
    raise AssertionError("no gridfld")
