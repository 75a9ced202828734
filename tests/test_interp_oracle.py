"""AutoDock 4.2.1's consumer of the maps (autogrid_original_4_2_1_ref/autodock:
readfield, readmap, trilinterp) restated in oracle/fginterp.c, pinned on the
CPU:

* against the reference's own trilinterp compiled from its source
  (tests/golden/interp/, written by tests/golden/make_interp_goldens.py):
  every per-atom output, total and returned value bit for bit, on AutoDock's
  shipped 1pgp maps with its 1pgp ligand;
* against AutoDock's own expectation for that ligand's crystal pose
  (Tests/test_autodock4.py:263: intermolecular energy -6.17 kcal/mol);
* against closed forms: maps linear in the grid indices are reproduced at
  any point, a grid point returns its map values, an atom outside the grid
  gets 500 x its squared distance from the centre.

The product's host readers (fg_read_fld, fg_read_map) are compared with the
oracle's readers on the same files."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests import fixtures as fx

GOLD = os.path.join(fx.GOLDEN, "interp", "1pgp_trilinterp.npz")


def write_fld(path, spacing, num_points, center):
    with open(path, "w") as f:
        f.write("# AVS field file\n#\n#SPACING %.3f\n#NELEMENTS %d %d %d\n#CENTER %.3f %.3f %.3f\n"
                % (spacing, *num_points, *center))
        f.write("#MACROMOLECULE rec.pdbqt\n#GRID_PARAMETER_FILE rec.gpf\n#\n")


def write_map(path, spacing, num_points, center, values_milli):
    """A .map as the writers print it: 6 header lines, "0" or "%.3f" of k/1000."""
    with open(path, "w") as f:
        f.write("GRID_PARAMETER_FILE rec.gpf\nGRID_DATA_FILE rec.maps.fld\nMACROMOLECULE rec.pdbqt\n")
        f.write("SPACING %.3f\nNELEMENTS %d %d %d\nCENTER %.3f %.3f %.3f\n" % (spacing, *num_points, *center))
        for k in values_milli.reshape(-1):
            f.write("0\n" if k == 0 else "%.3f\n" % (k / 1000.0))


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD)
    return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def mapset(gold, tmp_path_factory):
    """The 1pgp map set written as files and read back by the oracle."""
    d = str(tmp_path_factory.mktemp("pgpmaps"))
    sp, npts, c = float(gold["fld_spacing"]), tuple(gold["fld_num_points"]), tuple(gold["fld_center"])
    write_fld(os.path.join(d, "rec.maps.fld"), sp, npts, c)
    paths = []
    for slot in range(gold["maps_milli"].shape[3]):
        p = os.path.join(d, f"rec.{slot}.map")
        write_map(p, sp, npts, c, gold["maps_milli"][..., slot])
        paths.append(p)
    g = orc.read_fld(os.path.join(d, "rec.maps.fld"))
    return d, g, paths, orc.read_mapset(g, paths)


def test_reader_reads_the_printed_floats(gold, mapset):
    _, _, _, maps = mapset
    assert np.array_equal(maps, (gold["maps_milli"] / 1000.0).astype(np.float32))


@pytest.mark.parametrize("tag", ["a", "b"])
def test_oracle_matches_reference_trilinterp(gold, mapset, tag):
    _, g, _, maps = mapset
    elec, emap, et, mt, tot = orc.trilinterp(g, maps, 4, 5, gold[f"{tag}_xyz"], gold["charge"], gold["abs_charge"],
                                             gold["types"], ignore=gold[f"{tag}_ignore"])
    assert np.array_equal(elec, gold[f"{tag}_elec"])
    assert np.array_equal(emap, gold[f"{tag}_emap"])
    assert np.array_equal(et.astype(np.float32), gold[f"{tag}_elec_total"])
    assert np.array_equal(mt.astype(np.float32), gold[f"{tag}_emap_total"])
    assert np.array_equal(tot, gold[f"{tag}_total"])


def test_crystal_pose_is_autodocks_expected_energy(gold):
    """AutoDock's test of this map set and ligand expects an intermolecular
    energy of -6.17 kcal/mol (printed to 2 decimals)."""
    assert round(float(gold["a_total"][0]), 2) == -6.17


def linear_case():
    n1 = (9, 11, 13)
    g = orc.OiGrid()
    g.spacing, g.inv_spacing = 0.5, 2.0
    for k in range(3):
        g.num_points[k] = n1[k] - 1
        g.num_points1[k] = n1[k]
        g.center[k] = (1.25, -2.0, 3.5)[k]
        g.lo[k] = g.center[k] - (g.num_points[k] // 2) * g.spacing
        g.hi[k] = g.center[k] + (g.num_points[k] // 2) * g.spacing
    z, y, x = np.meshgrid(np.arange(n1[2]), np.arange(n1[1]), np.arange(n1[0]), indexing="ij")
    # slot 0: 1 + 2x - 3y + 0.5z; slot 1 (e): -x + 4z; slot 2 (d): 0.25 + y
    maps = np.stack([1 + 2 * x - 3 * y + 0.5 * z, -x + 4.0 * z, 0.25 + y], axis=-1).astype(np.float32)
    return g, maps


def test_linear_maps_are_reproduced():
    """Trilinear interpolation is exact for fields linear in the indices; a
    wrong weight, corner or summation term shows at ~1e-1."""
    g, maps = linear_case()
    rng = np.random.default_rng(5)
    n = 200
    lo, hi = np.array(list(g.lo)), np.array(list(g.hi))
    xyz = (lo + rng.uniform(0.01, 0.99, (n, 1, 3)) * (hi - lo)).astype(np.float32)
    q = np.array([-0.4], np.float32)
    elec, emap, et, mt, tot = orc.trilinterp(g, maps, 1, 2, xyz, q, np.abs(q), np.array([0]))
    u = (xyz[:, 0, :].astype(np.float64) - lo) * g.inv_spacing
    f0 = 1 + 2 * u[:, 0] - 3 * u[:, 1] + 0.5 * u[:, 2]
    fe = -u[:, 0] + 4 * u[:, 2]
    fd = 0.25 + u[:, 1]
    qd = float(q[0])  # the float charge, as the reference's Real
    assert np.allclose(et, fe * qd, rtol=1e-12, atol=1e-12)
    assert np.allclose(mt, f0 + fd * -qd, rtol=1e-12, atol=1e-12)


def test_grid_points_and_outside_penalty():
    g, maps = linear_case()
    # an atom on grid point (i, j, k) = (3, 4, 5) reads exactly that point
    pt = np.array([g.lo[0] + 3 * g.spacing, g.lo[1] + 4 * g.spacing, g.lo[2] + 5 * g.spacing], np.float32)
    q = np.array([0.5], np.float32)
    elec, emap, et, mt, _ = orc.trilinterp(g, maps, 1, 2, pt.reshape(1, 1, 3), q, np.abs(q), np.array([0]))
    assert et[0] == float(maps[5, 4, 3, 1]) * 0.5
    assert mt[0] == float(maps[5, 4, 3, 0]) + float(maps[5, 4, 3, 2]) * 0.5
    # outside: 500 * |x - centre|^2, the same in both totals
    out = np.array([[[g.hi[0] + 1.0, g.center[1], g.center[2] + 0.5]]], np.float32)
    elec, emap, et, mt, tot = orc.trilinterp(g, maps, 1, 2, out, q, np.abs(q), np.array([0]))
    d = out[0, 0].astype(np.float64) - np.array(list(g.center))
    want = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * 500.0
    assert et[0] == want and mt[0] == want and elec[0, 0] == np.float32(want)


def test_product_readers_match_the_oracle(mapset):
    import fastgrid_amd as fg
    d, g, paths, maps = mapset
    pg = fg.read_fld(os.path.join(d, "rec.maps.fld"))
    for f in ("spacing", "inv_spacing"):
        assert getattr(pg, f) == getattr(g, f)
    for f in ("num_points", "num_points1", "lo", "hi", "center"):
        assert list(getattr(pg, f)) == list(getattr(g, f)), f
    assert np.array_equal(fg.read_mapset(pg, paths), maps)


def test_product_reader_errors(mapset, tmp_path):
    import fastgrid_amd as fg
    d, g, paths, _ = mapset
    pg = fg.read_fld(os.path.join(d, "rec.maps.fld"))
    # a map whose centre differs from the .fld's, and one with a missing value
    text = open(paths[0]).read().split("\n")
    bad_c = text[:5] + ["CENTER 0.000 0.000 0.000"] + text[6:]
    short = text[:-3] + [""]
    for name, lines in (("center.map", bad_c), ("short.map", short)):
        p = str(tmp_path / name)
        open(p, "w").write("\n".join(lines))
        with pytest.raises(fg.FgError):
            fg.read_mapset(pg, [p])
    odd = str(tmp_path / "odd.fld")
    write_fld(odd, 0.375, (41, 40, 40), (0.0, 0.0, 0.0))
    with pytest.raises(fg.FgError):
        fg.read_fld(odd)


def test_text_values_parse_like_sscanf(tmp_path):
    """The product reader's fast decimal path returns strtof's float for
    printed map values and for longer decimals (checked against libc's
    strtof of the same text)."""
    import fastgrid_amd as fg
    rng = np.random.default_rng(9)
    vals = np.concatenate([rng.integers(-10 ** 8, 10 ** 8, 600) / 1000.0,
                           rng.uniform(-100, 100, 300), rng.uniform(-1e-3, 1e-3, 100)])
    texts = ["%.3f" % v for v in vals[:600]] + ["%.8f" % v for v in vals[600:900]] + ["%.12g" % v for v in vals[900:]]
    n = (9, 9, 9)
    texts = texts[:729]
    g = fg.FgMapGrid()
    g.spacing, g.inv_spacing = 0.375, 1 / 0.375
    for k in range(3):
        g.num_points[k], g.num_points1[k] = n[k] - 1, n[k]
    p = str(tmp_path / "v.map")
    with open(p, "w") as f:
        f.write("A\nB\nC\nSPACING 0.375\nNELEMENTS 8 8 8\nCENTER 0.000 0.000 0.000\n")
        f.write("\n".join(texts) + "\n")
    got = fg.read_mapset(g, [p]).reshape(-1)
    libc = ctypes.CDLL(None)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    want = np.array([libc.strtof(t.encode(), None) for t in texts], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
