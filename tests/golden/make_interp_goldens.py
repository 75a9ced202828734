"""Writes tests/golden/interp/: golden vectors of AutoDock 4.2.1's trilinear
interpolation (autogrid_original_4_2_1_ref/autodock/trilinterp.cc), computed
by the reference's own function (oracle/_ref/ref_trilinterp, built by
`make -f oracle/ref.mk` from the reference sources).  Run in a container where
the reference sits at /root/reference.

Inputs (data, stored in the fixture):
  * AutoDock's shipped 1pgp map set (Tests/1pgp_rec.{C,HD,OA,P,e,d}.map and
    1pgp_rec.maps.fld, 41^3 points), as int32 thousandths: the files print
    3 decimals, so float(k / 1000.0) is the float sscanf("%f") reads;
  * the 1pgp ligand of Tests/1pgp_lig.pdbqt (types C HD OA P -> slots 0..3,
    e = 4, d = 5, as Tests/1pgp.dpf loads them): coordinates, charges;
  * 240 poses: the crystal pose, then seeded rigid-body moves of it (rotation
    about its centre, translation up to 4 A, so some atoms leave the grid);
    a second set of 40 poses with atoms 3 and 7 ignored.  The crystal pose's
    total is AutoDock's own expected intermolecular energy of this test
    (Tests/test_autodock4.py:263, -6.17 kcal/mol).
Outputs: per pose the reference's per-atom elec/emap, elec/emap totals and
returned value (all float, the reference's Real).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
AD_TESTS = "/root/reference/autogrid_original_4_2_1_ref/autodock/Tests"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_trilinterp")
sys.path.insert(0, ROOT)

LIGAND_TYPES = ["C", "HD", "OA", "P"]


def ligand():
    xyz, q, types = [], [], []
    for line in open(os.path.join(AD_TESTS, "1pgp_lig.pdbqt")):
        if line.startswith(("ATOM", "HETATM")):
            xyz.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
            q.append(float(line[70:76]))
            types.append(LIGAND_TYPES.index(line[77:79].strip()))
    return np.array(xyz, np.float32), np.array(q, np.float32), np.array(types, np.int32)


def rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def poses(xyz, n, rng, shift):
    c = xyz.astype(np.float64).mean(0)
    out = [xyz]
    for _ in range(n - 1):
        t = rng.uniform(-shift, shift, 3)
        out.append(((xyz - c) @ rotation(rng).T + c + t).astype(np.float32))
    return np.stack(out)


def run_reference(g, maps, xyz, q, aq, types, ignore, outside=1):
    nx1, ny1, nz1 = g.num_points1
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nx1, ny1, nz1, maps.shape[3], 4, 5, xyz.shape[1], xyz.shape[0], outside], np.int32).tofile(f)
            np.array(list(g.lo) + list(g.hi) + list(g.center) + [g.inv_spacing], np.float64).tofile(f)
            maps.astype(np.float32).tofile(f)
            types.tofile(f)
            q.tofile(f)
            aq.tofile(f)
            ignore.astype(np.int32).tofile(f)
            xyz.astype(np.float32).tofile(f)
        subprocess.run([REF_BIN, fin, fout], check=True)
        rec = np.fromfile(fout, np.float32).reshape(xyz.shape[0], 2 * xyz.shape[1] + 3)
    na = xyz.shape[1]
    return rec[:, :na], rec[:, na:2 * na], rec[:, 2 * na], rec[:, 2 * na + 1], rec[:, 2 * na + 2]


def main():
    from oracle import oracle as orc
    out = os.path.join(HERE, "interp")
    os.makedirs(out, exist_ok=True)
    g = orc.read_fld(os.path.join(AD_TESTS, "1pgp_rec.maps.fld"))
    names = LIGAND_TYPES + ["e", "d"]
    maps = orc.read_mapset(g, [os.path.join(AD_TESTS, f"1pgp_rec.{n}.map") for n in names])
    milli = np.rint(maps.astype(np.float64) * 1000).astype(np.int32)
    assert np.array_equal((milli / 1000.0).astype(np.float32), maps)
    xyz, q, types = ligand()
    aq = np.abs(q)
    rng = np.random.default_rng(20240611)
    sets = {}
    for tag, n, ign, shift in (("a", 240, [], 4.0), ("b", 40, [3, 7], 3.0)):
        ignore = np.zeros(len(q), np.int32)
        ignore[ign] = 1
        P = poses(xyz, n, rng, shift)
        elec, emap, et, mt, tot = run_reference(g, maps, P, q, aq, types, ignore)
        sets.update({f"{tag}_xyz": P, f"{tag}_ignore": ignore, f"{tag}_elec": elec, f"{tag}_emap": emap,
                     f"{tag}_elec_total": et, f"{tag}_emap_total": mt, f"{tag}_total": tot})
    np.savez_compressed(os.path.join(out, "1pgp_trilinterp.npz"), maps_milli=milli, charge=q, abs_charge=aq,
                        types=types, fld_spacing=g.spacing, fld_num_points=np.array(list(g.num_points), np.int32),
                        fld_center=np.array(list(g.center)), **sets)
    print("wrote", os.path.join(out, "1pgp_trilinterp.npz"))


if __name__ == "__main__":
    main()
