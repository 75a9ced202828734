"""Helpers that materialise the golden fixtures of tests/golden/ into a
scratch directory, read map files, and apply the reference's 3-decimal
rounding (FG/autogrid.h:128-131)."""
from __future__ import annotations

import gzip
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def r3(a):
    a = np.asarray(a, dtype=np.float64)
    return np.where(np.abs(a) < 0.0005, 0.0, np.floor(a * 1000 + 0.5) * 0.001)


def milli(a):
    """Rounded map values as int thousandths."""
    return np.rint(r3(a) * 1000).astype(np.int64)


def read_map(path):
    lines = open(path).read().split("\n")
    return lines[:6], np.rint(np.array([float(v) for v in lines[6:] if v.strip()]) * 1000).astype(np.int64)


def goldens(name, which):
    with np.load(os.path.join(GOLDEN, name, which)) as z:
        return {k: z[k].astype(np.int64) for k in z.files}


def _copy(src, dst):
    """Copies file content only (fixture files may be read-only)."""
    with open(src, "rb") as f, open(dst, "wb") as g:
        g.write(f.read())


def colagen2(tmp):
    """Returns (run_dir, gpf) with the receptor at ../ as the GPF expects."""
    src = os.path.join(GOLDEN, "colagen2")
    _copy(os.path.join(src, "colagen2.gpf"), os.path.join(tmp, "colagen2.gpf"))
    with gzip.open(os.path.join(src, "macromol_1t60_half_prot.pdbqt.gz"), "rb") as g:
        open(os.path.join(tmp, "macromol_1t60_half_prot.pdbqt"), "wb").write(g.read())
    run = os.path.join(tmp, "run")
    os.makedirs(run, exist_ok=True)
    return run, "../colagen2.gpf"


def pgp(tmp):
    src = os.path.join(GOLDEN, "pgp")
    for f in ("1pgp.gpf", "AD4_parameters.dat"):
        _copy(os.path.join(src, f), os.path.join(tmp, f))
    with gzip.open(os.path.join(src, "1pgp_rec.pdbqt.gz"), "rb") as g:
        open(os.path.join(tmp, "1pgp_rec.pdbqt"), "wb").write(g.read())
    return tmp, "1pgp.gpf"


def synth(tmp, name):
    """Regenerates a synthetic fixture input and checks it is the one the
    goldens were made from."""
    import fastgrid_amd as fg
    meta = json.load(open(os.path.join(GOLDEN, "synth", "synth.json")))[name]
    fg.synth_write(tmp, name, meta["seed"], meta["n_atoms"], meta["shape"], meta["dims"], meta["npts"],
                   meta["ligand_types"])
    for ext in ("pdbqt", "gpf"):
        h = hashlib.sha256(open(os.path.join(tmp, f"{name}.{ext}"), "rb").read()).hexdigest()
        assert h == meta[f"{ext}_sha256"], f"generator output for {name}.{ext} changed"
    return tmp, f"{name}.gpf", meta["ligand_types"].split() + ["e", "d"]


def load_arrays(run_dir, gpf, **kw):
    import fastgrid_amd as fg
    cwd = os.getcwd()
    os.chdir(run_dir)
    try:
        return fg.Prepared(gpf, **kw).arrays()
    finally:
        os.chdir(cwd)


def variant(tmp, name):
    """The colagen2 receptor with one of the GPF variants of
    tests/golden/variants/variants.json; returns (run_dir, gpf, names, args)."""
    meta = json.load(open(os.path.join(GOLDEN, "variants", "variants.json")))[name]
    run, gpf = colagen2(tmp)
    path = os.path.join(tmp, "colagen2.gpf")
    text = open(path).read()
    for a, b in meta["edits"]:
        assert a in text, a
        text = text.replace(a, b)
    open(path, "w").write(text + "".join(l + "\n" for l in meta["extra"]))
    return run, gpf, meta["names"], meta["args"]


def fld_lines(path):
    """An AVS field file's lines without the "Created by <program>" line."""
    return [l for l in open(path).read().split("\n") if not l.startswith("# Created by")]


def variant_fld(name):
    return open(os.path.join(GOLDEN, "variants", f"{name}.fld")).read().split("\n")


VARIANTS = sorted(json.load(open(os.path.join(GOLDEN, "variants", "variants.json"))))


POCKET_DENSITY = 0.03  # atoms per A^3 of a synthetic pocket sphere


def pocket(tmp, name, seed, n_atoms, npts, ligand_types="A C HD N OA SA"):
    """A synthetic receptor pocket: n_atoms packed in a sphere of density
    POCKET_DENSITY, grid of npts points (odd per axis) at 0.375 A on its
    centroid.  Returns the prepared JobArrays."""
    import fastgrid_amd as fg
    radius = (3.0 * n_atoms / (4.0 * np.pi * POCKET_DENSITY)) ** (1.0 / 3.0)
    fg.synth_write(tmp, name, seed, n_atoms, "sphere", radius, npts, ligand_types)
    return load_arrays(tmp, name + ".gpf")


def value_digest(path):
    """sha256 of a map file's value lines (after the 6 header lines), line
    ends normalised: byte-level parity of the writer."""
    text = open(path, newline="").read().replace("\r\n", "\n")
    return hashlib.sha256("\n".join(text.split("\n")[6:]).encode()).hexdigest()


def text_digests(name, which):
    return json.load(open(os.path.join(GOLDEN, name, which.replace(".npz", ".text_sha256.json"))))


AGTESTS = os.path.join(GOLDEN, "agtests")


def agtests(tmp):
    """AutoGrid 4.2.1's own test inputs (receptors, GPFs, parameter file) in
    tmp; returns tmp and the fixture metadata (map names per x1hpv GPF, the
    reference's exit status for the GPFs it refuses)."""
    for f in os.listdir(AGTESTS):
        if f.endswith((".gpf", ".pdbqt", ".dat")):
            _copy(os.path.join(AGTESTS, f), os.path.join(tmp, f))
    with gzip.open(os.path.join(AGTESTS, "x1hpv.pdbqt.gz"), "rb") as g:
        open(os.path.join(tmp, "x1hpv.pdbqt"), "wb").write(g.read())
    return tmp, json.load(open(os.path.join(AGTESTS, "agtests.json")))


X1HPV = ["x1hpv_CASE0", "x1hpv_31", "x1hpv_hb4", "x1hpv_hb7"]
HSG1 = ["hsg1_sm", "hsg1_sm_no_parameter_file", "hsg1_no_receptor_types"]
HSG1_MAPS = ["A", "C", "HD", "N", "NA", "OA", "e", "d"]


def hsg1_close(got, want):
    """test_autogrid4.py:149-180: within 0.5 % where |value| > 0.3, else
    within 0.01 (values in thousandths)."""
    g, w = np.asarray(got) / 1000.0, np.asarray(want) / 1000.0
    big = np.abs(g) > 0.30
    rel = np.abs(g - w) / np.where(big, np.abs(g), 1.0) * 100.0
    return bool(np.all(np.where(big, rel < 0.5, np.abs(g - w) < 0.01)))


def z_slab(a, z0, nz):
    """The same job restricted to the z-planes z0 .. z0+nz-1 (the points keep
    their positions: div2 moves with the first plane); for oracle checks of
    sampled planes of grids too large for the oracle."""
    import copy
    b = copy.copy(a)
    b.n = (a.n[0], a.n[1], nz)
    b.div2 = (a.div2[0], a.div2[1], a.div2[2] - z0)
    assert a.covalent is None
    return b


def slab_of(v, a, z0, nz):
    """Values of the planes z0 .. z0+nz-1 of a map of job a (x fastest)."""
    plane = a.n[0] * a.n[1]
    return v[z0 * plane:(z0 + nz) * plane]
