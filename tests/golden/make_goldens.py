"""Writes the golden fixtures under tests/golden/ (run in a container where the
reference sources sit at /root/reference and oracle/_ref is built by
`make -f oracle/ref.mk`).

Fixtures (all data: inputs and expected outputs):
  colagen2/  the reference's own end-to-end test: tests/colagen2.gpf.orig,
             tests/macromol_1t60_half_prot.pdbqt.orig (gzipped) and the shipped
             AutoGrid 4.00 golden maps tests/original_win32/*.map (run with --v4).
  pgp/       autogrid_original_4_2_1_ref/autodock/Tests/1pgp* with
             AD4_parameters.dat: AutoGrid 4.2.1's shipped maps, and the maps the
             reference FastGrid binary writes for the same GPF (--v4), which
             differ on HD/OA because FastGrid reads parameter_file after the GPF.
  synth/     seeded synthetic receptors from the project generator and the maps
             the reference FastGrid binary (oracle/_ref/fastgrid4_ref, CPU path,
             MALLOC_PERTURB_=255 so its uninitialised bond vectors are zero)
             writes for them; with the sha256 of the generated input files.
Maps are stored as int32 thousandths (the 3-decimal values the files print).
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "fastgrid4_ref")
sys.path.insert(0, ROOT)

SYNTH = {
    # name: seed, atoms, shape, dims, npts, ligand types
    "micro": (11, 400, "sphere", 14.0, (21, 21, 21), "A C HD N OA"),
    "S": (1, 1500, "sphere", 24.0, (41, 41, 41), "A C HD N OA"),
}


def read_map(path):
    text = open(path, "rb").read().decode().replace("\r", "")
    lines = text.split("\n")
    header = lines[:6]
    vals = np.array([float(v) for v in lines[6:] if v.strip()], dtype=np.float64)
    return header, np.rint(vals * 1000).astype(np.int32)


def value_digest(path):
    text = open(path, newline="").read().replace("\r\n", "\n")
    return hashlib.sha256("\n".join(text.split("\n")[6:]).encode()).hexdigest()


def save_maps(path, files, names):
    arrays = {}
    headers = {}
    digests = {}
    for name, f in zip(names, files):
        h, v = read_map(f)
        arrays[name] = v
        headers[name] = h
        # the value lines byte for byte (the header names paths; line ends
        # normalised, the shipped win32 goldens use CRLF)
        digests[name] = value_digest(f)
    np.savez_compressed(path, **arrays)
    json.dump(digests, open(path.replace(".npz", ".text_sha256.json"), "w"), indent=1)
    return headers


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def run_ref(gpf, cwd, extra=()):
    env = dict(os.environ, MALLOC_PERTURB_="255")
    subprocess.run([REF_BIN, "-p", gpf, "-l", "ref.glg", *extra], cwd=cwd, env=env, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def colagen2():
    out = os.path.join(HERE, "colagen2")
    os.makedirs(out, exist_ok=True)
    shutil.copy(os.path.join(REF, "tests", "colagen2.gpf.orig"), os.path.join(out, "colagen2.gpf"))
    with open(os.path.join(REF, "tests", "macromol_1t60_half_prot.pdbqt.orig"), "rb") as f, \
            gzip.GzipFile(os.path.join(out, "macromol_1t60_half_prot.pdbqt.gz"), "wb", compresslevel=9, mtime=0) as g:
        g.write(f.read())
    names = ["C", "HD", "OA", "e", "d"]
    files = [os.path.join(REF, "tests", "original_win32", f"colagen2.{n}.map") for n in names]
    save_maps(os.path.join(out, "goldens_autogrid400.npz"), files, names)
    xyz = open(os.path.join(REF, "tests", "original_win32", "colagen2.maps.xyz"), "rb").read().decode().replace("\r", "")
    open(os.path.join(out, "colagen2.maps.xyz"), "w").write(xyz)
    # tables the golden log prints: the dielectric at 0.1 A steps and the
    # smoothed pair energies of each map against each receptor type
    log = open(os.path.join(REF, "tests", "original_win32", "log.glg"), "rb").read().decode().replace("\r", "").split("\n")
    tables = {"dielectric": [], "pair": []}
    i = next(k for k, l in enumerate(log) if l.startswith("  d   Dielectric"))
    for l in log[i + 2:i + 2 + 51]:
        r, v = l.split()
        tables["dielectric"].append([float(r), float(v)])
    for k, l in enumerate(log):
        if l.startswith("Finding the lowest pairwise interaction energy"):
            cols = log[k + 2].split()[1:]
            rows = [[float(x) for x in ll.split()] for ll in log[k + 4:k + 4 + 51]]
            tables["pair"].append({"receptor_types": cols, "rows": rows})
    json.dump(tables, open(os.path.join(out, "log_tables.json"), "w"))


# GPF variants of the colagen2 receptor, each run through the reference binary
VARIANTS = {
    "constant_dielectric": ([("dielectric -0.1465", "dielectric 40"), ("npts 40 40 40", "npts 20 20 20")], [], ["--v4"]),
    "disorder_h": ([("npts 40 40 40", "npts 24 24 24")], ["disorder_h"], ["--v4"]),
    "covalent_qasp": ([("npts 40 40 40", "npts 16 18 20"), ("spacing 1 ", "spacing 0.375 "),
                       ("ligand_types  C HD OA", "ligand_types  C HD OA Z"),
                       ("map colagen2.OA.map", "map colagen2.OA.map\nmap colagen2.Z.map\ncovalentmap 1.7 1000.0 64.5 -16.0 37.2")],
                      ["qasp 0.01234"], ["--v4"]),
    "odd_grid_ad41": ([("npts 40 40 40", "npts 21 34 9"), ("spacing 1 ", "spacing 0.4 ")], [], []),
    # floating grid: the receptor's atoms reach past the grid on every side
    "floating_grid": ([("npts 40 40 40", "npts 26 30 22")], ["fmap colagen2.fmap.map"], ["--v4"]),
}


def variant_gpf(text, edits, extra):
    for a, b in edits:
        assert a in text, a
        text = text.replace(a, b)
    return text + "".join(l + "\n" for l in extra)


def variants():
    out = os.path.join(HERE, "variants")
    os.makedirs(out, exist_ok=True)
    base = open(os.path.join(REF, "tests", "colagen2.gpf.orig")).read()
    meta = {}
    for name, (edits, extra, args) in VARIANTS.items():
        with tempfile.TemporaryDirectory() as d:
            shutil.copy(os.path.join(REF, "tests", "macromol_1t60_half_prot.pdbqt.orig"),
                        os.path.join(d, "macromol_1t60_half_prot.pdbqt"))
            run = os.path.join(d, "run")
            os.makedirs(run)
            text = variant_gpf(base, edits, extra)
            open(os.path.join(d, "colagen2.gpf"), "w").write(text)
            run_ref("../colagen2.gpf", run, args)
            names = [l.split()[1].split(".")[1] for l in text.split("\n") if l.startswith("map ")] + ["e", "d"]
            if any(l.startswith("fmap ") for l in extra):
                names.append("fmap")
            save_maps(os.path.join(out, f"{name}.npz"), [os.path.join(run, f"colagen2.{n}.map") for n in names], names)
            # the AVS field file without its "Created by <program>" line
            fld = [l for l in open(os.path.join(run, "colagen2.maps.fld")).read().split("\n")
                   if not l.startswith("# Created by")]
            open(os.path.join(out, f"{name}.fld"), "w").write("\n".join(fld))
            meta[name] = {"args": args, "names": names, "edits": edits, "extra": extra}
    json.dump(meta, open(os.path.join(out, "variants.json"), "w"), indent=1)


def pgp():
    out = os.path.join(HERE, "pgp")
    os.makedirs(out, exist_ok=True)
    t = os.path.join(REF, "autogrid_original_4_2_1_ref", "autodock", "Tests")
    shutil.copy(os.path.join(t, "1pgp.gpf"), out)
    shutil.copy(os.path.join(t, "AD4_parameters.dat"), out)
    with open(os.path.join(t, "1pgp_rec.pdbqt"), "rb") as f, \
            gzip.GzipFile(os.path.join(out, "1pgp_rec.pdbqt.gz"), "wb", compresslevel=9, mtime=0) as g:
        g.write(f.read())
    names = ["C", "HD", "OA", "P", "e", "d"]
    with tempfile.TemporaryDirectory() as d:
        for f in ("1pgp.gpf", "1pgp_rec.pdbqt", "AD4_parameters.dat"):
            shutil.copy(os.path.join(t, f), d)
        run_ref("1pgp.gpf", d, ["--v4"])
        fast = {n: read_map(os.path.join(d, f"1pgp_rec.{n}.map"))[1] for n in names}
    np.savez_compressed(os.path.join(out, "goldens_fastgrid_v4.npz"), **fast)
    # AutoGrid 4.2.1's shipped maps, stored as the difference from FastGrid's
    ag = {n: read_map(os.path.join(t, f"1pgp_rec.{n}.map"))[1] - fast[n] for n in names}
    np.savez_compressed(os.path.join(out, "goldens_autogrid421_minus_fastgrid.npz"), **ag)


def synth():
    import fastgrid_amd as fg
    out = os.path.join(HERE, "synth")
    os.makedirs(out, exist_ok=True)
    meta = {}
    for name, (seed, n, shape, dims, npts, lig) in SYNTH.items():
        with tempfile.TemporaryDirectory() as d:
            fg.synth_write(d, name, seed, n, shape, dims, npts, lig)
            run_ref(f"{name}.gpf", d)
            names = lig.split() + ["e", "d"]
            save_maps(os.path.join(out, f"{name}_goldens_fastgrid.npz"),
                      [os.path.join(d, f"{name}.{m}.map") for m in names], names)
            meta[name] = {"seed": seed, "n_atoms": n, "shape": shape, "dims": dims, "npts": npts,
                          "ligand_types": lig, "pdbqt_sha256": sha(os.path.join(d, f"{name}.pdbqt")),
                          "gpf_sha256": sha(os.path.join(d, f"{name}.gpf"))}
    json.dump(meta, open(os.path.join(out, "synth.json"), "w"), indent=1)


if __name__ == "__main__":
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    colagen2()
    variants()
    pgp()
    synth()
    print("goldens written under", HERE)
