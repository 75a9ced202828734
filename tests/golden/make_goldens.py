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
  agtests/   AutoGrid 4.2.1's own test cases (autogrid/Tests): x1hpv_CASE0-3 and the
             hsg1_sm GPFs with their receptors and parameter file, PyAutoDock's
             hsg1_sm maps, and the reference binary's maps for x1hpv_CASE0 and
             edits of it (larger grid; 4 and 7 maps with an H-bond min/max).
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


AGT = os.path.join(REF, "autogrid_original_4_2_1_ref", "autogrid", "Tests")
AGT_INPUTS = ["x1hpv_CASE0.gpf", "x1hpv_CASE1.gpf", "x1hpv_CASE2.gpf", "x1hpv_CASE3.gpf", "hsg1_sm.pdbqt",
              "hsg1_sm.gpf", "hsg1_sm_no_parameter_file.gpf", "hsg1_no_receptor_types.gpf",
              "hsg1_sm_minus_two_types.gpf", "AD4.1_bound.dat"]
HSG1_MAPS = ["A", "C", "HD", "N", "NA", "OA", "e", "d"]
# x1hpv_CASE0 (11 affinity maps) and edits of it: a larger grid, and ligand
# types repeated so that 4 and 7 maps keep an H-bond min/max (HD, HS: donor
# probes; NA: A1 acceptor probe).  Each edit list applies to CASE0's text.
X1HPV_MAPS = "A C HD N NA OA SA CL F S BR".split()


X1HPV = {
    "x1hpv_CASE0": ([], X1HPV_MAPS),
    "x1hpv_31": ([("npts 6 6 6", "npts 30 30 30")], X1HPV_MAPS),
    "x1hpv_hb4": ([("npts 6 6 6", "npts 24 24 24")], "A C HD HS N NA OA SA HD".split()),
    "x1hpv_hb7": ([("npts 6 6 6", "npts 20 20 20")], "A C HD HS N NA OA SA HD NA HS HD".split()),
}
# GPFs the reference refuses: exit status (FG/LogFile.cpp:147-149, FG/Exceptions.h:28-39)
AGT_FAILING = ["x1hpv_CASE1.gpf", "x1hpv_CASE2.gpf", "x1hpv_CASE3.gpf", "hsg1_sm_minus_two_types.gpf"]


def x1hpv_gpf(name):
    """The text of an x1hpv GPF of X1HPV (CASE0 with its edits)."""
    text = open(os.path.join(HERE, "agtests", "x1hpv_CASE0.gpf")).read()
    edits, types = X1HPV[name]
    for a, b in edits:
        assert a in text, a
        text = text.replace(a, b)
    if types != X1HPV_MAPS:
        # the new ligand_types line in place of the old one, the map lines
        # in place of the old ones (before elecmap)
        lines = [l for l in text.split("\n") if not l.startswith("map ")]
        lines = [f"ligand_types {' '.join(types)}" if l.startswith("ligand_types") else l for l in lines]
        i = next(k for k, l in enumerate(lines) if l.startswith("elecmap"))
        lines[i:i] = [f"map x1hpv.{n}.map" for n in x1hpv_names(name)[:-2]]
        text = "\n".join(lines)
    return text


def x1hpv_names(name):
    """Map names: the ligand types, a repeated type numbered from 2."""
    types = X1HPV[name][1]
    return [f"{t}{'' if types[:i].count(t) == 0 else types[:i].count(t) + 1}" for i, t in enumerate(types)] + ["e", "d"]


def agtests():
    """AutoGrid 4.2.1's own test cases (autogrid/Tests/test_autogrid4.py):
    inputs, PyAutoDock's hsg1_sm maps (the expected values of its hsg1 tests,
    :149-180), the reference binary's maps for x1hpv_CASE0 and the edits in
    X1HPV, and the exit status it gives the GPFs it refuses (:294-311)."""
    out = os.path.join(HERE, "agtests")
    os.makedirs(out, exist_ok=True)
    for f in AGT_INPUTS:
        shutil.copy(os.path.join(AGT, f), out)
    with open(os.path.join(AGT, "x1hpv.pdbqt"), "rb") as f, \
            gzip.GzipFile(os.path.join(out, "x1hpv.pdbqt.gz"), "wb", compresslevel=9, mtime=0) as g:
        g.write(f.read())
    py = {n: read_map(os.path.join(AGT, f"py_hsg1_sm.{n}.map"))[1] for n in HSG1_MAPS}
    np.savez_compressed(os.path.join(out, "py_hsg1_sm.npz"), **py)
    meta = {"failing": {}}
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(AGT, "x1hpv.pdbqt"), d)
        for f in ("hsg1_sm.pdbqt", "AD4.1_bound.dat"):
            shutil.copy(os.path.join(AGT, f), d)
        for name in X1HPV:
            text = x1hpv_gpf(name)
            open(os.path.join(d, name + ".gpf"), "w").write(text)
            if name != "x1hpv_CASE0":
                open(os.path.join(out, name + ".gpf"), "w").write(text)
            run_ref(name + ".gpf", d)
            names = x1hpv_names(name)
            save_maps(os.path.join(out, f"{name}.npz"), [os.path.join(d, f"x1hpv.{n}.map") for n in names], names)
            meta[name] = {"names": names}
        for f in AGT_FAILING:
            shutil.copy(os.path.join(AGT, f), d)
            env = dict(os.environ, MALLOC_PERTURB_="255")
            p = subprocess.run([REF_BIN, "-p", f, "-l", "fail.glg"], cwd=d, env=env, stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL)
            last = [l for l in open(os.path.join(d, "fail.glg")).read().split("\n") if l.strip()][-1]
            meta["failing"][f] = {"exit_status": p.returncode, "last_log_line": last.split(": ", 1)[1]}
    json.dump(meta, open(os.path.join(out, "agtests.json"), "w"), indent=1)


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
    agtests()
    print("goldens written under", HERE)
