"""AutoGrid 4.2.1's own test cases (autogrid_original_4_2_1_ref/autogrid/Tests/
test_autogrid4.py) on the CPU: the FP64 oracle and the product's host front
end and command line.

- x1hpv_CASE0 (:294-296) must complete: the oracle writes the 11 maps the
  reference binary writes (tests/golden/agtests/, made by
  tests/golden/make_goldens.py), bit for bit; likewise for edits of it with a
  larger grid and with 4 and 7 maps that keep an H-bond min/max.
- x1hpv_CASE1-3 (:299-311: a missing map, elecmap or dsolvmap keyword) and
  hsg1_sm_minus_two_types must fail as the reference fails: exit status 254
  (ExitProgram(-2) after FATAL_ERROR, FG/InputDataLoader.cpp:757-780,
  FG/LogFile.cpp:147-149) and no "Successful Completion" in the log.
- hsg1_sm, hsg1_sm_no_parameter_file and hsg1_no_receptor_types (:149-257)
  must reproduce PyAutoDock's py_hsg1_sm maps within the test's tolerance
  (0.5 % above 0.3, else 0.01); the reference binary reproduces them exactly,
  and so must the oracle.
- The command line's exit codes for an unknown receptor atom type (-1,
  FG/InputDataLoader.cpp:278-282), an unknown switch (1,
  FG/ProgramParameters.cpp:103-107) and a fatal GPF error (-2): all decided
  before any GPU call, so they run here.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from tests import fixtures as fx
from tests.test_host_frontend import compare_with_oracle


def _cli(run, *args):
    import fastgrid_amd as fg
    exe = os.path.join(fg.BIN_DIR, "fastgrid4")
    return subprocess.run([exe, *args], cwd=run, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120)


def _success_in_log(path):
    """test_autogrid4.py:99-113: "Successful Completion" in the last 10 lines."""
    return any("Successful Completion" in l for l in open(path).read().split("\n")[-10:])


@pytest.mark.parametrize("name", fx.X1HPV)
def test_oracle_x1hpv_matches_reference_binary(tmp_path, name):
    run, meta = fx.agtests(str(tmp_path))
    gold = fx.goldens("agtests", f"{name}.npz")
    dig = fx.text_digests("agtests", f"{name}.npz")
    assert list(gold) == meta[name]["names"]
    assert orc.run_cli(f"{name}.gpf", run) == 0
    assert _success_in_log(os.path.join(run, "oracle.glg"))
    for n in gold:
        path = os.path.join(run, f"x1hpv.{n}.map")
        assert np.array_equal(fx.read_map(path)[1], gold[n]), n
        assert fx.value_digest(path) == dig[n], n


def test_x1hpv_map_counts():
    """CASE0 has 11 affinity maps (more than 8: the 16-map kernels); the
    edits keep an H-bond min/max on 4 (HD HS NA HD) and 7 maps."""
    import json
    meta = json.load(open(os.path.join(fx.AGTESTS, "agtests.json")))
    assert len(meta["x1hpv_CASE0"]["names"]) == 13
    assert [n for n in meta["x1hpv_hb4"]["names"] if n.rstrip("23") in ("HD", "HS", "NA")] == ["HD", "HS", "NA", "HD2"]
    assert len([n for n in meta["x1hpv_hb7"]["names"] if n.rstrip("23") in ("HD", "HS", "NA")]) == 7


@pytest.mark.parametrize("name", fx.X1HPV + fx.HSG1)
def test_host_front_end_matches_oracle(tmp_path, name):
    run, _ = fx.agtests(str(tmp_path))
    compare_with_oracle(run, f"{name}.gpf")


@pytest.mark.parametrize("gpf", fx.HSG1)
def test_oracle_hsg1_matches_pyautodock(tmp_path, gpf):
    run, _ = fx.agtests(str(tmp_path))
    assert orc.run_cli(f"{gpf}.gpf", run) == 0
    py = fx.goldens("agtests", "py_hsg1_sm.npz")
    for n in fx.HSG1_MAPS:
        got = fx.read_map(os.path.join(run, f"hsg1_sm.{n}.map"))[1]
        assert fx.hsg1_close(got, py[n]), n
        assert np.array_equal(got, py[n]), n


@pytest.mark.parametrize("gpf", ["x1hpv_CASE1.gpf", "x1hpv_CASE2.gpf", "x1hpv_CASE3.gpf",
                                 "hsg1_sm_minus_two_types.gpf"])
def test_refused_gpfs_fail_like_the_reference(tmp_path, gpf):
    import fastgrid_amd as fg
    run, meta = fx.agtests(str(tmp_path))
    want = meta["failing"][gpf]["exit_status"]
    assert want == 254
    assert orc.run_cli(gpf, run) == want
    p = _cli(run, "-p", gpf, "-l", "fg.glg")
    assert p.returncode == want, p.stderr
    assert not _success_in_log(os.path.join(run, "fg.glg"))
    assert "ERROR:" in open(os.path.join(run, "fg.glg")).read()
    cwd = os.getcwd()
    os.chdir(run)
    try:
        with pytest.raises(fg.FgError) as ei:
            fg.Prepared(gpf)
        assert ei.value.exit_code == -2
    finally:
        os.chdir(cwd)


def test_unknown_receptor_type_exits_minus_one(tmp_path):
    run, _ = fx.agtests(str(tmp_path))
    lines = open(os.path.join(run, "hsg1_sm.pdbqt")).read().split("\n")
    i = [k for k, l in enumerate(lines) if l.startswith("ATOM")][2]
    lines[i] = lines[i][:77] + "Xx"  # type column (FG/InputDataLoader.cpp:252-261)
    open(os.path.join(run, "hsg1_sm.pdbqt"), "w").write("\n".join(lines))
    gpf = "hsg1_sm_no_parameter_file.gpf"
    assert orc.run_cli(gpf, run) == 255
    p = _cli(run, "-p", gpf, "-l", "fg.glg")
    assert p.returncode == 255 and "unknown type: 'Xx'" in p.stderr
    assert not _success_in_log(os.path.join(run, "fg.glg"))


def test_unknown_switch_and_missing_gpf(tmp_path):
    run, _ = fx.agtests(str(tmp_path))
    p = _cli(run, "--bogus")
    assert p.returncode == 1 and "unknown switch" in p.stderr
    p = _cli(run, "-p", "missing.gpf", "-l", "fg.glg")
    assert p.returncode == 911 & 0xff  # exit(911) as the shell sees it
