"""The C ABI: libfgmaps loads, exports every entry point include/fgmaps.h
declares, and refuses to compute without a GPU (no CPU fallback)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fgmaps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("fg_create", "fg_elec_map", "fg_affinity_d_maps", "fg_job_create", "fg_batch_run", "fg_prep_load"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import fastgrid_amd as fg
    lib = ctypes.CDLL(fg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared set
    assert set(fg.SIGNATURES) == set(declared_functions())


def test_round3_known_answers():
    import fastgrid_amd as fg
    cases = {0.0: 0.0, 0.0004999: 0.0, -0.0004999: 0.0, 0.0005: 0.001, -0.0005: 0.0, 1.2345: 1.235,
             -1.2345: -1.234, 2.0004: 2.0, 100000.0: 100000.0, -0.0015: -0.001}
    for a, want in cases.items():
        got = fg.round3(a)
        assert abs(got - want) < 1e-12, (a, got, want)


@pytest.mark.skipif(bool(int(os.environ.get("FG_EXPECT_GPU", "0"))), reason="GPU expected in this run")
def test_no_cpu_fallback_without_gpu():
    import fastgrid_amd as fg
    if fg.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(fg.FgError) as ei:
        fg.Context(0)
    assert "no GPU" in str(ei.value)
