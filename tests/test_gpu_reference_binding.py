"""The backend drop-in, end to end: the reference's own front end (parsing,
tables, bond vectors, covalent maps, floating grid, writers; built from its
sources by oracle/ref.mk) with libfgmaps bound in place of its two map
backends through fastgrid_amd/integration/amd_backend.cpp.  Its maps must
equal the goldens the unmodified reference writes (the e-map, computed in the
compensated form, within one printed unit).

The binding is built only where the reference's sources are (the container
that builds this repository); without it these tests are skipped."""
import os
import subprocess

import pytest

from tests import fixtures as fx
from tests.conftest import ROOT
from tests.test_gpu_parity import compare_files

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "oracle", "_ref", "fastgrid4_amd")
need = pytest.mark.skipif(not os.path.exists(EXE), reason="reference binding not built (no reference sources)")


def run_binding(run, gpf, *args):
    # zero-filled allocations: the reference leaves unbonded atoms' bond
    # vectors uninitialised (FG/BondVectors.cpp:34); the goldens were made so
    env = dict(os.environ, MALLOC_PERTURB_="255")
    p = subprocess.run([EXE, "-p", gpf, "-l", "ref.glg", *args], cwd=run, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert "Successful Completion" in open(os.path.join(run, "ref.glg")).read()


@need
def test_binding_colagen2_goldens(tmp_path):
    run, gpf = fx.colagen2(str(tmp_path))
    run_binding(run, gpf, "--v4")
    compare_files(run, "colagen2", fx.goldens("colagen2", "goldens_autogrid400.npz"),
                  digests=fx.text_digests("colagen2", "goldens_autogrid400.npz"))


@need
@pytest.mark.parametrize("name", ["covalent_qasp", "floating_grid", "constant_dielectric"])
def test_binding_gpf_variants(tmp_path, name):
    run, gpf, names, args = fx.variant(str(tmp_path), name)
    run_binding(run, gpf, *args)
    compare_files(run, "colagen2", fx.goldens("variants", f"{name}.npz"),
                  digests=fx.text_digests("variants", f"{name}.npz"))


@need
@pytest.mark.parametrize("name", ["S"])
def test_binding_synthetic(tmp_path, name):
    run, gpf, _ = fx.synth(str(tmp_path), name)
    run_binding(run, gpf)
    compare_files(run, name, fx.goldens("synth", f"{name}_goldens_fastgrid.npz"),
                  digests=fx.text_digests("synth", f"{name}_goldens_fastgrid.npz"))


@need
@pytest.mark.parametrize("name", ["x1hpv_CASE0", "x1hpv_hb7"])
def test_binding_x1hpv(tmp_path, name):
    run, _ = fx.agtests(str(tmp_path))
    run_binding(run, f"{name}.gpf")
    compare_files(run, "x1hpv", fx.goldens("agtests", f"{name}.npz"), digests=fx.text_digests("agtests", f"{name}.npz"))
