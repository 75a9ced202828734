"""Shared pytest set-up.  Tests marked `gpu` need an MI355X; everything else
runs on the CPU (oracle, host front end, C-ABI symbol checks)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (runs on the MI355X box)")


@pytest.fixture(scope="session")
def ctx():
    import fastgrid_amd as fg
    c = fg.Context(0)
    yield c
    c.close()
