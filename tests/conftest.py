import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sep-graph_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (HIP device)")


def _ensure_built():
    # Tests use what is built; build only when nothing is there yet.
    lib = os.path.join(PKG, "lib", "libsepgraph.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", PKG, "-j8"], check=True)
    if not os.path.exists(os.path.join(ORACLE, "liboracle.so")):
        subprocess.run(["make", "-C", ORACLE], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def sg():
    import sepgraph

    return sepgraph


@pytest.fixture(scope="session")
def orc():
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def gpu(sg):
    if sg.lib().sg_device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


_graph_cache = {}


@pytest.fixture(scope="session")
def rmat16_sym(sg):
    """Config 1: RMAT scale 16, edge factor 16, symmetrised, deduplicated."""
    if "rmat16" not in _graph_cache:
        _graph_cache["rmat16"] = sg.rmat_graph(16, 16, symmetrize=True, seed=1)
    return _graph_cache["rmat16"]
