import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X (lib2d4 on cuda:0)")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("2d4_amd")


@pytest.fixture(scope="session")
def orc():
    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    if not oracle.ref_available("o2"):
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    return oracle.Ref("o2")


@pytest.fixture(scope="session")
def golden():
    return {k: np.load(os.path.join(GOLDEN, f"{k}_golden.npz")) for k in ("field", "utils", "dconf")}


@pytest.fixture(scope="session")
def ctx(pkg):
    """A lib2d4 context on cuda:0 (GPU tests only; fails loudly if the library or GPU is missing)."""
    c = pkg.Context(0)
    yield c
    c.close()
