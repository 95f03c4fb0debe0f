"""Test configuration: import paths, the `gpu` marker, shared fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ngsolve-sparsesolv_amd", "oracle", "tests", os.path.join("tests", "golden")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle("oracle")


@pytest.fixture(scope="session")
def ref():
    import pyoracle
    if not pyoracle.available("ref"):
        pytest.skip("reference-header build oracle/_ref/libssref.so not present")
    return pyoracle.Oracle("ref")


def load_golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
