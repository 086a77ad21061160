"""Shared test set-up: import paths, markers, and the libraries under test.

* ``oracle`` -- the CPU checker (oracle/liboracle.so, built on demand).
* ``tc`` -- the product's host package (glsl-tree-construction_amd/tc), which
  loads the in-tree libtc.so.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "glsl-tree-construction_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def tclib():
    import tc

    tc.lib()
    return tc


@pytest.fixture(scope="session")
def gpu_ctx():
    """A tc Context on device 0 (GPU tests only)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a visible GPU")
    import tc

    ctx = tc.Context(0)
    yield ctx
    ctx.close()


def rng(seed):
    return np.random.default_rng(seed)
