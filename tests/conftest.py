import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_NAME = "multiple-double-arithmetic-matrix-matrix-multiplication_amd"
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def mdg():
    """The product library; raises (never skips) when it is not built."""
    return importlib.import_module(PKG_NAME)


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def bits(a):
    """uint64 view for bitwise comparison (distinguishes -0.0 and NaNs)."""
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bits_equal(got, want, what=""):
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        bad = np.nonzero(g != w)[0]
        i = int(bad[0])
        raise AssertionError(f"{what}: {bad.size} of {g.size} doubles differ; first at {i}: "
                             f"got {np.asarray(got).ravel()[i]!r} want {np.asarray(want).ravel()[i]!r}")


def to_host(t):
    return t.detach().cpu().numpy()


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
