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


def assert_bits_equal_nan(got, want, what=""):
    """Bitwise equality, except that any NaN matches any NaN: the sign and
    payload of a NaN made by 0/0 or Inf - Inf depend on the compiler's
    operand order and on the hardware, not on the algorithm."""
    g = np.array(got, dtype=np.float64).ravel()
    w = np.array(want, dtype=np.float64).ravel()
    gn, wn = np.isnan(g), np.isnan(w)
    if not np.array_equal(gn, wn):
        i = int(np.nonzero(gn != wn)[0][0])
        raise AssertionError(f"{what}: NaN at different positions; first at {i}: got {g[i]!r} want {w[i]!r}")
    g[gn] = 0.0
    w[wn] = 0.0
    assert_bits_equal(g, w, what)


def to_host(t):
    return t.detach().cpu().numpy()


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
