"""The C ABI library loads and exports every entry point include/mdgemm.h
declares; argument errors are reported without touching a GPU."""
import ctypes
import subprocess

import pytest


def test_library_exports_every_declared_symbol(mdg):
    lib = mdg.lib()
    declared = mdg.declared_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_compat_symbols_exported(mdg):
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", mdg.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in ("matmulTCnt(", "matmulqdf(", "matmulhdf(", "split4pd(", "split8pd(", "mixsplit2(",
                 "pllsqueeze_old(", "bigB2(", "mat(", "zeros(", "part("):
        assert name in out, name


def test_reference_kernels_exported(mdg):
    """The __global__ ddmm/qdmm/odmm/hdmm of include/mdgemm_kernels.hpp: their
    host-side launch stubs are exported, so hipcc-built callers link."""
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", mdg.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in ("ddmm", "qdmm", "odmm", "hdmm"):
        assert f"__device_stub__{name}(double*, double*, double*, int)" in out, name


@pytest.mark.parametrize("L", [0, 1, 3, 5, 32])
def test_bad_limb_count_rejected(mdg, L):
    lib = mdg.lib()
    rc = lib.mdg_vector_gemm(L, 4, 4, 4, None, 4, None, 4, None, 4, 0, None)
    assert rc == 1 and b"limb count" in lib.mdg_last_error()
    rc = lib.mdg_gemm(L, 4, 4, 4, None, 4, None, 4, None, 4, 2, 0, None, None)
    assert rc == 1


def test_bad_shapes_rejected(mdg):
    lib = mdg.lib()
    assert lib.mdg_vector_gemm(4, -1, 4, 4, None, 4, None, 4, None, 4, 0, None) == 1
    assert lib.mdg_vector_gemm(4, 4, 4, 8, None, 4, None, 4, None, 4, 0, None) == 1  # lda < k
    assert b"leading dimension" in lib.mdg_last_error()
    assert lib.mdg_gemm(4, 4, 4, 4, None, 4, None, 4, None, 4, 9, 0, None, None) == 1
    assert lib.mdg_split4pd(None, 7, 2, None, None) == 1     # count not a multiple of p
    assert lib.mdg_mixsplit2(None, 3, None, None) == 1
    assert lib.mdg_tc_gemm(4, 4, 4, 4, None, None, None, 12, None) == 1  # mix only for L=2
    assert lib.mdg_fill_random(7, 4, 1, 0, 0, 0, None, None) == 1
    assert lib.mdg_matmulhost(0, 4, None, None, None, None) == 1          # p < 1
    assert lib.mdg_matmulhost(4, -1, None, None, None, None) == 1
    assert lib.mdg_renormhost(None, 4, 4, 2, None) == 1                   # big is 0 or 1
    assert b"big" in lib.mdg_last_error()


def test_null_pointers_rejected(mdg):
    lib = mdg.lib()
    rc = lib.mdg_vector_gemm(4, 4, 4, 4, None, 4, None, 4, None, 4, 0, None)
    assert rc == 1 and b"NULL" in lib.mdg_last_error()


def test_no_device_reports_hip_error(mdg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = mdg.lib()
    assert lib.mdg_init() == 3
    assert b"device" in lib.mdg_last_error()
