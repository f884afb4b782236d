"""The reference's C++ interface (include/mdgemm_compat.hpp) end to end, driven
by tests/cpp/tc_compat_test.cpp the way C++/TCtest.cu drives the reference,
checked against the oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, assert_bits_equal

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "cpp", "tc_compat_test")


@pytest.mark.parametrize("n", [16, 64])   # 64: the size C++/TCtest.cu:7-19 runs
@pytest.mark.parametrize("p", [2, 4, 8, 16])
def test_compat_driver(orc, tmp_path, p, n):
    assert os.path.exists(EXE), "tests/cpp/tc_compat_test is not built; run __graft_entry__.build()"
    pp = 4 * p
    r = subprocess.run([EXE, str(tmp_path), str(n), str(p)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout)
    ld = {f[:-4]: np.fromfile(tmp_path / f, dtype=np.float64) for f in os.listdir(tmp_path)}
    A = orc.generate(orc.GEN_REF, p, n, n, 2024, 0, 0)
    B = orc.generate(orc.GEN_REF, p, n, n, 2025, 0, 0)
    assert_bits_equal(ld["A"], A, "mat A")
    assert_bits_equal(ld["B"], B, "mat B")
    Aq, Bq = orc.split4pd(A, p), orc.split4pd(B, p)
    assert_bits_equal(ld["Aq"], Aq, "split4pd")
    assert_bits_equal(ld["Bq"], Bq, "split4pd")
    assert_bits_equal(ld["B2"], orc.bigB2(n, pp, Bq), "bigB2")
    C1q = orc.conv_gemm(n, n, n, pp, Aq, Bq)
    assert_bits_equal(ld["C1q"], C1q, "Aq * bigB2(Bq)")
    assert_bits_equal(ld["C1"], orc.squeeze(p, pp, C1q), "pllsqueeze_old")
    assert_bits_equal(ld["C2"], orc.gemm(p, n, n, n, A, B, mode=orc.TP_FMA), "matmulTCnt")
    assert_bits_equal(ld["C3"], orc.gemm(p, n, n, n, A, B, mode=orc.TP_DEKKER), "matmul*f")
    assert_bits_equal(ld["part1"], A[1::p], "part")
    A3 = A.reshape(n, n, p)
    assert_bits_equal(ld["bigA"], A3.transpose(0, 2, 1).ravel(), "bigA")
    B3 = B.reshape(n, n, p)
    want = np.zeros((p, n, p, n))     # [i][c][j][r] column-major blocks
    for i in range(p):
        for j in range(i + 1):
            want[i, :, j, :] = B3[:, :, i - j].T
    assert_bits_equal(ld["bigB"], want.ravel(), "bigB")
    assert_bits_equal(ld["mh"], orc.matmulhost(p, n, A, B), "matmulhost")
    assert_bits_equal(ld["rn"], orc.renormhost(A, n, p), "renormhost")
    assert_bits_equal(ld["rnbig"], orc.renormhost(A3.transpose(0, 2, 1).ravel(), n, p, big=True),
                      "renormhostbig")
    want_sq = orc.squeeze(p, pp, C1q)
    assert_bits_equal(ld["sq_p"], want_sq, "pllsqueeze")
    assert_bits_equal(ld["sqnt_p"], want_sq, "pllntsqueeze")
    if p == 2:
        assert_bits_equal(ld["x8"], orc.split4pd(A, 2), "split4pd")
        assert_bits_equal(ld["x16"], orc.split8pd(A, 2), "split8pd")
        assert_bits_equal(ld["x12"], orc.mixsplit2(A), "mixsplit2")
        for pp2 in (8, 12, 16):
            want2 = orc.squeeze(2, pp2, ld[f"x{pp2}"])
            assert_bits_equal(ld[f"sq_2_{pp2}"], want2, f"pllsqueeze_2_{pp2}")
            assert_bits_equal(ld[f"sqnt_2_{pp2}"], want2, f"pllntsqueeze_2_{pp2}")
        assert_bits_equal(ld["sq_12"], orc.squeeze(2, 12, ld["x12"]), "pllsqueeze(x, 2, 12)")


REF_KERNELS_EXE = os.path.join(ROOT, "tests", "cpp", "ref_kernels_test")


@pytest.mark.parametrize("p,n,nfrag", [(2, 64, 16), (4, 64, 8), (8, 48, 16), (16, 32, 32), (4, 96, 32)])
def test_exported_kernels_reference_launch(orc, tmp_path, p, n, nfrag):
    """The exported __global__ ddmm/qdmm/odmm/hdmm (C++/TCcuda.h:62-65),
    launched by a HIP program with the reference's grid Gr(n/nfrag,
    n/nfrag) x Bl(nfrag, nfrag) (C++/TCcuda.cu:265-293): bit-identical to
    the oracle's multi-double GEMM with FMA TwoProd."""
    assert os.path.exists(REF_KERNELS_EXE), "tests/cpp/ref_kernels_test is not built; run __graft_entry__.build()"
    A = orc.generate(orc.GEN_QD, p, n, n, 700 + p, -8, 8)
    B = orc.generate(orc.GEN_QD, p, n, n, 800 + p, -8, 8)
    A.tofile(tmp_path / "A.bin")
    B.tofile(tmp_path / "B.bin")
    r = subprocess.run([REF_KERNELS_EXE, str(p), str(n), str(nfrag), str(tmp_path / "A.bin"),
                        str(tmp_path / "B.bin"), str(tmp_path / "C.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    C = np.fromfile(tmp_path / "C.bin", dtype=np.float64)
    assert_bits_equal(C, orc.gemm(p, n, n, n, A, B, mode=orc.TP_FMA), f"exported kernel p={p}")
