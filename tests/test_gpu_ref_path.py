"""The reference's split ("tensor core") path on the GPU: splitter, block
Toeplitz layout, FP64 accumulate GEMM, limb convolution, squeeze, and the
fused chain, against the oracle."""
import numpy as np
import pytest

from conftest import assert_bits_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [2, 4, 8, 16])
def test_splitters_bit_identical(mdg, orc, torch_cuda, L):
    rng = np.random.default_rng(L)
    cases = {
        "ref": orc.generate(orc.GEN_REF, L, 300, 1, 5, -5, 5),
        "signed": orc.generate(orc.GEN_QD, L, 300, 1, 6, -30, 30),
    }
    z = orc.generate(orc.GEN_REF, L, 300, 1, 7, 0, 0)
    z[rng.random(z.size) < 0.3] = 0.0
    cases["zeros"] = z
    for tag, x in cases.items():
        print(f"split L={L} {tag}", flush=True)
        dx = to_dev(x)
        assert_bits_equal(to_host(mdg.split4pd(dx, L)), orc.split4pd(x, L), f"split4pd {tag}")
        assert_bits_equal(to_host(mdg.split8pd(dx, L)), orc.split8pd(x, L), f"split8pd {tag}")
        if L == 2:
            assert_bits_equal(to_host(mdg.mixsplit2(dx)), orc.mixsplit2(x), f"mixsplit2 {tag}")


@pytest.mark.parametrize("L", [2, 4, 8, 16])
def test_squeeze_bit_identical(mdg, orc, torch_cuda, L):
    rng = np.random.default_rng(10 + L)
    pp = 4 * L
    x = orc.generate(orc.GEN_QD, 1, 257 * pp, 1, 9, -20, 20)
    x[rng.random(x.size) < 0.2] = 0.0
    assert_bits_equal(to_host(mdg.squeeze(L, pp, to_dev(x))), orc.squeeze(L, pp, x), "squeeze")


def test_bigB2_layout(mdg, orc, torch_cuda):
    n, pp = 5, 6
    Bq = np.arange(n * n * pp, dtype=np.float64) + 1.0
    assert_bits_equal(to_host(mdg.bigB2(to_dev(Bq), n, pp)), orc.bigB2(n, pp, Bq), "bigB2")


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (37, 29, 45), (130, 70, 33), (256, 128, 64)])
def test_dgemm_acc_exact_on_integers(mdg, torch_cuda, M, N, K):
    # integer operands: every FP64 partial sum is exact, so any summation
    # order must give numpy's result exactly; C accumulates like the
    # reference's WMMA kernel (C++/TCtest.cu:32,45)
    rng = np.random.default_rng(M + N + K)
    A = rng.integers(-50, 51, (M, K)).astype(np.float64)
    B = rng.integers(-50, 51, (K, N)).astype(np.float64)
    C0 = rng.integers(-50, 51, (M, N)).astype(np.float64)
    dC = to_dev(C0.ravel())
    mdg.dgemm_acc(M, N, K, to_dev(A.ravel()), to_dev(B.T.ravel()), dC)   # B column-major
    assert np.array_equal(to_host(dC).reshape(M, N), C0 + A @ B)


@pytest.mark.parametrize("L,pp", [(2, 8), (4, 16), (8, 32), (2, 12)])
def test_conv_gemm_exact_domain(mdg, orc, torch_cuda, L, pp):
    # parts of the reference generator at equal leading exponents: every
    # anti-diagonal sum is exact, so C1q is order-independent
    n = 24
    A = orc.generate(orc.GEN_REF, L, n, n, 21, 0, 0)
    B = orc.generate(orc.GEN_REF, L, n, n, 22, 0, 0)
    if pp == 12:
        Aq, Bq = orc.mixsplit2(A), orc.mixsplit2(B)
    else:
        Aq, Bq = orc.split4pd(A, L), orc.split4pd(B, L)
        if pp == 8 * L:
            Aq, Bq = orc.split8pd(A, L), orc.split8pd(B, L)
    got = to_host(mdg.conv_gemm(n, n, n, pp, to_dev(Aq), to_dev(Bq)))
    assert_bits_equal(got, orc.conv_gemm(n, n, n, pp, Aq, Bq), f"conv L={L} pp={pp}")


@pytest.mark.parametrize("L,kind", [(2, 4), (2, 8), (2, 12), (4, 4), (8, 4), (16, 4), (4, 8)])
def test_tc_gemm_chain_matches_reference_chain(mdg, orc, torch_cuda, L, kind):
    # the reference's C++/TCtest.cu:53-114 chain, restated on the CPU:
    # split -> Aq * bigB2(Bq) -> pllsqueeze_old
    n = 20
    A = orc.generate(orc.GEN_REF, L, n, n, 31 + L, 0, 0)
    B = orc.generate(orc.GEN_REF, L, n, n, 41 + L, 0, 0)
    split = {4: lambda x: orc.split4pd(x, L), 8: lambda x: orc.split8pd(x, L),
             12: orc.mixsplit2}[kind]
    pp = {4: 4, 8: 8, 12: 6}[kind] * L
    C1q = orc.conv_gemm(n, n, n, pp, split(A), split(B))
    want = orc.squeeze(L, pp, C1q)
    got = to_host(mdg.tc_gemm(L, n, n, n, to_dev(A), to_dev(B), split_kind=kind))
    assert_bits_equal(got, want, f"tc chain L={L} kind={kind}")


@pytest.mark.parametrize("p", [1, 2, 4, 16])
def test_matmulhost_bit_identical(mdg, orc, torch_cuda, p):
    # mixed signs and exponents: the sums round, so only the reference's
    # order (k ascending, j = 0 .. i within it) gives the same bits
    n = 37
    rng = np.random.default_rng(40 + p)
    A = rng.standard_normal(n * n * p) * np.exp2(rng.integers(-30, 30, n * n * p))
    B = rng.standard_normal(n * n * p) * np.exp2(rng.integers(-30, 30, n * n * p))
    print(f"matmulhost p={p} n={n}", flush=True)
    got = to_host(mdg.matmulhost(p, n, to_dev(A), to_dev(B)))
    assert_bits_equal(got, orc.matmulhost(p, n, A, B), f"matmulhost p={p}")


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("p", [2, 4, 8, 16])
def test_renormhost_bit_identical(mdg, orc, torch_cuda, p, big):
    n = 45
    rng = np.random.default_rng(50 + p)
    A = rng.standard_normal(n * n * p) * np.exp2(rng.integers(-60, 60, n * n * p))
    print(f"renormhost p={p} big={big}", flush=True)
    d = to_dev(A)
    mdg.renormhost(d, n, p, big=big)
    assert_bits_equal(to_host(d), orc.renormhost(A, n, p, big=big), f"renormhost p={p} big={big}")
