"""Vector GEMM on the GPU against the oracle's restatement of the reference's
host loops (C++/TCcuda.cu:320-372): bit-identical."""
import numpy as np
import pytest

from conftest import assert_bits_equal, to_dev, to_host

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (5, 7, 9), (16, 32, 16), (33, 47, 29), (40, 24, 70)]


def _inputs(orc, kind, L, m, n, k, seed):
    if kind == "qd":
        return (orc.generate(orc.GEN_QD, L, m, k, seed, -8, 8),
                orc.generate(orc.GEN_QD, L, k, n, seed + 1, -8, 8))
    if kind == "wide":
        return (orc.generate(orc.GEN_QD, L, m, k, seed, -60, 60),
                orc.generate(orc.GEN_QD, L, k, n, seed + 1, -60, 60))
    if kind == "ref":
        return (orc.generate(orc.GEN_REF, L, m, k, seed, 0, 0),
                orc.generate(orc.GEN_REF, L, k, n, seed + 1, 0, 0))
    if kind == "dd":
        return (orc.generate(orc.GEN_DD, L, m, k, seed),
                orc.generate(orc.GEN_DD, L, k, n, seed + 1))
    # small integers and zero limbs: every renormalisation takes the
    # zero-skipping branches
    rng = np.random.default_rng(seed)
    A = rng.integers(-3, 4, m * k * L).astype(np.float64)
    B = rng.integers(-3, 4, k * n * L).astype(np.float64)
    A[rng.random(A.size) < 0.4] = 0.0
    return A, B


@pytest.mark.parametrize("L", [2, 4, 8, 16])
@pytest.mark.parametrize("kind", ["qd", "wide", "ref", "dd", "int"])
def test_vector_gemm_bit_identical(mdg, orc, torch_cuda, L, kind):
    for (m, n, k) in SHAPES:
        if L == 16 and m * n * k > 30000:
            continue
        A, B = _inputs(orc, kind, L, m, n, k, 1000 * L + m)
        print(f"vector L={L} kind={kind} m={m} n={n} k={k}", flush=True)
        dA, dB = to_dev(A), to_dev(B)
        for tp_dev, tp_orc in ((mdg.TWOPROD_DEKKER, orc.TP_DEKKER), (mdg.TWOPROD_FMA, orc.TP_FMA)):
            C = to_host(mdg.vector_gemm(L, m, n, k, dA, dB, twoprod=tp_dev))
            want = orc.gemm(L, m, n, k, A, B, mode=tp_orc)
            assert_bits_equal(C, want, f"L={L} {kind} {m}x{n}x{k} tp={tp_dev}")


def test_vector_gemm_hexa_double_64x64x256(mdg, orc, torch_cuda):
    # hexa-double at a tile-filling shape: 4 x 4 tiles of 16, k spanning 16
    # k-blocks, both two_prod flavours
    L, m, n, k = 16, 64, 64, 256
    A = orc.generate(orc.GEN_QD, L, m, k, 16064, -20, 20)
    B = orc.generate(orc.GEN_QD, L, k, n, 16065, -20, 20)
    dA, dB = to_dev(A), to_dev(B)
    for tp_dev, tp_orc in ((mdg.TWOPROD_DEKKER, orc.TP_DEKKER), (mdg.TWOPROD_FMA, orc.TP_FMA)):
        C = to_host(mdg.vector_gemm(L, m, n, k, dA, dB, twoprod=tp_dev))
        want = orc.gemm(L, m, n, k, A, B, mode=tp_orc, nthreads=orc.num_threads())
        assert_bits_equal(C, want, f"hd 64x64x256 tp={tp_dev}")


def test_vector_gemm_leading_dimensions(mdg, orc, torch_cuda):
    L, m, n, k = 4, 9, 11, 13
    lda, ldb, ldc = 17, 15, 19
    rng = np.random.default_rng(3)
    A = rng.standard_normal(m * lda * L)
    B = rng.standard_normal(k * ldb * L)
    A.reshape(m, lda, L)[:, :, 1:] *= 2.0 ** -60
    B.reshape(k, ldb, L)[:, :, 1:] *= 2.0 ** -60
    C0 = np.full(m * ldc * L, 7.0)
    dC = to_dev(C0)
    mdg.vector_gemm(L, m, n, k, to_dev(A), to_dev(B), C=dC, lda=lda, ldb=ldb, ldc=ldc)
    C = to_host(dC).reshape(m, ldc, L)
    Ad = np.ascontiguousarray(A.reshape(m, lda, L)[:, :k, :]).ravel()
    Bd = np.ascontiguousarray(B.reshape(k, ldb, L)[:, :n, :]).ravel()
    want = orc.gemm(L, m, n, k, Ad, Bd, mode=orc.TP_FMA).reshape(m, n, L)
    assert_bits_equal(C[:, :n, :], want, "ld")
    assert np.all(C[:, n:, :] == 7.0), "wrote outside the n columns"


def test_vector_gemm_k_zero_gives_zero(mdg, torch_cuda):
    dA = to_dev(np.zeros(4))
    C = mdg.vector_gemm(2, 3, 2, 0, dA, dA, C=to_dev(np.ones(12)))
    assert np.all(to_host(C) == 0.0)


def test_fill_random_matches_oracle(mdg, orc, torch_cuda):
    for kind, L, emin, emax in ((0, 2, 0, 0), (1, 4, -8, 8), (1, 16, -60, 60), (2, 8, -5, 5)):
        got = to_host(mdg.fill_random(kind, L, 37, 11, 99, emin, emax))
        want = orc.generate(kind, L, 37, 11, 99, emin, emax)
        assert_bits_equal(got, want, f"generator kind={kind} L={L}")
