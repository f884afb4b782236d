"""Batched GEMMs of uneven shape (config 4's mix of double- and quad-double
problems): the fused persistent pipeline with hipGraph replay against the
per-problem loop and the oracle."""
import numpy as np
import pytest

from conftest import assert_bits_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


def _batch(mdg, orc, count, seed, lo=16, hi=160):
    rng = np.random.default_rng(seed)
    probs, host = [], []
    for i in range(count):
        L = 2 if i % 2 == 0 else 4
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, 3))
        A = orc.generate(orc.GEN_QD, L, m, k, seed * 1000 + 2 * i, -8, 8)
        B = orc.generate(orc.GEN_QD, L, k, n, seed * 1000 + 2 * i + 1, -8, 8)
        probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B),
                      "C": to_dev(np.full(m * n * L, np.nan))})
        host.append((L, m, n, k, A, B))
    return probs, host


def _refill(probs, host, orc, seed):
    # new data in the same device buffers (a graph replay reads them afresh)
    for i, (p, (L, m, n, k, _, _)) in enumerate(zip(probs, host)):
        A = orc.generate(orc.GEN_QD, L, m, k, seed + 2 * i, -8, 8)
        B = orc.generate(orc.GEN_QD, L, k, n, seed + 2 * i + 1, -8, 8)
        p["A"].copy_(to_dev(A))
        p["B"].copy_(to_dev(B))
        host[i] = (L, m, n, k, A, B)


def test_batched_vector_bit_identical(mdg, orc, torch_cuda):
    probs, host = _batch(mdg, orc, 24, 3)
    for persistent in (0, 1, 1):   # loop, graph build, graph replay
        for p in probs:
            p["C"].fill_(np.nan)
        mdg.gemm_batched(probs, method="vector", persistent=persistent)
        torch_cuda.cuda.synchronize()
        for p, (L, m, n, k, A, B) in zip(probs, host):
            assert_bits_equal(to_host(p["C"]), orc.gemm(L, m, n, k, A, B, mode=orc.TP_FMA),
                              f"vector batch persistent={persistent} L={L} {m}x{n}x{k}")


def test_batched_int8_persistent_equals_loop(mdg, orc, torch_cuda):
    probs, host = _batch(mdg, orc, 24, 5)
    mdg.gemm_batched(probs, method="int8", persistent=0, spread_bits=16)
    torch_cuda.cuda.synchronize()
    loop = [to_host(p["C"]).copy() for p in probs]
    for p, (L, m, n, k, A, B), C in zip(probs, host, loop):
        err, _ = orc.normwise_error(L, m, n, k, A, B, C)
        assert err <= 4.0 * k * 2.0 ** (-53 * L), (L, m, n, k, err)
    for rep in range(3):           # build + capture, then two graph replays
        for p in probs:
            p["C"].fill_(np.nan)
        mdg.gemm_batched(probs, method="int8", persistent=1, spread_bits=16)
        torch_cuda.cuda.synchronize()
        for i, (p, C) in enumerate(zip(probs, loop)):
            assert_bits_equal(to_host(p["C"]), C, f"int8 batch rep={rep} problem {i}")


def test_batched_int8_graph_replay_sees_new_data(mdg, orc, torch_cuda):
    probs, host = _batch(mdg, orc, 12, 7)
    mdg.gemm_batched(probs, method="int8", persistent=1, spread_bits=16)   # build
    for rep in range(3):
        _refill(probs, host, orc, 9000 + 100 * rep)
        torch_cuda.cuda.synchronize()
        mdg.gemm_batched(probs, method="int8", persistent=1, spread_bits=16)   # replay
        torch_cuda.cuda.synchronize()
        for p, (L, m, n, k, A, B) in zip(probs, host):
            err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(p["C"]))
            assert err <= 4.0 * k * 2.0 ** (-53 * L), (rep, L, m, n, k, err)


@pytest.mark.parametrize("method", ["f64", "bf16"])
def test_batched_loop_methods(mdg, orc, torch_cuda, method):
    probs, host = _batch(mdg, orc, 8, 11)
    for persistent in (0, 1, 1):
        mdg.gemm_batched(probs, method=method, persistent=persistent, spread_bits=16)
        torch_cuda.cuda.synchronize()
        for p, (L, m, n, k, A, B) in zip(probs, host):
            err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(p["C"]))
            assert err <= 4.0 * k * 2.0 ** (-53 * L), (method, persistent, L, m, n, k, err)


def _batch_all_l(orc, count, seed, lo=16, hi=200):
    rng = np.random.default_rng(seed)
    probs, host = [], []
    for i in range(count):
        L = (2, 4, 8, 16)[i % 4]
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, 3))
        emax = 8 if L <= 4 else 30
        A = orc.generate(orc.GEN_QD, L, m, k, seed * 1000 + 2 * i, -emax, emax)
        B = orc.generate(orc.GEN_QD, L, k, n, seed * 1000 + 2 * i + 1, -emax, emax)
        probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B),
                      "C": to_dev(np.full(m * n * L, np.nan))})
        host.append((L, m, n, k, A, B))
    return probs, host


def test_batched_int8_residue_persistent_equals_loop(mdg, orc, torch_cuda):
    # every limb count in one fused residue batch, bit for bit against the
    # per-problem calls, then two graph replays on fresh data
    probs, host = _batch_all_l(orc, 16, 13)
    mdg.gemm_batched(probs, method="int8_residue", persistent=0, spread_bits=16)
    torch_cuda.cuda.synchronize()
    loop = [to_host(p["C"]).copy() for p in probs]
    for p, (L, m, n, k, A, B), C in zip(probs, host, loop):
        err, _ = orc.normwise_error(L, m, n, k, A, B, C, nsamp=512)
        assert err <= 4.0 * k * 2.0 ** (-53 * L), (L, m, n, k, err)
    for rep in range(3):           # build + capture, then two graph replays
        for p in probs:
            p["C"].fill_(np.nan)
        mdg.gemm_batched(probs, method="int8_residue", persistent=1, spread_bits=16)
        torch_cuda.cuda.synchronize()
        for i, (p, C) in enumerate(zip(probs, loop)):
            assert_bits_equal(to_host(p["C"]), C, f"int8_residue batch rep={rep} problem {i} L={p['L']}")
    _refill(probs, host, orc, 7700)
    torch_cuda.cuda.synchronize()
    mdg.gemm_batched(probs, method="int8_residue", persistent=1, spread_bits=16)   # replay
    torch_cuda.cuda.synchronize()
    for p, (L, m, n, k, A, B) in zip(probs, host):
        err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(p["C"]), nsamp=512)
        assert err <= 4.0 * k * 2.0 ** (-53 * L), (L, m, n, k, err)


def test_batched_int8_residue_deep_problem_falls_back(mdg, orc, torch_cuda):
    # a problem deeper than one INT32 plane per GEMM: the batch runs as
    # per-problem calls inside the graph, with the same results
    probs, host = _batch(mdg, orc, 3, 17, lo=16, hi=40)
    L, m, n, k = 2, 16, 8, 70000
    A = orc.generate(orc.GEN_QD, L, m, k, 1717, -8, 8)
    B = orc.generate(orc.GEN_QD, L, k, n, 1718, -8, 8)
    probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B),
                  "C": to_dev(np.full(m * n * L, np.nan))})
    host.append((L, m, n, k, A, B))
    mdg.gemm_batched(probs, method="int8_residue", persistent=0, spread_bits=16)
    torch_cuda.cuda.synchronize()
    loop = [to_host(p["C"]).copy() for p in probs]
    for rep in range(2):
        for p in probs:
            p["C"].fill_(np.nan)
        mdg.gemm_batched(probs, method="int8_residue", persistent=1, spread_bits=16)
        torch_cuda.cuda.synchronize()
        for i, (p, C) in enumerate(zip(probs, loop)):
            assert_bits_equal(to_host(p["C"]), C, f"deep batch rep={rep} problem {i}")
    err, _ = orc.normwise_error(L, m, n, k, A, B, loop[-1], nsamp=128)
    assert err <= 4.0 * k * 2.0 ** (-53 * L)



@pytest.mark.parametrize("method", ["int8", "int8_residue", "vector"])
def test_batched_padded_leading_dimensions(mdg, orc, torch_cuda, method):
    # operands and results inside wider buffers: the fused pipeline against
    # the per-problem calls, bit for bit, padding untouched
    rng = np.random.default_rng(23)
    probs, pads = [], []
    for i in range(10):
        L = (2, 4)[i % 2]
        m, n, k = (int(v) for v in rng.integers(16, 150, 3))
        lda, ldb, ldc = k + 2, n + 3, n + 1
        A = orc.generate(orc.GEN_QD, L, m, lda, 500 + 2 * i, -8, 8)
        B = orc.generate(orc.GEN_QD, L, k, ldb, 501 + 2 * i, -8, 8)
        probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B), "lda": lda, "ldb": ldb,
                      "ldc": ldc, "C": to_dev(np.full(m * ldc * L, 9.0))})
        pads.append((m, n, ldc, L))
    mdg.gemm_batched(probs, method=method, persistent=0, spread_bits=16)
    torch_cuda.cuda.synchronize()
    loop = [to_host(p["C"]).copy() for p in probs]
    for (m, n, ldc, L), C in zip(pads, loop):
        assert np.all(C.reshape(m, ldc, L)[:, n:] == 9.0), "loop wrote padding"
    for rep in range(2):
        for p in probs:
            p["C"].fill_(9.0)
        mdg.gemm_batched(probs, method=method, persistent=1, spread_bits=16)
        torch_cuda.cuda.synchronize()
        for i, (p, C) in enumerate(zip(probs, loop)):
            assert_bits_equal(to_host(p["C"]), C, f"{method} padded batch rep={rep} problem {i}")
