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


def test_batched_int8_edge_shapes(mdg, orc, torch_cuda):
    """Row counts that end a 128-row tile in each of its four 32-row MFMA
    tiles (the band kernel leaves out the rows past m), depths that end
    anywhere in a 128-byte k-block (the slicing stage stores only the chunks
    that hold digits of k; the plan zeroed the rest once), one-row and
    one-column problems: the persistent pipeline, built and replayed on new
    data, is bit-identical to the loop of per-problem calls (the single-problem
    kernels compute every padded row) and within the exact bound."""
    shapes = [(1, 1, 1), (32, 5, 15), (33, 40, 127), (64, 1, 129), (65, 33, 300), (96, 128, 7),
              (97, 129, 200), (128, 64, 128), (129, 17, 33), (200, 96, 257)]
    rng = np.random.default_rng(21)
    probs, host = [], []
    for i, (m, n, k) in enumerate(shapes):
        L = (2, 4)[i % 2]
        A = orc.generate(orc.GEN_QD, L, m, k, 2100 + 2 * i, -8, 8)
        B = orc.generate(orc.GEN_QD, L, k, n, 2101 + 2 * i, -8, 8)
        probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B),
                      "C": to_dev(np.full(m * n * L, np.nan))})
        host.append((L, m, n, k, A, B))
    for rep in range(3):           # build + capture, then replays on new data
        if rep:
            _refill(probs, host, orc, int(rng.integers(1, 1 << 30)))
        mdg.gemm_batched(probs, method="int8", persistent=0, spread_bits=16)
        torch_cuda.cuda.synchronize()
        loop = [to_host(p["C"]).copy() for p in probs]
        for p in probs:
            p["C"].fill_(np.nan)
        mdg.gemm_batched(probs, method="int8", persistent=1, spread_bits=16)
        torch_cuda.cuda.synchronize()
        for i, (p, C, (L, m, n, k, A, B)) in enumerate(zip(probs, loop, host)):
            assert_bits_equal(to_host(p["C"]), C, f"int8 edge batch rep={rep} problem {i} {m}x{n}x{k}")
            err, _ = orc.normwise_error(L, m, n, k, A, B, C)
            assert err <= 4.0 * k * 2.0 ** (-53 * L), (rep, L, m, n, k, err)


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


def _gen_batch(mdg, orc, shapes, seed, emin, emax):
    probs, host = [], []
    for i, (L, m, n, k) in enumerate(shapes):
        A = orc.generate(orc.GEN_QD, L, m, k, seed + 2 * i, emin, emax)
        B = orc.generate(orc.GEN_QD, L, k, n, seed + 2 * i + 1, emin, emax)
        probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B),
                      "C": to_dev(np.zeros(m * n * L))})
        host.append((L, m, n, k, A, B))
    return probs, host


@pytest.mark.parametrize("method,persistent", [("int8", 1), ("int8_residue", 1), ("bf16", 1), ("int8", 0),
                                               ("int8_residue", 0)])
def test_batched_spread_beyond_allowance_is_flagged(mdg, orc, torch_cuda, method, persistent):
    # the slices are sized from spread_bits without a read-back; a problem
    # whose measured exponent spread is larger (here leading exponents in
    # [-30, 30] against an allowance of 2 bits) must not come back silently
    # inaccurate: its C is NaN and the flags report it
    shapes = [(2, 40, 36, 64), (4, 33, 20, 48), (2, 24, 24, 32)]
    # leading exponent 0 everywhere: measured spread 1 bit, within the allowance
    probs, host = _gen_batch(mdg, orc, shapes, 70, 0, 0)
    wide, _ = _gen_batch(mdg, orc, [(4, 40, 36, 256)], 90, -30, 30)
    probs[1] = wide[0]
    for rep in range(2):   # build, then replay
        for p in probs:
            p["C"].fill_(0.0)
        mdg.gemm_batched(probs, method=method, persistent=persistent, spread_bits=2)
        torch_cuda.cuda.synchronize()
        status, flags = mdg.batched_last_flags(len(probs))
        assert status == 6 and flags[1] & 2, (rep, status, flags)
        assert bool(torch_cuda.isnan(probs[1]["C"]).all()), "flagged problem not NaN-filled"
        for i in (0, 2):
            assert not bool(torch_cuda.isnan(probs[i]["C"]).any())
            L, m, n, k, A, B = host[i]
            err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(probs[i]["C"]))
            assert err <= 4.0 * k * 2.0 ** (-53 * L), (i, err)


def test_batched_default_allowance_covers_wide_exponent_ranges(mdg, orc, torch_cuda):
    # the measured spread E - log2(mean |x0|) of a row of k entries is at most
    # log2(k) + 1 (the mean is at least max / k), so the default 16-bit
    # allowance covers every k <= 2^15: leading exponents spread over 40
    # bits are not flagged and come back within the error bound
    shapes = [(4, 48, 40, 512), (2, 64, 32, 300)]
    probs, host = _gen_batch(mdg, orc, shapes, 110, -40, 0)
    mdg.gemm_batched(probs, method="int8", persistent=1)
    torch_cuda.cuda.synchronize()
    status, flags = mdg.batched_last_flags(len(probs))
    assert status == 0 and flags == [0, 0], flags
    for p, (L, m, n, k, A, B) in zip(probs, host):
        err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(p["C"]))
        assert err <= 4.0 * k * 2.0 ** (-53 * L), err


@pytest.mark.parametrize("method", ["int8", "int8_residue", "f64"])
def test_batched_non_finite_input_is_flagged(mdg, orc, torch_cuda, method):
    shapes = [(2, 20, 18, 40), (4, 17, 30, 25)]
    probs, host = _gen_batch(mdg, orc, shapes, 130, -8, 8)
    Ah = host[0][4].copy()
    Ah[7] = np.inf
    probs[0]["A"].copy_(to_dev(Ah))
    mdg.gemm_batched(probs, method=method, persistent=1, spread_bits=16)
    torch_cuda.cuda.synchronize()
    status, flags = mdg.batched_last_flags(2)
    assert status == 5 and flags[0] & 1 and flags[1] == 0, (status, flags)
    assert bool(torch_cuda.isnan(probs[0]["C"]).all())
    L, m, n, k, A, B = host[1]
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(probs[1]["C"]))
    assert err <= 4.0 * k * 2.0 ** (-53 * L), err


def test_batched_plan_cache_is_bounded(mdg, orc, torch_cuda):
    # three buffer sets cycled through persistent=1 twice with a two-plan
    # cache: device memory after the second cycle equals that after the
    # first (plans are evicted, not accumulated), every result equals the
    # loop's, and releasing the workspace frees the plans
    mdg.release_workspace()
    mdg.set_batch_plan_cache(2, 1 << 40)
    try:
        sets = [_gen_batch(mdg, orc, [(2, 60, 50, 70), (4, 40, 40, 90)], 200 + 10 * j, -8, 8) for j in range(3)]
        want = []
        for probs, _ in sets:
            mdg.gemm_batched(probs, method="int8", persistent=0)
            torch_cuda.cuda.synchronize()
            want.append([to_host(p["C"]).copy() for p in probs])
        free = []
        for cycle in range(2):
            for (probs, _), w in zip(sets, want):
                for p in probs:
                    p["C"].fill_(0.0)
                mdg.gemm_batched(probs, method="int8", persistent=1)
                torch_cuda.cuda.synchronize()
                for p, c in zip(probs, w):
                    assert_bits_equal(to_host(p["C"]), c, f"cycle {cycle}")
            free.append(torch_cuda.cuda.mem_get_info()[0])
        assert free[1] == free[0], free
        mdg.release_workspace()
        assert torch_cuda.cuda.mem_get_info()[0] > free[1]
    finally:
        mdg.set_batch_plan_cache(32, 16 << 30)


@pytest.mark.parametrize("method", ["bf16", "f64"])
def test_batched_loop_plan_survives_scratch_growth_and_release(mdg, orc, torch_cuda, method):
    # a persistent batch of a method without fused stages replays the
    # per-problem calls from a graph; it must not depend on the per-stream
    # scratch, which a larger call on the same stream regrows and
    # mdg_release_workspace frees
    probs, host = _gen_batch(mdg, orc, [(2, 40, 30, 50), (4, 36, 28, 44)], 300, -8, 8)
    mdg.gemm_batched(probs, method=method, persistent=0)
    torch_cuda.cuda.synchronize()
    want = [to_host(p["C"]).copy() for p in probs]
    mdg.gemm_batched(probs, method=method, persistent=1)   # build and capture
    torch_cuda.cuda.synchronize()
    L, n = 4, 512
    X = mdg.fill_random(mdg.GEN_QD, L, n, n, 5, -8, 8)
    mdg.gemm(L, n, n, n, X, X, method=method)                # regrows the stream's scratch
    for rep in range(2):
        if rep == 1:
            mdg.release_workspace()                           # frees scratch and plans
        for p in probs:
            p["C"].fill_(0.0)
        mdg.gemm_batched(probs, method=method, persistent=1)
        torch_cuda.cuda.synchronize()
        for p, w in zip(probs, want):
            assert_bits_equal(to_host(p["C"]), w, f"{method} replay {rep}")


@pytest.mark.parametrize("method", ["int8", "int8_residue", "vector"])
def test_batched_random_mix_persistent_equals_loop(mdg, orc, torch_cuda, method):
    # seeded batches of every limb count, ragged shapes (1 .. 200), exponent
    # ranges and padded leading dimensions: the replayed graph against the
    # loop of single calls bit for bit, and each result within the bound
    for trial in range(3):
        rng = np.random.default_rng(700 + trial)
        probs, host = [], []
        for i in range(int(rng.integers(8, 20))):
            L = int(rng.choice([2, 4, 8, 16]))
            m, n, k = (int(v) for v in rng.integers(1, 201, 3))
            e = int(rng.integers(0, 7))
            pad = int(rng.integers(0, 3))
            lda, ldb, ldc = k + pad, n + pad, n + 2 * pad
            A = orc.generate(orc.GEN_QD, L, m, lda, 9000 * trial + 2 * i, -e, e)
            B = orc.generate(orc.GEN_QD, L, k, ldb, 9000 * trial + 2 * i + 1, -e, e)
            probs.append({"L": L, "m": m, "n": n, "k": k, "A": to_dev(A), "B": to_dev(B), "lda": lda,
                          "ldb": ldb, "ldc": ldc, "C": to_dev(np.full(m * ldc * L, np.nan))})
            host.append((L, m, n, k, lda, ldb, ldc, A, B))
        print(f"batch mix {method} trial {trial}: {len(probs)} problems", flush=True)
        mdg.gemm_batched(probs, method=method, persistent=0, spread_bits=16)
        torch_cuda.cuda.synchronize()
        loop = [to_host(p["C"]).copy() for p in probs]
        for rep in range(2):
            for p in probs:
                p["C"].fill_(np.nan)
            mdg.gemm_batched(probs, method=method, persistent=1, spread_bits=16)
            torch_cuda.cuda.synchronize()
            for i, (p, C) in enumerate(zip(probs, loop)):
                assert_bits_equal(to_host(p["C"]), C, f"{method} trial {trial} rep {rep} problem {i}")
        for (L, m, n, k, lda, ldb, ldc, A, B), C in zip(host, loop):
            A3 = A.reshape(m, lda, L)[:, :k].ravel()
            B3 = B.reshape(k, ldb, L)[:, :n].ravel()
            C3 = C.reshape(m, ldc, L)[:, :n].ravel()
            if method == "vector":
                assert_bits_equal(C3, orc.gemm(L, m, n, k, A3, B3, mode=orc.TP_FMA), f"vector {L} {m}x{n}x{k}")
            else:
                err, _ = orc.normwise_error(L, m, n, k, A3, B3, C3, nsamp=32)
                assert err <= 4.0 * k * 2.0 ** (-53 * L), (method, L, m, n, k, err)
