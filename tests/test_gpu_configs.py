"""The five workload configurations at their full sizes, checked through
size-independent properties: the exact checker's normwise error on sampled
entries (<= c k 2^(-53 L), c = 4), bit-identity of the vector GEMM with the
oracle on sampled rows, agreement between paths, and batched == loop.
Timings are printed (pytest -s) for the record."""
import time

import numpy as np
import pytest

from conftest import assert_bits_equal, to_host

pytestmark = pytest.mark.gpu


def _bound(L, k):
    return 4.0 * k * 2.0 ** (-53 * L)


def _timed(torch, fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, (time.perf_counter() - t0) * 1e3


def _err(orc, L, n, k, A, B, C, nsamp, seed=3):
    return orc.normwise_error(L, n, n, k, A, B, C, nsamp=nsamp, seed=seed)[0]


def test_config1_dd_n128(mdg, orc, torch_cuda):
    # double-double n = 128, hi ~ U[-1,1], lo = hi 2^-53 U[-.5,.5] renormalised
    L, n = 2, 128
    A = mdg.fill_random(mdg.GEN_DD, L, n, n, 101)
    B = mdg.fill_random(mdg.GEN_DD, L, n, n, 102)
    Ah, Bh = to_host(A), to_host(B)
    assert_bits_equal(Ah, orc.generate(orc.GEN_DD, L, n, n, 101), "generator")
    want = orc.gemm(L, n, n, n, Ah, Bh, mode=orc.TP_DEKKER)   # the reference's CPU path
    got = to_host(mdg.vector_gemm(L, n, n, n, A, B, twoprod=mdg.TWOPROD_DEKKER))
    assert_bits_equal(got, want, "dd vector GEMM vs host GEMM")
    assert_bits_equal(to_host(mdg.vector_gemm(L, n, n, n, A, B)), want, "fma two_prod")
    for method in ("int8", "f64", "bf16", "int8_residue", "f64_residue"):
        C, info = mdg.gemm(L, n, n, n, A, B, method=method)
        e = _err(orc, L, n, n, Ah, Bh, to_host(C), n * n)
        print(f"cfg1 dd n=128 {method}: err {e:.3e} S={info['slices']}")
        assert e <= _bound(L, n)


def test_config2_qd_n1024_f64_split_vs_vector(mdg, orc, torch_cuda):
    L, n = 4, 1024
    A = mdg.fill_random(mdg.GEN_QD, L, n, n, 201, -8, 8)
    B = mdg.fill_random(mdg.GEN_QD, L, n, n, 202, -8, 8)
    Ah, Bh = to_host(A), to_host(B)
    Cv, tv = _timed(torch_cuda, lambda: mdg.vector_gemm(L, n, n, n, A, B))
    (Cs, info), ts = _timed(torch_cuda, lambda: mdg.gemm(L, n, n, n, A, B, method="f64"))
    (Cr, _), tr = _timed(torch_cuda, lambda: mdg.gemm(L, n, n, n, A, B, method="int8_residue"))
    (Cf, _), tf = _timed(torch_cuda, lambda: mdg.gemm(L, n, n, n, A, B, method="f64_residue"))
    Cv, Cs, Cr, Cf = to_host(Cv), to_host(Cs), to_host(Cr), to_host(Cf)
    # vector GEMM bit-identical to the host loops on 3 sampled rows
    for r in (0, 517, 1023):
        want = orc.gemm_rows(L, r, r + 1, n, n, Ah, Bh, mode=orc.TP_FMA, nthreads=8)
        assert_bits_equal(Cv[r * n * L:(r + 1) * n * L], want, f"row {r}")
    ev = _err(orc, L, n, n, Ah, Bh, Cv, 1024)
    es = _err(orc, L, n, n, Ah, Bh, Cs, 1024)
    er = _err(orc, L, n, n, Ah, Bh, Cr, 1024)
    print(f"cfg2 qd n=1024: vector {tv:.1f} ms err {ev:.3e}; f64 split {ts:.1f} ms err {es:.3e} "
          f"S={info['slices']} w={info['slice_bits']}; int8 residues {tr:.1f} ms err {er:.3e}; "
          f"f64 residues {tf:.1f} ms")
    assert ev <= _bound(L, n) and es <= _bound(L, n) and er <= _bound(L, n)
    # the split paths are at least as accurate as the reference arithmetic
    assert es <= ev and er <= ev
    assert np.array_equal(Cr.view(np.uint64), Cf.view(np.uint64))   # same exact product


def test_config3_od_2048_int8_bf16_vs_f64(mdg, orc, torch_cuda):
    L, n = 8, 2048
    A = mdg.fill_random(mdg.GEN_QD, L, n, n, 301, -30, 30)
    B = mdg.fill_random(mdg.GEN_QD, L, n, n, 302, -30, 30)
    Ah, Bh = to_host(A), to_host(B)
    res = {}
    for method in ("f64", "int8", "bf16", "int8_residue", "f64_residue"):
        (C, info), t = _timed(torch_cuda, lambda: mdg.gemm(L, n, n, n, A, B, method=method))
        e = _err(orc, L, n, n, Ah, Bh, to_host(C), 4096)
        res[method] = to_host(C)
        print(f"cfg3 od n=2048 {method}: {t:.1f} ms err {e:.3e} S={info['slices']} "
              f"w={info['slice_bits']} P={info['pair_gemms']}")
        assert e <= _bound(L, n)
    # INT8 and BF16 slice identically and sum exactly: identical bits; so do
    # the two residue methods
    assert_bits_equal(res["int8"], res["bf16"], "int8 vs bf16")
    assert_bits_equal(res["int8_residue"], res["f64_residue"], "int8 vs f64 residues")


def test_config4_batched_1024_mixed(mdg, orc, torch_cuda):
    rng = np.random.default_rng(404)
    probs = []
    for i in range(1024):
        L = 2 if i < 512 else 4
        m, n, k = (int(v) for v in rng.integers(16, 513, 3))
        probs.append({"L": L, "m": m, "n": n, "k": k,
                      "A": mdg.fill_random(mdg.GEN_QD, L, m, k, 40000 + 2 * i, -8, 8),
                      "B": mdg.fill_random(mdg.GEN_QD, L, k, n, 40001 + 2 * i, -8, 8),
                      "C": mdg.empty(m * n * L)})
    eff = sum(2.0 * p["m"] * p["n"] * p["k"] for p in probs)
    _, t_loop = _timed(torch_cuda, lambda: mdg.gemm_batched(probs, method="int8", persistent=0,
                                                            spread_bits=16))
    # every problem against the exact checker, and every problem of the
    # persistent pipeline bit for bit against the loop (compared on the device)
    for i, p in enumerate(probs):
        L, m, n, k = p["L"], p["m"], p["n"], p["k"]
        e = orc.normwise_error(L, m, n, k, to_host(p["A"]), to_host(p["B"]), to_host(p["C"]), nsamp=64,
                               seed=i)[0]
        assert e <= _bound(L, k), (i, L, m, n, k, e)
    loop = [p["C"].clone() for p in probs]
    _, t_build = _timed(torch_cuda, lambda: mdg.gemm_batched(probs, method="int8", persistent=1,
                                                             spread_bits=16))
    _, t_replay = _timed(torch_cuda, lambda: mdg.gemm_batched(probs, method="int8", persistent=1,
                                                              spread_bits=16))
    print(f"cfg4 1024 problems ({eff:.3e} eff. flops): loop {t_loop:.1f} ms, graph build "
          f"{t_build:.1f} ms, graph replay {t_replay:.1f} ms ({eff / t_replay / 1e6:.1f} GFLOP/s)")
    for i, (p, C) in enumerate(zip(probs, loop)):
        assert torch_cuda.equal(p["C"].view(torch_cuda.int64), C.view(torch_cuda.int64)), f"problem {i}"
    del loop
    # the same batch through the INT8 residue method: fused pipeline against
    # its own per-problem loop
    _, r_loop = _timed(torch_cuda, lambda: mdg.gemm_batched(probs, method="int8_residue", persistent=0,
                                                            spread_bits=16))
    for i, p in enumerate(probs):
        L, m, n, k = p["L"], p["m"], p["n"], p["k"]
        e = orc.normwise_error(L, m, n, k, to_host(p["A"]), to_host(p["B"]), to_host(p["C"]), nsamp=64,
                               seed=i)[0]
        assert e <= _bound(L, k), (i, L, m, n, k, e)
    rloop = [p["C"].clone() for p in probs]
    _, r_build = _timed(torch_cuda, lambda: mdg.gemm_batched(probs, method="int8_residue", persistent=1,
                                                             spread_bits=16))
    _, r_replay = _timed(torch_cuda, lambda: mdg.gemm_batched(probs, method="int8_residue", persistent=1,
                                                              spread_bits=16))
    print(f"cfg4 INT8 residues: loop {r_loop:.1f} ms, graph build {r_build:.1f} ms, graph replay "
          f"{r_replay:.1f} ms ({eff / r_replay / 1e6:.1f} GFLOP/s)")
    for i, (p, C) in enumerate(zip(probs, rloop)):
        assert torch_cuda.equal(p["C"].view(torch_cuda.int64), C.view(torch_cuda.int64)), f"residue problem {i}"
    del rloop


def test_metric_config_qd_4096(mdg, orc, torch_cuda):
    # the benchmark's workload: quad-double m = n = k = 4096, leading-limb
    # exponents in [-8, 8], seeds 1 and 2 (bench.py).  The INT8 slice path on
    # its default engine (the band kernel) and on the 256 x 256 single-diagonal
    # kernel, the FP64-MFMA split and the vector GEMM, each against the exact
    # checker on 1024 sampled entries; the two INT8 engines bit for bit on
    # the whole matrix; two vector rows bit for bit against the oracle's
    # restatement of the reference's host loops
    L, n = 4, 4096
    A = mdg.fill_random(mdg.GEN_QD, L, n, n, 1, -8, 8)
    B = mdg.fill_random(mdg.GEN_QD, L, n, n, 2, -8, 8)
    Ah, Bh = to_host(A), to_host(B)
    out = {}
    for method in ("int8", "int8_tile256", "f64", "vector"):
        if method == "int8_tile256":
            mdg.set_int8_backend(4)
        try:
            (C, info), t = _timed(torch_cuda, lambda: mdg.gemm(L, n, n, n, A, B,
                                                               method="int8" if method == "int8_tile256" else method))
        finally:
            mdg.set_int8_backend(0)
        e = _err(orc, L, n, n, Ah, Bh, to_host(C), 1024, seed=17)
        print(f"metric qd n=4096 {method}: {t:.1f} ms err {e:.3e} backend {info['backend']}")
        assert e <= _bound(L, n), (method, e)
        out[method] = C
    assert torch_cuda.equal(out["int8"].view(torch_cuda.int64), out["int8_tile256"].view(torch_cuda.int64))
    Cv = to_host(out["vector"]).reshape(n, n, L)
    for r in (0, 2999):
        want = orc.gemm_rows(L, r, r + 1, n, n, Ah, Bh, mode=orc.TP_FMA, nthreads=orc.num_threads())
        assert_bits_equal(Cv[r].ravel(), want, f"vector row {r}")


def test_metric_config_qd_4096_invariants(mdg, orc, torch_cuda):
    # size-independent properties of the INT8 slice path at the benchmark's
    # full size, compared on the device: permuting the depth (A's columns and
    # B's rows alike) leaves C unchanged bit for bit (exact integer sums), and
    # permuting A's rows and B's columns permutes C; scaling A by 2^37 and B
    # by 2^-90 scales C exactly
    torch = torch_cuda
    L, n = 4, 4096
    A = mdg.fill_random(mdg.GEN_QD, L, n, n, 1, -8, 8)
    B = mdg.fill_random(mdg.GEN_QD, L, n, n, 2, -8, 8)
    C, _ = mdg.gemm(L, n, n, n, A, B, method="int8")
    g = torch.Generator(device="cpu").manual_seed(5)
    pk = torch.randperm(n, generator=g).cuda()
    pr = torch.randperm(n, generator=g).cuda()
    pc = torch.randperm(n, generator=g).cuda()
    A3, B3 = A.view(n, n, L), B.view(n, n, L)
    Ak = A3.index_select(1, pk).contiguous().view(-1)
    Bk = B3.index_select(0, pk).contiguous().view(-1)
    Ck, _ = mdg.gemm(L, n, n, n, Ak, Bk, method="int8")
    print("metric qd n=4096: depth permutation", flush=True)
    assert torch.equal(Ck.view(torch.int64), C.view(torch.int64))
    del Ak, Bk, Ck
    Ar = A3.index_select(0, pr).contiguous().view(-1)
    Bc = B3.index_select(1, pc).contiguous().view(-1)
    Crc, _ = mdg.gemm(L, n, n, n, Ar, Bc, method="int8")
    want = C.view(n, n, L).index_select(0, pr).index_select(1, pc).contiguous().view(-1)
    print("metric qd n=4096: row and column permutations", flush=True)
    assert torch.equal(Crc.view(torch.int64), want.view(torch.int64))
    del Ar, Bc, Crc, want
    Cs, _ = mdg.gemm(L, n, n, n, A * 2.0 ** 37, B * 2.0 ** -90, method="int8")
    print("metric qd n=4096: power-of-two scaling", flush=True)
    assert torch.equal(Cs.view(torch.int64), (C * 2.0 ** -53).view(torch.int64))


@pytest.fixture(scope="module")
def hd4096(mdg, torch_cuda):
    L, n = 16, 4096
    A = mdg.fill_random(mdg.GEN_QD, L, n, n, 501, -60, 60)
    B = mdg.fill_random(mdg.GEN_QD, L, n, n, 502, -60, 60)
    return L, n, A, B, to_host(A), to_host(B)


@pytest.mark.parametrize("method", ["int8", "f64", "int8_residue", "f64_residue"])
def test_config5_hd_4096_split(mdg, orc, torch_cuda, hd4096, method):
    L, n, A, B, Ah, Bh = hd4096
    (C, info), t = _timed(torch_cuda, lambda: mdg.gemm(L, n, n, n, A, B, method=method))
    Ch = to_host(C)
    del C
    e = _err(orc, L, n, n, Ah, Bh, Ch, 256)
    print(f"cfg5 hd n=4096 {method}: {t:.1f} ms ({2.0 * n ** 3 / t / 1e6:.1f} GFLOP/s eff) err {e:.3e} "
          f"S={info['slices']} w={info['slice_bits']} P={info['pair_gemms']}")
    assert e <= _bound(L, n)
    mdg.lib().mdg_release_workspace()


def test_config5_hd_4096_vector(mdg, orc, torch_cuda, hd4096):
    L, n, A, B, Ah, Bh = hd4096
    C, t = _timed(torch_cuda, lambda: mdg.vector_gemm(L, n, n, n, A, B))
    Ch = to_host(C)
    e = _err(orc, L, n, n, Ah, Bh, Ch, 256)
    print(f"cfg5 hd n=4096 vector: {t:.1f} ms ({2.0 * n ** 3 / t / 1e6:.1f} GFLOP/s eff) err {e:.3e}")
    assert e <= _bound(L, n)
    # 16 entries of row 1234 against the oracle's host loops
    r, c0, w = 1234, 2000, 16
    Arow = Ah[r * n * L:(r + 1) * n * L]
    Bsub = np.ascontiguousarray(Bh.reshape(n, n, L)[:, c0:c0 + w, :]).ravel()
    want = orc.gemm(L, 1, w, n, Arow, Bsub, mode=orc.TP_FMA)
    assert_bits_equal(Ch.reshape(n, n, L)[r, c0:c0 + w, :].ravel(), want, "hd row 1234")
