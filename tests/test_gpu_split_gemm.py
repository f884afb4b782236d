"""Split-product GEMMs (FP64 / INT8 / BF16 MFMA slices, INT8 residues)
against the exact checker: normwise error || C_exact - C || / || |A||B| ||
within c*k*2^(-53L), and agreement with the vector GEMM (the reference's
arithmetic)."""
import numpy as np
import pytest

from conftest import assert_bits_equal, to_dev, to_host

pytestmark = pytest.mark.gpu

METHODS = ["f64", "int8", "int8_rocblas", "int8_tile256", "int8_hipblaslt", "bf16", "bf16_diag", "int8_residue",
           "f64_residue"]


def gemm(mdg, L, m, n, k, A, B, method, **kw):
    """int8 and bf16 run the default engine (the hand-written band kernels),
    int8_mfma forces it, int8_tile256 and bf16_diag force the hand-written
    single-diagonal kernels (256 x 256 INT8, INT64-plane BF16),
    int8_rocblas / int8_hipblaslt the vendor libraries (comparison
    engines); all give the same exact diagonal sums."""
    forced = {"int8_rocblas": (1, "int8"), "int8_mfma": (2, "int8"), "int8_hipblaslt": (3, "int8"),
              "int8_tile256": (4, "int8"), "bf16_diag": (4, "bf16")}.get(method)
    if forced:
        mdg.set_int8_backend(forced[0])
        try:
            return mdg.gemm(L, m, n, k, A, B, method=forced[1], **kw)
        finally:
            mdg.set_int8_backend(0)
    return mdg.gemm(L, m, n, k, A, B, method=method, **kw)


def _bound(L, k):
    # c * k * 2^(-53 L) with c = 4
    return 4.0 * k * 2.0 ** (-53 * L)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("L", [2, 4, 8, 16])
def test_split_gemm_error_bound(mdg, orc, torch_cuda, method, L):
    for (m, n, k, emin, emax) in ((37, 29, 53, -8, 8), (64, 128, 96, -30, 30), (5, 3, 200, -2, 2)):
        print(f"split {method} L={L} {m}x{n}x{k} e=[{emin},{emax}]", flush=True)
        A = orc.generate(orc.GEN_QD, L, m, k, 7 * L + m, emin, emax)
        B = orc.generate(orc.GEN_QD, L, k, n, 9 * L + n, emin, emax)
        C, info = gemm(mdg, L, m, n, k, to_dev(A), to_dev(B), method)
        C = to_host(C)
        err, _ = orc.normwise_error(L, m, n, k, A, B, C)
        assert err <= _bound(L, k), f"{method} L={L}: err={err:.3e} bound={_bound(L, k):.3e} {info}"
        S = info["slices"]
        if method == "int8_residue":
            # narrow moduli: one INT8 GEMM each, wide moduli: four
            assert info["pair_gemms"] == (S - info["diagonals"]) + 4 * info["diagonals"]
        elif method == "f64_residue":
            assert info["pair_gemms"] == S and info["slice_bits"] == 21
        else:
            assert S >= 1 and info["pair_gemms"] == S * (S + 1) // 2
        V = orc.gemm(L, m, n, k, A, B, mode=orc.TP_FMA)
        errv, _ = orc.normwise_error(L, m, n, k, A, B, V)
        assert errv <= _bound(L, k)


@pytest.mark.parametrize("method", METHODS)
def test_split_gemm_dd_config1_inputs(mdg, orc, torch_cuda, method):
    # config 1's generator (hi ~ U[-1,1], lo = hi 2^-53 U[-.5,.5]) at n = 128
    L, n = 2, 128
    A = orc.generate(orc.GEN_DD, L, n, n, 1)
    B = orc.generate(orc.GEN_DD, L, n, n, 2)
    C, info = gemm(mdg, L, n, n, n, to_dev(A), to_dev(B), method)
    err, _ = orc.normwise_error(L, n, n, n, A, B, to_host(C))
    assert err <= _bound(L, n), (err, info)


@pytest.mark.parametrize("method", METHODS)
def test_split_gemm_zero_rows_and_columns(mdg, orc, torch_cuda, method):
    L, m, n, k = 4, 20, 18, 33
    A = orc.generate(orc.GEN_QD, L, m, k, 3, -8, 8).reshape(m, k, L)
    B = orc.generate(orc.GEN_QD, L, k, n, 4, -8, 8).reshape(k, n, L)
    A[3] = 0.0
    B[:, 5] = 0.0
    C, _ = gemm(mdg, L, m, n, k, to_dev(A.ravel()), to_dev(B.ravel()), method)
    C = to_host(C).reshape(m, n, L)
    assert np.all(C[3] == 0.0) and np.all(C[:, 5] == 0.0)
    err, _ = orc.normwise_error(L, m, n, k, A.ravel(), B.ravel(), C.ravel())
    assert err <= _bound(L, k)


@pytest.mark.parametrize("method", METHODS)
def test_split_gemm_all_zero(mdg, torch_cuda, method):
    L, m, n, k = 2, 9, 7, 5
    z = to_dev(np.zeros(max(m, n) * k * L))
    C, _ = gemm(mdg, L, m, n, k, z, z, method, C=to_dev(np.ones(m * n * L)))
    assert np.all(to_host(C) == 0.0)


@pytest.mark.parametrize("method", METHODS)
def test_split_gemm_small_integers_exact(mdg, torch_cuda, method):
    # exactly representable products and sums: the result must be exact
    rng = np.random.default_rng(8)
    L, m, n, k = 4, 33, 17, 65
    A = np.zeros((m, k, L))
    B = np.zeros((k, n, L))
    A[:, :, 0] = rng.integers(-100, 101, (m, k))
    B[:, :, 0] = rng.integers(-100, 101, (k, n))
    C, _ = gemm(mdg, L, m, n, k, to_dev(A.ravel()), to_dev(B.ravel()), method)
    C = to_host(C).reshape(m, n, L)
    assert np.array_equal(C[:, :, 0], A[:, :, 0] @ B[:, :, 0])
    assert np.all(C[:, :, 1:] == 0.0)


def test_split_gemm_rejects_non_finite(mdg, torch_cuda):
    L, m, n, k = 2, 4, 4, 4
    A = np.ones(m * k * L)
    A[5] = np.nan
    with pytest.raises(mdg.MdgError) as ei:
        mdg.gemm(L, m, n, k, to_dev(A), to_dev(np.ones(k * n * L)), method="int8")
    assert ei.value.status == 5


@pytest.mark.parametrize("method", ["int8_rocblas", "int8_mfma", "int8_tile256"])
def test_int8_chunked_diagonals(mdg, orc, torch_cuda, method):
    # S * kp > 2^17: the deepest diagonals are cut into K-chunks whose INT32
    # planes are added in INT64 by the reconstruction
    L, m, n, k = 4, 64, 64, 4608
    A = orc.generate(orc.GEN_QD, L, m, k, 5, -8, 8)
    B = orc.generate(orc.GEN_QD, L, k, n, 6, -8, 8)
    C, info = gemm(mdg, L, m, n, k, to_dev(A), to_dev(B), method)
    assert info["flushed"] == 1, info
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C), nsamp=512)
    assert err <= _bound(L, k), (err, info)


@pytest.mark.parametrize("L", [2, 4])
def test_int8_engines_agree_bitwise(mdg, orc, torch_cuda, L):
    # every engine computes the same exact diagonal sums, so the reconstructed
    # results must be identical bit for bit; the default engine is the band
    # kernel (backend 2), with no library in the default dispatch
    m, n, k = 300, 260, 333
    A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 17, -20, 20))
    B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 18, -20, 20))
    C1, i1 = gemm(mdg, L, m, n, k, A, B, "int8_rocblas")
    C2, i2 = gemm(mdg, L, m, n, k, A, B, "int8_mfma")
    C4, i4 = gemm(mdg, L, m, n, k, A, B, "int8_hipblaslt")
    C5, i5 = gemm(mdg, L, m, n, k, A, B, "int8_tile256")
    assert i1["backend"] == 1 and i2["backend"] == 2 and i4["backend"] == 3 and i5["backend"] == 4
    for C in (C2, C4, C5):
        assert np.array_equal(to_host(C1).view(np.uint64), to_host(C).view(np.uint64))
    for _ in range(2):
        C3, i3 = gemm(mdg, L, m, n, k, A, B, "int8")
        assert i3["backend"] == 2
        assert np.array_equal(to_host(C3).view(np.uint64), to_host(C1).view(np.uint64))


def test_int8_engines_agree_multi_tile(mdg, orc, torch_cuda):
    # 3 x 3 tiles of 256 (5 x 5 of the band kernel's 128) with ragged last
    # rows and columns, and kp = 1152 (9 k-blocks per slice), bit for bit
    # across the hand-written kernels and rocBLAS
    L, m, n, k = 2, 600, 530, 1100
    A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 71, -12, 12))
    B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 72, -12, 12))
    C1, i1 = gemm(mdg, L, m, n, k, A, B, "int8_rocblas")
    C2, i2 = gemm(mdg, L, m, n, k, A, B, "int8_mfma")
    C3, i3 = gemm(mdg, L, m, n, k, A, B, "int8_tile256")
    assert i1["backend"] == 1 and i2["backend"] == 2 and i3["backend"] == 4
    assert np.array_equal(to_host(C1).view(np.uint64), to_host(C2).view(np.uint64))
    assert np.array_equal(to_host(C1).view(np.uint64), to_host(C3).view(np.uint64))


@pytest.mark.parametrize("L,m,n,k,emax", [
    (2, 1, 1, 1, 0),          # S = 5 slices: one band of 4 and a band of 1
    (4, 130, 3, 257, 8),      # band of 4 x 7 and a remainder band; one-row tiles
    (8, 129, 257, 128, 30),   # S not a multiple of 4, ragged tiles
    (16, 64, 72, 300, 60),    # ~110 slices: 28 bands
    (4, 24, 40, 40000, 4),    # kp = 40064: chunks of 3 slices, padding steps
    (4, 40, 36, 20000, 6),    # kp = 20096: chunks of 6 slices (BF16: moves on full steps, padding steps)
    (2, 20, 17, 70000, 2),    # chunks of one slice
])
def test_int8_band_kernel_shapes(mdg, orc, torch_cuda, L, m, n, k, emax):
    # the band kernel (default) against the 256 x 256 single-diagonal kernel
    # bit for bit, and against the exact checker
    print(f"band L={L} {m}x{n}x{k}", flush=True)
    A = orc.generate(orc.GEN_QD, L, m, k, 31 + m, -emax, emax)
    B = orc.generate(orc.GEN_QD, L, k, n, 32 + n, -emax, emax)
    dA, dB = to_dev(A), to_dev(B)
    C2, i2 = gemm(mdg, L, m, n, k, dA, dB, "int8")
    C4, i4 = gemm(mdg, L, m, n, k, dA, dB, "int8_tile256")
    C5, i5 = gemm(mdg, L, m, n, k, dA, dB, "bf16")
    assert i2["backend"] == 2 and i4["backend"] == 4, (i2, i4)
    assert np.array_equal(to_host(C2).view(np.uint64), to_host(C4).view(np.uint64))
    # the BF16 band kernel sums the same slice products exactly
    assert i5["slices"] == i2["slices"], (i5, i2)
    assert np.array_equal(to_host(C2).view(np.uint64), to_host(C5).view(np.uint64))
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C2), nsamp=256)
    assert err <= _bound(L, k), (err, i2)


@pytest.mark.parametrize("L,m,n,k", [(4, 150, 140, 160), (2, 40, 33, 70000)])
def test_band_kernel_every_leftover_band(mdg, orc, torch_cuda, L, m, n, k):
    # slice counts S with every S mod 4 (INT8 bands) and S mod 3 (BF16
    # bands): the leftover diagonals run as a band shifted below diagonal 0;
    # at k = 70000 (one slice per K-chunk) that band is itself cut into
    # chunks.  Band kernels against the 256 x 256 single-diagonal kernel,
    # bit for bit.
    A = orc.generate(orc.GEN_QD, L, m, k, 41 + k, -6, 6)
    B = orc.generate(orc.GEN_QD, L, k, n, 42 + k, -6, 6)
    dA, dB = to_dev(A), to_dev(B)
    seen = set()
    for tb in range(53 * L - 40, 53 * L + 16, 4):
        C2, i2 = gemm(mdg, L, m, n, k, dA, dB, "int8", target_bits=tb)
        S = i2["slices"]
        if S in seen:
            continue
        seen.add(S)
        print(f"L={L} {m}x{n}x{k} target {tb} S={S}", flush=True)
        C4, i4 = gemm(mdg, L, m, n, k, dA, dB, "int8_tile256", target_bits=tb)
        C5, i5 = gemm(mdg, L, m, n, k, dA, dB, "bf16", target_bits=tb)
        assert i2["backend"] == 2 and i4["backend"] == 4 and i4["slices"] == S == i5["slices"], (i2, i4, i5)
        assert np.array_equal(to_host(C2).view(np.uint64), to_host(C4).view(np.uint64)), S
        assert np.array_equal(to_host(C2).view(np.uint64), to_host(C5).view(np.uint64)), S
    assert {S % 4 for S in seen} == {0, 1, 2, 3} and {S % 3 for S in seen} == {0, 1, 2}, sorted(seen)


@pytest.mark.parametrize("method", ["vector"] + METHODS)
def test_empty_shapes(mdg, orc, torch_cuda, method):
    # k = 0: C = 0 in every limb; m = 0 or n = 0: nothing to compute and C
    # left as it was
    import torch
    L = 4
    A, B = to_dev(np.ones(4 * L)), to_dev(np.ones(4 * L))

    def run(m, n, k, C):
        print(f"empty {method} {m}x{n}x{k}", flush=True)
        if method == "vector":
            return mdg.vector_gemm(L, m, n, k, A, B, C=C)
        return gemm(mdg, L, m, n, k, A, B, method, C=C)[0]

    C = torch.full((3 * 5 * L,), 7.0, dtype=torch.float64, device="cuda")
    out = run(3, 5, 0, C)
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0
    C = torch.full((2 * 5 * L,), 7.0, dtype=torch.float64, device="cuda")
    run(0, 5, 4, C)
    run(2, 0, 4, C)
    torch.cuda.synchronize()
    assert bool(torch.all(C == 7.0))


@pytest.mark.parametrize("method", ["int8", "bf16", "int8_residue", "f64"])
def test_concurrent_calls_on_two_streams(mdg, orc, torch_cuda, method):
    # two host threads, each on a stream of its own (scratch is per stream),
    # three calls each at once: every result bit-identical to the same call
    # made alone
    import threading
    import torch
    L = 4
    shapes = [(150, 130, 300), (96, 200, 257)]
    data = []
    for i, (m, n, k) in enumerate(shapes):
        A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 71 + i, -8, 8))
        B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 81 + i, -8, 8))
        want, _ = mdg.gemm(L, m, n, k, A, B, method=method)
        data.append((m, n, k, A, B, to_host(want)))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [[], []]
    errs = []

    def work(i):
        try:
            m, n, k, A, B, _ = data[i]
            for _ in range(3):
                C = mdg.empty(m * n * L)
                mdg.gemm(L, m, n, k, A, B, method=method, C=C, stream=streams[i])
                outs[i].append(C)
        except Exception as e:   # reported by the main thread
            errs.append(e)

    print(f"concurrent {method}", flush=True)
    ts = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    torch.cuda.synchronize()
    assert not errs, errs
    for i in range(2):
        assert len(outs[i]) == 3
        for C in outs[i]:
            assert np.array_equal(to_host(C).view(np.uint64), data[i][5].view(np.uint64)), (method, i)


def _digits_127(L, rows, cols):
    # every entry 1/2 + sum_{t=1}^{6L-1} 127 * 2^(-8t-6), limb l holding the
    # terms t = 6l .. 6l+5 (at most 48 bits, exact): the row scale is 2^0,
    # so slice t >= 1 of every entry is the digit 127 and the top slice 32
    x = np.zeros(L)
    for t in range(1, 6 * L):
        x[t // 6] += 127.0 * 2.0 ** (-8 * t - 6)
    x[0] += 0.5
    return np.tile(x, rows * cols).reshape(rows * cols * L).copy()


@pytest.mark.parametrize("method", ["bf16", "int8"])
def test_band_kernel_largest_partial_sums(mdg, orc, torch_cuda, method):
    # every product of two slices is 127^2 and they all add up, so the BF16
    # kernel's FP32 sums grow as fast as they can between moves into INT32
    # (at most 13 steps of 64 products: 832 * 127^2 < 2^24, also where the
    # steps without a slice pair on every diagonal leave MFMAs out; a sum
    # held past 1040 products, 1041 * 127^2 odd and above 2^24, would
    # round); the result must match the INT8 path bit for bit and the exact
    # product
    L, m, n, k = 4, 64, 96, 4096
    A = _digits_127(L, m, k)
    B = _digits_127(L, k, n)
    dA, dB = to_dev(A), to_dev(B)
    C, info = gemm(mdg, L, m, n, k, dA, dB, method)
    C8, _ = gemm(mdg, L, m, n, k, dA, dB, "int8_tile256")
    assert np.array_equal(to_host(C).view(np.uint64), to_host(C8).view(np.uint64)), info
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C), nsamp=128)
    assert err <= _bound(L, k), (err, info)


def test_int8_deep_k_band_chunks(mdg, orc, torch_cuda):
    # one slice pair deeper than an INT32 plane allows (kp > 131071): the
    # band kernels cut the depth into K-chunks of whole k-blocks, the 256 x
    # 256 kernel cuts it at other depths (130944 products), and every
    # chunking must give the same exact sums
    L, m, n, k = 2, 8, 12, 131100
    A = orc.generate(orc.GEN_QD, L, m, k, 41, -2, 2)
    B = orc.generate(orc.GEN_QD, L, k, n, 42, -2, 2)
    dA, dB = to_dev(A), to_dev(B)
    C, info = gemm(mdg, L, m, n, k, dA, dB, "int8")
    assert info["backend"] == 2, info
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C), nsamp=96)
    assert err <= _bound(L, k), (err, info)
    C4, info4 = gemm(mdg, L, m, n, k, dA, dB, "int8_tile256")
    assert info4["backend"] == 4, info4
    assert np.array_equal(to_host(C).view(np.uint64), to_host(C4).view(np.uint64)), info4
    C2, info2 = gemm(mdg, L, m, n, k, dA, dB, "bf16")
    assert info2["backend"] == 2, info2
    assert np.array_equal(to_host(C).view(np.uint64), to_host(C2).view(np.uint64)), info2


def test_deep_k_band_chunks_quad_double(mdg, orc, torch_cuda):
    # a quad-double product three INT32 planes deep (k = 400000, 3125
    # k-blocks of 128): INT8 and BF16 band kernels against the 256 x 256
    # kernel, bit for bit, and against the exact product
    L, m, n, k = 4, 20, 16, 400000
    A = orc.generate(orc.GEN_QD, L, m, k, 51, -3, 3)
    B = orc.generate(orc.GEN_QD, L, k, n, 52, -3, 3)
    dA, dB = to_dev(A), to_dev(B)
    C, info = gemm(mdg, L, m, n, k, dA, dB, "int8")
    assert info["backend"] == 2, info
    C4, _ = gemm(mdg, L, m, n, k, dA, dB, "int8_tile256")
    assert np.array_equal(to_host(C).view(np.uint64), to_host(C4).view(np.uint64)), info
    C2, info2 = gemm(mdg, L, m, n, k, dA, dB, "bf16")
    assert info2["backend"] == 2, info2
    assert np.array_equal(to_host(C).view(np.uint64), to_host(C2).view(np.uint64)), info2
    C3, info3 = gemm(mdg, L, m, n, k, dA, dB, "bf16_diag")
    assert info3["backend"] == 4, info3
    assert np.array_equal(to_host(C).view(np.uint64), to_host(C3).view(np.uint64)), info3
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C), nsamp=24)
    assert err <= _bound(L, k), (err, info)


def test_bf16_many_band_items(mdg, orc, torch_cuda):
    # 107 slices at depth 34197: more BF16 band items than one launch's
    # table holds, so the band kernel runs in several launches; it, the
    # single-diagonal BF16 kernel and the INT8 path give the same bits
    L, m, n, k = 16, 24, 20, 34197
    A = orc.generate(orc.GEN_QD, L, m, k, 100, -46, 46)
    B = orc.generate(orc.GEN_QD, L, k, n, 200, -46, 46)
    dA, dB = to_dev(A), to_dev(B)
    C8, i8 = gemm(mdg, L, m, n, k, dA, dB, "int8", target_bits=824)
    C5, i5 = gemm(mdg, L, m, n, k, dA, dB, "bf16", target_bits=824)
    C6, i6 = gemm(mdg, L, m, n, k, dA, dB, "bf16_diag", target_bits=824)
    assert i5["slices"] == i8["slices"] > 100, (i5, i8)
    assert i5["backend"] == 2 and i6["backend"] == 4, (i5, i6)
    assert np.array_equal(to_host(C8).view(np.uint64), to_host(C5).view(np.uint64))
    assert np.array_equal(to_host(C8).view(np.uint64), to_host(C6).view(np.uint64))


def test_f64_residue_deep_flushes(mdg, orc, torch_cuda):
    # k > 8192: the FP64 partial sums are moved into INT64 more than once
    L, m, n, k = 4, 40, 36, 20000
    A = orc.generate(orc.GEN_QD, L, m, k, 25, -6, 6)
    B = orc.generate(orc.GEN_QD, L, k, n, 26, -6, 6)
    C, info = mdg.gemm(L, m, n, k, to_dev(A), to_dev(B), method="f64_residue")
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C), nsamp=128)
    assert err <= _bound(L, k), (err, info)


def test_residue_chunked_depth(mdg, orc, torch_cuda):
    # 2 kp > 130944: the depth-2k GEMMs of the wide moduli are cut into
    # K-chunks whose INT32 planes are added in INT64 before the reduction
    L, m, n, k = 4, 16, 24, 70001
    A = orc.generate(orc.GEN_QD, L, m, k, 15, -4, 4)
    B = orc.generate(orc.GEN_QD, L, k, n, 16, -4, 4)
    C, info = mdg.gemm(L, m, n, k, to_dev(A), to_dev(B), method="int8_residue")
    assert info["flushed"] == 1 and info["diagonals"] > 0, info
    err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C), nsamp=96)
    assert err <= _bound(L, k), (err, info)



@pytest.mark.parametrize("L,m,n,k", [(2, 200, 150, 300), (4, 200, 150, 300), (2, 600, 530, 1100),
                                     (4, 257, 3, 129), (8, 1, 300, 5)])
def test_residue_engines_agree_bitwise(mdg, orc, torch_cuda, L, m, n, k):
    # the residue products are exact integers on every engine: the default
    # hand-written kernel (backend 2; 256 x 256 tiles, ragged edges, several
    # tiles, one-row and one-column tiles) against rocBLAS and hipBLASLt
    print(f"residue engines L={L} {m}x{n}x{k}", flush=True)
    A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 27 + m, -20, 20))
    B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 28 + n, -20, 20))
    out = {}
    for engine, want in ((1, 1), (3, 3), (0, 2)):
        mdg.set_int8_backend(engine)
        try:
            C, info = mdg.gemm(L, m, n, k, A, B, method="int8_residue")
        finally:
            mdg.set_int8_backend(0)
        assert info["backend"] == want, (engine, info)
        out[engine] = to_host(C).view(np.uint64)
    assert np.array_equal(out[1], out[3]) and np.array_equal(out[1], out[0])


@pytest.mark.parametrize("L", [2, 4, 8, 16])
def test_int8_and_f64_residues_bit_identical(mdg, orc, torch_cuda, L):
    # both residue methods compute the same exact integer product of the same
    # fixed-point inputs (different moduli, different MFMA formats), so the
    # reconstructed results must agree bit for bit
    m, n, k = 70, 90, 110
    A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 91 + L, -25, 25))
    B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 92 + L, -25, 25))
    C8, i8 = mdg.gemm(L, m, n, k, A, B, method="int8_residue")
    C64, i64 = mdg.gemm(L, m, n, k, A, B, method="f64_residue")
    assert i8["total_bits"] == i64["total_bits"]
    assert np.array_equal(to_host(C8).view(np.uint64), to_host(C64).view(np.uint64))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("L", [2, 8])
def test_split_gemm_leading_dimensions(mdg, orc, torch_cuda, method, L):
    # operands inside wider row-major buffers (lda > k, ldb > n) and a result
    # with ldc > n: the same bits as the packed layout, padding untouched
    m, n, k = 45, 38, 71
    lda, ldb, ldc = k + 3, n + 5, n + 2
    A = orc.generate(orc.GEN_QD, L, m, k, 301 + L, -8, 8)
    B = orc.generate(orc.GEN_QD, L, k, n, 302 + L, -8, 8)
    Aw = np.full((m, lda, L), 7.0)
    Aw[:, :k] = A.reshape(m, k, L)
    Bw = np.full((k, ldb, L), -3.0)
    Bw[:, :n] = B.reshape(k, n, L)
    C0, _ = gemm(mdg, L, m, n, k, to_dev(A), to_dev(B), method)
    Cw = to_dev(np.full(m * ldc * L, 5.0))
    gemm(mdg, L, m, n, k, to_dev(Aw.ravel()), to_dev(Bw.ravel()), method, C=Cw, lda=lda, ldb=ldb, ldc=ldc)
    torch_cuda.cuda.synchronize()
    got = to_host(Cw).reshape(m, ldc, L)
    assert np.array_equal(got[:, :n].view(np.uint64), to_host(C0).reshape(m, n, L).view(np.uint64)), method
    assert np.all(got[:, n:] == 5.0), "padding columns of C were written"


def test_random_shapes_engines_agree(mdg, orc, torch_cuda):
    # a seeded sweep over limb counts, ragged shapes, exponent ranges,
    # depths (a few deep enough to cut the bands into K-chunks) and
    # precision targets: the INT8 band kernel, the 256 x 256 kernel and the
    # BF16 band kernel give the same bits, so do the two residue methods,
    # and the slice result is within the bound
    rng = np.random.default_rng(2026)
    for case in range(36):
        L = int(rng.choice([2, 4, 8, 16]))
        m, n = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        k = int(rng.integers(1, 2500)) if case % 6 else int(rng.integers(15000, 45000))
        if L == 16 and k > 2500:
            m, n = min(m, 64), min(n, 64)
        e = int(rng.integers(0, 8 * L))
        tb = int(rng.choice([0, 53 * L - 24, 53 * L + 16]))
        print(f"sweep {case}: L={L} {m}x{n}x{k} e=+-{e} target={tb}", flush=True)
        A = orc.generate(orc.GEN_QD, L, m, k, 100 + case, -e, e)
        B = orc.generate(orc.GEN_QD, L, k, n, 200 + case, -e, e)
        dA, dB = to_dev(A), to_dev(B)
        C2, i2 = gemm(mdg, L, m, n, k, dA, dB, "int8", target_bits=tb)
        C4, i4 = gemm(mdg, L, m, n, k, dA, dB, "int8_tile256", target_bits=tb)
        C5, i5 = gemm(mdg, L, m, n, k, dA, dB, "bf16", target_bits=tb)
        h2 = to_host(C2)
        assert np.array_equal(h2.view(np.uint64), to_host(C4).view(np.uint64)), (case, i2, i4)
        assert np.array_equal(h2.view(np.uint64), to_host(C5).view(np.uint64)), (case, i2, i5)
        R1, _ = gemm(mdg, L, m, n, k, dA, dB, "int8_residue", target_bits=tb)
        R2, _ = gemm(mdg, L, m, n, k, dA, dB, "f64_residue", target_bits=tb)
        assert np.array_equal(to_host(R1).view(np.uint64), to_host(R2).view(np.uint64)), case
        if tb == 0:
            err, _ = orc.normwise_error(L, m, n, k, A, B, h2, nsamp=32)
            assert err <= _bound(L, k), (case, err, i2)


@pytest.mark.parametrize("method", ["vector", "f64", "int8", "bf16", "int8_residue", "f64_residue"])
def test_power_of_two_scaling_is_exact(mdg, orc, torch_cuda, method):
    # A 2^a times B 2^b is (A B) 2^(a+b) bit for bit on every path: row and
    # column scales, slicing, the integer products and the reconstruction
    # all commute with powers of two, far from 1 in both directions (as long
    # as no limb leaves the normal range)
    for L, scales in ((2, ((400, -700), (-500, 480))), (4, ((400, -700), (-500, 480))),
                      (8, ((300, -500), (-400, 380))), (16, ((150, -180), (-120, 200)))):
        m, n, k = 37, 29, 45
        A = orc.generate(orc.GEN_QD, L, m, k, 60 + L, -8, 8)
        B = orc.generate(orc.GEN_QD, L, k, n, 70 + L, -8, 8)

        def run(X, Y):
            if method == "vector":
                return to_host(mdg.vector_gemm(L, m, n, k, to_dev(X), to_dev(Y)))
            return to_host(gemm(mdg, L, m, n, k, to_dev(X), to_dev(Y), method)[0])

        C = run(A, B)
        for a, b in scales:
            print(f"scaling {method} L={L} 2^{a} x 2^{b}", flush=True)
            Cs = run(np.ldexp(A, a), np.ldexp(B, b))
            assert_bits_equal(Cs, np.ldexp(C, a + b), f"{method} L={L} 2^{a} 2^{b}")


@pytest.mark.parametrize("method", ["f64", "int8", "bf16", "int8_residue", "f64_residue"])
def test_transposed_product_is_the_transpose(mdg, orc, torch_cuda, method):
    # B^T A^T = (A B)^T bit for bit: A's rows and B's columns get their scales
    # and slices (or residues) the same way, the exact integer sums commute,
    # and the reconstruction of entry (r, c) depends on E_r + F_c only
    for L in (2, 4, 8, 16):
        m, n, k = 41, 27, 66
        A = orc.generate(orc.GEN_QD, L, m, k, 80 + L, -20, 20)
        B = orc.generate(orc.GEN_QD, L, k, n, 90 + L, -20, 20)
        At = np.ascontiguousarray(A.reshape(m, k, L).transpose(1, 0, 2)).ravel()
        Bt = np.ascontiguousarray(B.reshape(k, n, L).transpose(1, 0, 2)).ravel()
        print(f"transpose {method} L={L}", flush=True)
        C, _ = gemm(mdg, L, m, n, k, to_dev(A), to_dev(B), method)
        Ct, _ = gemm(mdg, L, n, m, k, to_dev(Bt), to_dev(At), method)
        want = np.ascontiguousarray(to_host(C).reshape(m, n, L).transpose(1, 0, 2)).ravel()
        assert_bits_equal(to_host(Ct), want, f"{method} L={L}")


@pytest.mark.parametrize("method", ["f64", "int8", "bf16", "int8_residue", "f64_residue", "vector"])
def test_permutation_invariance(mdg, orc, torch_cuda, method):
    # permuting A's rows and B's columns permutes C's entries (every path);
    # permuting the depth (A's columns and B's rows alike) leaves C unchanged
    # on the split paths, whose depth sums are exact integers (the vector path
    # rounds as it goes, so its sums depend on the order)
    rng = np.random.default_rng(31)
    for L in (2, 4, 16):
        m, n, k = 33, 21, 150
        A = orc.generate(orc.GEN_QD, L, m, k, 110 + L, -12, 12).reshape(m, k, L)
        B = orc.generate(orc.GEN_QD, L, k, n, 120 + L, -12, 12).reshape(k, n, L)
        pr, pc, pk = rng.permutation(m), rng.permutation(n), rng.permutation(k)

        def run(X, Y):
            X, Y = to_dev(np.ascontiguousarray(X).ravel()), to_dev(np.ascontiguousarray(Y).ravel())
            if method == "vector":
                return to_host(mdg.vector_gemm(L, m, n, k, X, Y)).reshape(m, n, L)
            return to_host(gemm(mdg, L, m, n, k, X, Y, method)[0]).reshape(m, n, L)

        print(f"permutations {method} L={L}", flush=True)
        C = run(A, B)
        assert_bits_equal(run(A[pr], B[:, pc]).ravel(), C[pr][:, pc].ravel(), f"{method} L={L} rows/cols")
        if method != "vector":
            assert_bits_equal(run(A[:, pk], B[pk]).ravel(), C.ravel(), f"{method} L={L} depth")


@pytest.mark.parametrize("method", ["f64", "int8", "bf16", "int8_residue", "f64_residue"])
def test_extreme_exponent_ranges_within_the_bound(mdg, orc, torch_cuda, method):
    # leading limbs spread over 2^-150 .. 2^150 inside every row and column:
    # the measured spread (E - log2(mean |x0|) <= log2 k + 1) stays small, the
    # entries far below a row's scale fall under its fixed point, and the
    # normwise error against |A| |B| still meets the bound at every L
    for L in (2, 4, 8, 16):
        m, n, k = 12, 10, 30
        A = orc.generate(orc.GEN_QD, L, m, k, 5 + L, -150, 150)
        B = orc.generate(orc.GEN_QD, L, k, n, 6 + L, -150, 150)
        print(f"extreme range {method} L={L}", flush=True)
        C, info = gemm(mdg, L, m, n, k, to_dev(A), to_dev(B), method)
        err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C))
        assert err <= _bound(L, k), (method, L, err, info)


@pytest.mark.parametrize("method", ["f64", "int8", "bf16", "int8_residue", "f64_residue"])
def test_repeated_calls_are_bit_identical(mdg, orc, torch_cuda, method):
    # no step of a call depends on the order in which workgroups finish (the
    # column scan adds its row-chunk sums in a fixed order; the GEMMs' sums
    # are exact): three calls give the same bits and the same slice count
    L, m, n, k = 4, 70, 300, 3000
    A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 501, -20, 20))
    B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 502, -20, 20))
    runs = [gemm(mdg, L, m, n, k, A, B, method) for _ in range(3)]
    for C, info in runs[1:]:
        assert info["slices"] == runs[0][1]["slices"] and info["spread_b"] == runs[0][1]["spread_b"]
        assert np.array_equal(to_host(C).view(np.uint64), to_host(runs[0][0]).view(np.uint64))
