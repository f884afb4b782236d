"""mdg_gemm_noalloc: mdg_gemm with all workspace in the library's static
device array.  Checked through the C ABI against mdg_gemm bit for bit on two
calls in a row (the second smaller, so it runs over what the first left in
the shared workspace), against the exact checker, for its size limit and its
unsupported methods, and for leaving device memory untouched."""
import numpy as np
import pytest

from conftest import to_dev, to_host

pytestmark = pytest.mark.gpu


def _bound(L, k):
    return 4.0 * k * 2.0 ** (-53 * L)


@pytest.mark.parametrize("method", ["int8", "f64", "bf16", "vector"])
@pytest.mark.parametrize("L", [2, 4])
def test_noalloc_matches_gemm_on_two_calls(mdg, orc, torch_cuda, method, L):
    for rep, (m, n, k) in enumerate(((300, 260, 333), (96, 80, 120))):
        print(f"noalloc {method} L={L} {m}x{n}x{k}", flush=True)
        A = orc.generate(orc.GEN_QD, L, m, k, 11 + m + L, -12, 12)
        B = orc.generate(orc.GEN_QD, L, k, n, 12 + n + L, -12, 12)
        dA, dB = to_dev(A), to_dev(B)
        C1, i1 = mdg.gemm(L, m, n, k, dA, dB, method=method)
        C2, i2 = mdg.gemm_noalloc(L, m, n, k, dA, dB, method=method)
        assert np.array_equal(to_host(C1).view(np.uint64), to_host(C2).view(np.uint64)), (rep, i1, i2)
        assert i2["slices"] == i1["slices"] and i2["pair_gemms"] == i1["pair_gemms"]
        if method == "int8":
            assert i2["backend"] == 2, i2
        err, _ = orc.normwise_error(L, m, n, k, A, B, to_host(C2), nsamp=256)
        assert err <= _bound(L, k), (err, i2)


def test_noalloc_allocates_no_device_memory(mdg, orc, torch_cuda):
    L, m, n, k = 4, 256, 256, 256
    A = to_dev(orc.generate(orc.GEN_QD, L, m, k, 21, -8, 8))
    B = to_dev(orc.generate(orc.GEN_QD, L, k, n, 22, -8, 8))
    C = mdg.empty(m * n * L)
    mdg.gemm_noalloc(L, m, n, k, A, B, method="int8", C=C)      # loads the kernels
    torch_cuda.cuda.synchronize()
    free0, _ = torch_cuda.cuda.mem_get_info()
    for (mm, nn, kk) in ((256, 256, 256), (200, 130, 250), (64, 64, 64)):
        mdg.gemm_noalloc(L, mm, nn, kk, A, B, method="int8", C=C, lda=k, ldb=n, ldc=n)
    torch_cuda.cuda.synchronize()
    free1, _ = torch_cuda.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)


def test_noalloc_rejects_oversize_and_residue(mdg, orc, torch_cuda):
    cap = mdg.lib().mdg_noalloc_workspace_bytes()
    assert cap == 256 << 20
    L, n = 4, 2048   # 29 INT32 planes of 2048^2 alone exceed 256 MiB
    A = mdg.fill_random(mdg.GEN_QD, L, n, n, 31, -8, 8)
    C = mdg.empty(n * n * L)
    C.fill_(1.5)
    with pytest.raises(mdg.MdgError) as ei:
        mdg.gemm_noalloc(L, n, n, n, A, A, method="int8", C=C)
    assert ei.value.status == 4 and "static workspace" in str(ei.value), str(ei.value)
    assert bool((C == 1.5).all()), "C was written by a rejected call"
    for method in ("int8_residue", "f64_residue"):
        with pytest.raises(mdg.MdgError) as ei:
            mdg.gemm_noalloc(L, 64, 64, 64, A, A, method=method, lda=n, ldb=n)
        assert ei.value.status == 2
    # a rejected call leaves the workspace usable
    m = 128
    B = mdg.fill_random(mdg.GEN_QD, L, m, m, 32, -8, 8)
    C1, _ = mdg.gemm(L, m, m, m, B, B, method="int8")
    C2, _ = mdg.gemm_noalloc(L, m, m, m, B, B, method="int8")
    assert np.array_equal(to_host(C1).view(np.uint64), to_host(C2).view(np.uint64))


def test_noalloc_random_shapes_match_gemm(mdg, orc, torch_cuda):
    # a seeded sweep (limb counts, ragged shapes, exponent ranges, one depth
    # cut into K-chunks): the no-allocation call gives mdg_gemm's bits
    rng = np.random.default_rng(88)
    for case in range(12):
        L = int(rng.choice([2, 4, 8, 16]))
        m, n = int(rng.integers(1, 160)), int(rng.integers(1, 160))
        k = int(rng.integers(1, 2000)) if case else 20000
        if case == 0:   # the deep one: double-double, so its slices fit the workspace
            L, m, n = 2, min(m, 40), min(n, 40)
        e = int(rng.integers(0, 4 * L))
        A = orc.generate(orc.GEN_QD, L, m, k, 300 + case, -e, e)
        B = orc.generate(orc.GEN_QD, L, k, n, 400 + case, -e, e)
        dA, dB = to_dev(A), to_dev(B)
        for method in ("int8", "bf16", "f64", "vector"):
            print(f"noalloc sweep {case} {method} L={L} {m}x{n}x{k}", flush=True)
            C1, _ = mdg.gemm(L, m, n, k, dA, dB, method=method)
            C2, _ = mdg.gemm_noalloc(L, m, n, k, dA, dB, method=method)
            assert np.array_equal(to_host(C1).view(np.uint64), to_host(C2).view(np.uint64)), (case, method)
