"""The CPU oracle, pinned: against the reference's own code (golden vectors
from tests/golden/gen_golden.py, and the reference-built oracle/_ref where it
exists), against exact rational arithmetic, and against closed forms."""
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import assert_bits_equal, assert_bits_equal_nan

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")
LS = (2, 4, 8, 16)


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ------------------------------------------------ golden vectors (reference) --
@pytest.mark.parametrize("L", LS)
def test_ops_match_reference_vectors(orc, gold, L):
    a, b, y = gold[f"L{L}_a"], gold[f"L{L}_b"], gold[f"L{L}_y"]
    assert_bits_equal(orc.op_mul(L, a, b), gold[f"L{L}_mul"], f"mul L={L}")
    assert_bits_equal(orc.op_add(L, a, b), gold[f"L{L}_add"], f"add L={L}")
    assert_bits_equal(orc.op_inc_d(L, a, y), gold[f"L{L}_inc_d"], f"inc_d L={L}")
    if L > 2:
        f = gold[f"L{L}_f"]
        assert_bits_equal(orc.op_fast_renorm(L, f), gold[f"L{L}_fast_renorm"], "fast_renorm")
        assert_bits_equal(orc.op_renorm(L, f), gold[f"L{L}_renorm"], "renorm")


@pytest.mark.parametrize("L", LS)
@pytest.mark.parametrize("tag", ["qd", "ref"])
def test_gemm_matches_reference_vectors(orc, gold, L, tag):
    A, B = gold[f"L{L}_gemm_{tag}_A"], gold[f"L{L}_gemm_{tag}_B"]
    C = orc.gemm(L, 5, 7, 9, A, B)
    assert_bits_equal(C, gold[f"L{L}_gemm_{tag}_C"], f"gemm L={L} {tag}")
    # the fma two_prod agrees with Dekker's away from underflow
    assert_bits_equal(orc.gemm(L, 5, 7, 9, A, B, mode=orc.TP_FMA), C, "fma gemm")


@pytest.mark.parametrize("L", LS)
def test_split_and_squeeze_match_reference_vectors(orc, gold, L):
    for tag in ("ref", "qd"):
        x = gold[f"L{L}_split_{tag}_x"]
        assert_bits_equal(orc.split4pd(x, L), gold[f"L{L}_split4pd_{tag}"], f"split4pd {tag}")
        assert_bits_equal(orc.split8pd(x, L), gold[f"L{L}_split8pd_{tag}"], f"split8pd {tag}")
    assert_bits_equal(orc.squeeze(L, 4 * L, gold[f"L{L}_squeeze_x"]), gold[f"L{L}_squeeze_y"],
                      "squeeze")


def test_mixsplit2_matches_reference_vectors(orc, gold):
    for tag in ("ref", "qd"):
        assert_bits_equal(orc.mixsplit2(gold[f"mix_{tag}_x"]), gold[f"mix_{tag}"], tag)


# --------------------------------------- reference-built library (fresh data) --
def _ref_or_skip(orc):
    if orc.ref() is None and not orc.build_ref():
        pytest.skip("oracle/_ref not built (needs /root/reference)")


@pytest.mark.parametrize("L", LS)
def test_oracle_matches_reference_build_fresh(orc, L):
    _ref_or_skip(orc)
    rng = np.random.default_rng(L)
    a = orc.generate(orc.GEN_QD, L, 3000, 1, 11 * L, -30, 30)
    b = orc.generate(orc.GEN_QD, L, 3000, 1, 13 * L, -30, 30)
    a[rng.random(a.size) < 0.1] = 0.0
    assert_bits_equal(orc.op_mul(L, a, b), orc.op_mul(L, a, b, use_ref=True), "mul")
    assert_bits_equal(orc.op_add(L, a, b), orc.op_add(L, a, b, use_ref=True), "add")
    n = 10
    A = orc.generate(orc.GEN_QD, L, n, n, 3, -4, 4)
    B = orc.generate(orc.GEN_QD, L, n, n, 4, -4, 4)
    assert_bits_equal(orc.gemm(L, n, n, n, A, B), orc.gemm(L, n, n, n, A, B, use_ref=True), "gemm")


# ------------------------------------------------- exact checker, pinned ------
def _frac_entry(L, k, A, B, n, r, c):
    s = Fraction(0)
    for kk in range(k):
        a = sum(Fraction(float(v)) for v in A[(r * k + kk) * L:(r * k + kk + 1) * L])
        b = sum(Fraction(float(v)) for v in B[(kk * n + c) * L:(kk * n + c + 1) * L])
        s += a * b
    return s


@pytest.mark.parametrize("L", LS)
def test_exact_checker_equals_rational_arithmetic(orc, L):
    m, n, k = 3, 4, 6
    A = orc.generate(orc.GEN_QD, L, m, k, 21, -40, 40)
    B = orc.generate(orc.GEN_QD, L, k, n, 22, -40, 40)
    for r in range(m):
        for c in range(n):
            v, base = orc.exact_entry(L, k, A, k, B, n, r, c)
            assert Fraction(v) * Fraction(2) ** base == _frac_entry(L, k, A, B, n, r, c)


def test_exact_checker_agrees_with_mpmath_1100_bits(orc):
    # the normwise error is specified against an MPFR-class oracle at >= 1100
    # bits; the checker is exact, so mpmath at 1100 bits must agree with it to
    # that precision, and the per-entry error (exact - C) must round the same
    mpmath = pytest.importorskip("mpmath")
    L, m, n, k = 16, 3, 4, 7
    A = orc.generate(orc.GEN_QD, L, m, k, 71, -60, 60)
    B = orc.generate(orc.GEN_QD, L, k, n, 72, -60, 60)
    C = orc.gemm(L, m, n, k, A, B)
    rr, cc = np.divmod(np.arange(m * n), n)
    err, _, _ = orc.exact_error(L, m, n, k, A, B, C, rr, cc)
    with mpmath.workprec(1100):
        for e in range(m * n):
            r, c = int(rr[e]), int(cc[e])
            s = mpmath.mpf(0)
            for kk in range(k):
                a = mpmath.fsum(mpmath.mpf(float(v)) for v in A[(r * k + kk) * L:(r * k + kk + 1) * L])
                b = mpmath.fsum(mpmath.mpf(float(v)) for v in B[(kk * n + c) * L:(kk * n + c + 1) * L])
                s += a * b
            v, base = orc.exact_entry(L, k, A, k, B, n, r, c)
            exact = mpmath.mpf(v) * mpmath.mpf(2) ** base
            assert abs(exact - s) <= abs(exact) * mpmath.mpf(2) ** -1090
            chat = mpmath.fsum(mpmath.mpf(float(x)) for x in C[e * L:(e + 1) * L])
            assert float(s - chat) == err[e]


def test_exact_checker_error_of_exact_result_is_zero(orc):
    # small integers: every product and sum is exact in one double
    rng = np.random.default_rng(5)
    m, n, k, L = 4, 5, 6, 4
    A = np.zeros(m * k * L)
    B = np.zeros(k * n * L)
    A[::L] = rng.integers(-9, 10, m * k)
    B[::L] = rng.integers(-9, 10, k * n)
    Cint = (A[::L].reshape(m, k) @ B[::L].reshape(k, n)).ravel()
    C = np.zeros(m * n * L)
    C[::L] = Cint
    rr, cc = np.divmod(np.arange(m * n), n)
    err, den, val = orc.exact_error(L, m, n, k, A, B, C, rr, cc)
    assert np.all(err == 0.0) and np.array_equal(val, Cint)
    # a planted mistake in one limb shows up exactly
    C2 = C.copy()
    C2[(2 * n + 3) * L + 1] = 2.0 ** -60
    err2, _, _ = orc.exact_error(L, m, n, k, A, B, C2, rr, cc)
    assert err2[2 * n + 3] == -(2.0 ** -60) and np.count_nonzero(err2) == 1


# ------------------------------------------------------------ closed forms --
@pytest.mark.parametrize("L", LS)
def test_gemm_identity_is_exact(orc, L):
    # A*I has exactly A's value (the reference generator's limbs overlap, so
    # the first pass renormalises them), and a renormalised A is a fixed point
    n = 6
    A = orc.generate(orc.GEN_REF, L, n, n, 31, 0, 0)
    ident = np.zeros(n * n * L)
    ident[(np.arange(n) * n + np.arange(n)) * L] = 1.0
    AI = orc.gemm(L, n, n, n, A, ident)
    for e in range(n * n):
        assert (sum(Fraction(float(v)) for v in AI[e * L:(e + 1) * L])
                == sum(Fraction(float(v)) for v in A[e * L:(e + 1) * L]))
    assert_bits_equal(orc.gemm(L, n, n, n, AI, ident), AI, "(A*I)*I")
    assert_bits_equal(orc.gemm(L, n, n, n, ident, AI), AI, "I*(A*I)")


@pytest.mark.parametrize("L", LS)
def test_gemm_error_within_bound(orc, L):
    m, n, k = 6, 5, 40
    A = orc.generate(orc.GEN_QD, L, m, k, 41, -8, 8)
    B = orc.generate(orc.GEN_QD, L, k, n, 42, -8, 8)
    C = orc.gemm(L, m, n, k, A, B)
    err, _ = orc.normwise_error(L, m, n, k, A, B, C)
    assert err <= 8 * k * 2.0 ** (-53 * L + (2 if L == 2 else 6))


def test_split_parts_sum_to_stored_value(orc):
    # positive inputs of the reference generator: the parts add back exactly
    for L in (2, 4):
        x = orc.generate(orc.GEN_REF, L, 50, 1, 51, -3, 3)
        for parts, per in ((orc.split4pd(x, L), 4), (orc.split8pd(x, L), 8)):
            per_entry = parts.reshape(-1, per * L)
            for i, v in enumerate(x.reshape(-1, L)):
                want = sum(Fraction(float(t)) for t in v)
                got = sum(Fraction(float(t)) for t in per_entry[i])
                assert got == want


def test_normwise_error_does_not_underflow(orc):
    # hexa-double-sized errors (2^-800 absolute) must not vanish when squared
    L, m, n, k = 16, 2, 2, 3
    A = orc.generate(orc.GEN_QD, L, m, k, 61, 0, 0)
    B = orc.generate(orc.GEN_QD, L, k, n, 62, 0, 0)
    C = orc.gemm(L, m, n, k, A, B)
    C2 = C.copy()
    C2[L - 1] += 2.0 ** -800
    e1, _ = orc.normwise_error(L, m, n, k, A, B, C2)
    assert e1 > 0.0


# ------------------------------------- scalar operations off the GEMM path --
def _scalar_cases(orc, gold, L):
    a, b, y = gold[f"L{L}_a"], gold[f"L{L}_b"], gold[f"L{L}_y"]
    sq, pw = gold[f"L{L}_sqrt_x"], gold[f"L{L}_pw"]
    return {"sub": (orc.OP_SUB, a, b), "div": (orc.OP_DIV, a, b), "sqr": (orc.OP_SQR, a, None),
            "sqrt": (orc.OP_SQRT, sq, None), "sqrt_signed": (orc.OP_SQRT, a, None),
            "mul_d": (orc.OP_MUL_D, a, y), "mul_pwr2": (orc.OP_MUL_PWR2, a, pw),
            "neg": (orc.OP_NEG, a, None)}


@pytest.mark.parametrize("L", LS)
def test_scalar_ops_match_reference_vectors(orc, gold, L):
    for name, (code, x, bb) in _scalar_cases(orc, gold, L).items():
        assert_bits_equal_nan(orc.op(code, L, x, bb), gold[f"L{L}_op_{name}"], f"{name} L={L}")
    a = gold[f"L{L}_a"]
    got = orc.op(orc.OP_ABS, L, a).reshape(-1, L)
    want = gold[f"L{L}_op_abs"].reshape(-1, L).copy()
    if L == 16:
        # the reference's hdf_abs writes limb 5 into limb 9
        # (C++/hexa_double_functions.cu:1479,1489); the oracle and the
        # library return |a| limb-wise
        a2 = a.reshape(-1, L)
        sgn = np.where(a2[:, 0] < 0.0, -1.0, 1.0)
        assert_bits_equal(want[:, 9], sgn * a2[:, 5], "reference's limb-9 copy")
        want[:, 9] = sgn * a2[:, 9]
    assert_bits_equal(got, want, f"abs L={L}")


@pytest.mark.parametrize("L", LS)
def test_scalar_ops_match_reference_build_fresh(orc, L):
    _ref_or_skip(orc)
    a = orc.generate(orc.GEN_QD, L, 2000, 1, 17 * L, -40, 40)
    b = orc.generate(orc.GEN_QD, L, 2000, 1, 19 * L, -40, 40)
    y = orc.generate(orc.GEN_QD, 1, 2000, 1, 23 * L, -40, 40)
    pos = (a.reshape(-1, L) * np.where(a.reshape(-1, L)[:, :1] < 0, -1.0, 1.0)).ravel()
    for code, x, bb in ((orc.OP_SUB, a, b), (orc.OP_DIV, a, b), (orc.OP_SQR, a, None),
                        (orc.OP_SQRT, pos, None), (orc.OP_MUL_D, a, y), (orc.OP_INC_D, a, y)):
        assert_bits_equal_nan(orc.op(code, L, x, bb), orc.op(code, L, x, bb, use_ref=True), f"op {code}")


def _val(v):
    return sum(Fraction(float(t)) for t in v)


@pytest.mark.parametrize("L", LS)
def test_scalar_ops_against_exact_rationals(orc, L):
    # relative error of every operation against exact rational arithmetic:
    # 2^(4 - 53 L) (the CAMPARY/QD algorithms lose at most a few bits), and
    # for sqrt the residual |c^2 - a| / a within 2^(12 - 53 L) (one Heron
    # step from a half-length root); a dropped term or a wrong sign fails by
    # dozens of orders of magnitude
    n = 40
    a = orc.generate(orc.GEN_QD, L, n, 1, 31 * L, -30, 30)
    b = orc.generate(orc.GEN_QD, L, n, 1, 37 * L, -30, 30)
    d = orc.generate(orc.GEN_QD, 1, n, 1, 41 * L, -30, 30)
    pos = (a.reshape(-1, L) * np.where(a.reshape(-1, L)[:, :1] < 0, -1.0, 1.0)).ravel()
    tol = Fraction(2) ** (4 - 53 * L)
    res = {name: orc.op(code, L, x, bb) for name, code, x, bb in (
        ("sub", orc.OP_SUB, a, b), ("div", orc.OP_DIV, a, b), ("sqr", orc.OP_SQR, a, None),
        ("mul_d", orc.OP_MUL_D, a, d), ("sqrt", orc.OP_SQRT, pos, None))}
    for i in range(n):
        sl = slice(i * L, (i + 1) * L)
        A, B, D = _val(a[sl]), _val(b[sl]), Fraction(float(d[i]))
        assert abs(_val(res["sub"][sl]) - (A - B)) <= tol * (abs(A) + abs(B))
        assert abs(_val(res["div"][sl]) - A / B) <= tol * abs(A / B)
        assert abs(_val(res["sqr"][sl]) - A * A) <= tol * A * A
        assert abs(_val(res["mul_d"][sl]) - A * D) <= tol * abs(A * D)
        P, R = _val(pos[sl]), _val(res["sqrt"][sl])
        assert R > 0 and abs(R * R - P) <= Fraction(2) ** (12 - 53 * L) * P


@pytest.mark.parametrize("L", LS)
def test_scalar_ops_closed_forms(orc, L):
    # small integers and exact quotients/roots come out exact
    k = np.arange(1, 41, dtype=np.float64)
    sq = np.zeros(k.size * L)
    sq[::L] = k * k
    root = orc.op(orc.OP_SQRT, L, sq)
    assert np.array_equal(root[::L], k) and not np.any(root.reshape(-1, L)[:, 1:])
    num = np.zeros(k.size * L)
    den = np.zeros(k.size * L)
    num[::L] = k * 7.0
    den[::L] = k
    q = orc.op(orc.OP_DIV, L, num, den)
    assert np.all(q[::L] == 7.0) and not np.any(q.reshape(-1, L)[:, 1:])
    a = orc.generate(orc.GEN_QD, L, 30, 1, 43 * L, -30, 30)
    one = orc.op(orc.OP_DIV, L, a, a).reshape(-1, L)
    assert np.all(one[:, 0] == 1.0) and not np.any(one[:, 1:])
    assert not np.any(orc.op(orc.OP_SUB, L, a, a))
    # x^2 via sqr and via mul(x, x): the same bits at L >= 4 (the reference's
    # sqr bodies are its mul bodies with b = a), not at L = 2
    if L > 2:
        assert_bits_equal(orc.op(orc.OP_SQR, L, a), orc.op_mul(L, a, a), "sqr == mul(a, a)")


def test_matmulhost_small_integers_exact(orc):
    # every product and partial sum an exact small integer: the convolution
    # C[r][c][i] = sum_k sum_{j<=i} A[r][k][j] B[k][c][i-j], whatever the order
    rng = np.random.default_rng(5)
    for p, n in ((1, 5), (4, 6), (8, 3)):
        A = rng.integers(-100, 101, size=(n, n, p))
        B = rng.integers(-100, 101, size=(n, n, p))
        want = np.zeros((n, n, p), dtype=np.int64)
        for i in range(p):
            for j in range(i + 1):
                want[:, :, i] += A[:, :, j] @ B[:, :, i - j]
        got = orc.matmulhost(p, n, A.astype(np.float64).ravel(), B.astype(np.float64).ravel())
        assert np.array_equal(got, want.ravel().astype(np.float64)), (p, n)


def test_matmulhost_summation_order(orc):
    # C++/TCcuda.cu:304-317 adds term by term, k ascending and j = 0 .. i
    # within each k.  Entry (0, 0), part 1 of a 2 x 2 product of 2-part
    # entries: terms (k, j) = (0, 0) 2^53, (0, 1) 1, (1, 0) -2^53, (1, 1) 1
    # give ((2^53 + 1) - 2^53) + 1 = 1 in that order (j outer would give 2)
    n, p = 2, 2
    A = np.zeros((n, n, p))
    B = np.zeros((n, n, p))
    A[0, 0] = [1.0, 1.0]
    B[0, 0] = [1.0, 2.0 ** 53]
    A[0, 1] = [1.0, 1.0]
    B[1, 0] = [1.0, -(2.0 ** 53)]
    C = orc.matmulhost(p, n, A.ravel(), B.ravel()).reshape(n, n, p)
    assert C[0, 0, 1] == 1.0
    # three terms of one k: (2^53 + 1) + 1 = 2^53 forwards, 2^53 + 2 backwards
    A1 = np.array([1.0, 1.0, 1.0])
    B1 = np.array([1.0, 1.0, 2.0 ** 53])
    assert orc.matmulhost(3, 1, A1, B1)[2] == 2.0 ** 53


def test_renormhost_closed_forms(orc):
    # exact sums: x_k + x_{k+1} and x_{k+1} - ... leave a_k + a_{k+1} in
    # place k and the last part unchanged; a part below the sum's rounding is
    # lost (the second value is hi - x_k, not the error of the sum)
    n, p = 3, 5
    rng = np.random.default_rng(9)
    A = rng.integers(-1000, 1001, size=(n, n, p)).astype(np.float64)
    want = A.copy()
    want[:, :, :-1] = A[:, :, :-1] + A[:, :, 1:]
    assert np.array_equal(orc.renormhost(A.ravel(), n, p), want.ravel())
    x = np.array([1.0, 2.0 ** -60, 0.5])
    assert np.array_equal(orc.renormhost(x, 1, 3), np.array([1.0, 0.5, 0.5]))
    # the bigA layout (renormhostbig) holds the same entries with part k of
    # row i at i n p + k n + j
    Ar = rng.standard_normal((n, n, p))
    big = orc.renormhost(Ar.transpose(0, 2, 1).ravel(), n, p, big=True)
    assert np.array_equal(big, orc.renormhost(Ar.ravel(), n, p).reshape(n, n, p).transpose(0, 2, 1).ravel())
