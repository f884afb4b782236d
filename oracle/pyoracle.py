"""ctypes binding of the CPU oracle (md_oracle.c) and of the reference-built
harness (oracle/_ref/libmdref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
imports it.

Arrays are numpy float64 with interleaved limbs: entry (r, c) of an m x n
matrix of L-limb numbers lives at [(r * ld + c) * L + limb], limb 0 the most
significant (reference layout, C++/TCcuda.cu:130-131).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdoracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmdref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None
_ref = None

TP_DEKKER = 0  # the reference's two_prod (Veltkamp split, no FMA)
TP_FMA = 1     # two_prod through fma; equal to Dekker unless a product underflows

GEN_DD = 0     # hi ~ U[-1,1), lo = hi*2^-53*U[-.5,.5)
GEN_QD = 1     # lead s*2^e, e in [emin,emax], next = prev*2^-53*U[-.5,.5)
GEN_REF = 2    # the reference's random_pd layout (positive, 52-bit limbs)

# elementwise operation codes (the product library's mdg_op values)
OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_SQR, OP_SQRT = 0, 1, 2, 3, 4, 5
OP_MUL_D, OP_INC_D, OP_MUL_PWR2, OP_ABS, OP_NEG = 6, 7, 8, 9, 10
BINARY_OPS = (OP_ADD, OP_SUB, OP_MUL, OP_DIV)
DOUBLE_OPS = (OP_MUL_D, OP_INC_D, OP_MUL_PWR2)


def build(force=False):
    """Compile libmdoracle.so (and _ref when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def build_ref():
    """Compile oracle/_ref/libmdref.so from /root/reference (this container)."""
    if os.path.isdir("/root/reference/C++"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return os.path.exists(REF_PATH)


def _p(a):
    return a.ctypes.data_as(_dp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        L = _lib
        L.orc_generate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                   ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp]
        for name in ("orc_add",):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp, _dp]
        L.orc_mul.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp, _dp, ctypes.c_int]
        L.orc_inc_d.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp, _dp]
        L.orc_fast_renorm.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp]
        L.orc_renorm.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp]
        L.orc_two_prod.argtypes = [ctypes.c_long, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp, _dp, _dp,
                             ctypes.c_int]
        L.orc_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               _dp, ctypes.c_long, _dp, ctypes.c_long, _dp, ctypes.c_long,
                               ctypes.c_int, ctypes.c_int]
        L.orc_gemm_rows.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _dp, ctypes.c_long, _dp, ctypes.c_long,
                                    _dp, ctypes.c_long, ctypes.c_int, ctypes.c_int]
        L.orc_squeeze.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_int, _dp, _dp]
        L.orc_conv_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    _dp, _dp, _dp]
        L.orc_bigB2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_renormhost.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_split4pd.argtypes = [ctypes.c_long, ctypes.c_int, _dp, _dp]
        L.orc_split8pd.argtypes = [ctypes.c_long, ctypes.c_int, _dp, _dp]
        L.orc_mixsplit2.argtypes = [ctypes.c_long, _dp, _dp]
        L.orc_exact_error.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _dp, ctypes.c_long, _dp, ctypes.c_long, _dp,
                                      ctypes.c_long, ctypes.c_int, _ip, _ip, _dp, _dp,
                                      _dp, ctypes.c_int]
        L.orc_exact_entry_words.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long,
                                            _dp, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                            _u64p, _ip, _ip]
        L.orc_num_threads.restype = ctypes.c_int
    return _lib


def ref():
    """The reference-built harness, or None where it is not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = ctypes.CDLL(REF_PATH)
        R = _ref
        for name in ("ref_mul", "ref_add", "ref_inc_d"):
            getattr(R, name).argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp, _dp]
        R.ref_fast_renorm.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp]
        R.ref_renorm.argtypes = [ctypes.c_int, ctypes.c_long, _dp, _dp]
        R.ref_two_prod.argtypes = [ctypes.c_long, _dp, _dp, _dp, _dp]
        R.ref_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp, _dp, _dp]
        R.ref_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               _dp, ctypes.c_long, _dp, ctypes.c_long, _dp, ctypes.c_long]
        R.ref_squeeze.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_int, _dp, _dp]
        R.ref_split4pd.argtypes = [ctypes.c_long, ctypes.c_int, _dp, _dp]
        R.ref_split8pd.argtypes = [ctypes.c_long, ctypes.c_int, _dp, _dp]
        R.ref_mixsplit2.argtypes = [ctypes.c_long, _dp, _dp]
    return _ref


def _chk(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with status {rc}")


def generate(kind, L, rows, cols, seed, emin=0, emax=0):
    out = np.empty(rows * cols * L, dtype=np.float64)
    _chk(lib().orc_generate(kind, L, rows * cols, seed, emin, emax, _p(out)), "generate")
    return out


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def op_mul(L, a, b, mode=TP_DEKKER, use_ref=False):
    a, b = _c(a), _c(b)
    c = np.empty_like(a)
    cnt = a.size // L
    if use_ref:
        _chk(ref().ref_mul(L, cnt, _p(a), _p(b), _p(c)), "ref_mul")
    else:
        _chk(lib().orc_mul(L, cnt, _p(a), _p(b), _p(c), mode), "mul")
    return c


def op_add(L, a, b, use_ref=False):
    a, b = _c(a), _c(b)
    c = np.empty_like(a)
    fn = ref().ref_add if use_ref else lib().orc_add
    _chk(fn(L, a.size // L, _p(a), _p(b), _p(c)), "add")
    return c


def op_inc_d(L, a, y, use_ref=False):
    a, y = _c(a), _c(y)
    c = np.empty_like(a)
    fn = ref().ref_inc_d if use_ref else lib().orc_inc_d
    _chk(fn(L, a.size // L, _p(a), _p(y), _p(c)), "inc_d")
    return c


def op(code, L, a, b=None, mode=TP_DEKKER, use_ref=False):
    """Elementwise operation `code` (OP_*) on L-limb entries a (and b: L-limb
    entries for the binary ops, one double per entry for the _D/_PWR2 ops)."""
    a = _c(a)
    cnt = a.size // L
    c = np.empty_like(a)
    bp = None
    if code in BINARY_OPS or code in DOUBLE_OPS:
        b = _c(b)
        assert b.size == (a.size if code in BINARY_OPS else cnt)
        bp = _p(b)
    if use_ref:
        _chk(ref().ref_op(code, L, cnt, _p(a), bp, _p(c)), f"ref_op {code}")
    else:
        _chk(lib().orc_op(code, L, cnt, _p(a), bp, _p(c), mode), f"op {code}")
    return c


def op_fast_renorm(L, x, use_ref=False):
    x = _c(x)
    cnt = x.size // (L + 1)
    r = np.full(cnt * L, np.nan)
    fn = ref().ref_fast_renorm if use_ref else lib().orc_fast_renorm
    _chk(fn(L, cnt, _p(x), _p(r)), "fast_renorm")
    return r


def op_renorm(L, x, use_ref=False):
    x = _c(x)
    cnt = x.size // (L + 1)
    r = np.full(cnt * L, np.nan)
    fn = ref().ref_renorm if use_ref else lib().orc_renorm
    _chk(fn(L, cnt, _p(x), _p(r)), "renorm")
    return r


def gemm(L, m, n, k, A, B, mode=TP_DEKKER, nthreads=1, use_ref=False, lda=None,
         ldb=None, ldc=None):
    A, B = _c(A), _c(B)
    lda = k if lda is None else lda
    ldb = n if ldb is None else ldb
    ldc = n if ldc is None else ldc
    C = np.zeros(m * ldc * L)
    if use_ref:
        _chk(ref().ref_gemm(L, m, n, k, _p(A), lda, _p(B), ldb, _p(C), ldc), "ref_gemm")
    else:
        _chk(lib().orc_gemm(L, m, n, k, _p(A), lda, _p(B), ldb, _p(C), ldc, mode, nthreads),
             "gemm")
    return C


def gemm_rows(L, r0, r1, n, k, A, B, mode=TP_DEKKER, nthreads=1):
    """Rows [r0, r1) of the host GEMM (the CPU-baseline sample)."""
    A, B = _c(A), _c(B)
    C = np.zeros((r1 - r0) * n * L)
    Asub = A[r0 * k * L:r1 * k * L]
    _chk(lib().orc_gemm_rows(L, 0, r1 - r0, n, k, _p(Asub), k, _p(B), n, _p(C), n, mode,
                             nthreads), "gemm_rows")
    return C


def gemm_rows_ref(L, r0, r1, n, k, A, B):
    """Rows [r0, r1) of the host GEMM through the reference-built library
    (oracle/_ref), or None where it is not built."""
    R = ref()
    if R is None:
        return None
    A, B = _c(A), _c(B)
    Asub = np.ascontiguousarray(A[r0 * k * L:r1 * k * L])
    C = np.zeros((r1 - r0) * n * L)
    _chk(R.ref_gemm(L, r1 - r0, n, k, _p(Asub), k, _p(B), n, _p(C), n), "ref_gemm rows")
    return C


def squeeze(L, pp, x, use_ref=False):
    x = _c(x)
    nent = x.size // pp
    y = np.empty(nent * L)
    fn = ref().ref_squeeze if use_ref else lib().orc_squeeze
    _chk(fn(L, nent, pp, _p(x), _p(y)), "squeeze")
    return y


def conv_gemm(m, n, k, pp, Aq, Bq):
    Aq, Bq = _c(Aq), _c(Bq)
    C = np.empty(m * n * pp)
    _chk(lib().orc_conv_gemm(m, n, k, pp, _p(Aq), _p(Bq), _p(C)), "conv_gemm")
    return C


def matmulhost(p, n, A, B):
    """The reference's matmulhost (C++/TCcuda.cu:304-317): orc_conv_gemm with
    m = n = k, the same loop order."""
    return conv_gemm(n, n, n, p, A, B)


def renormhost(A, n, p, big=False):
    """The reference's renormhost / renormhostbig (C++/TCcuda.cu:374-398) on
    a copy of A."""
    A = _c(A).copy()
    _chk(lib().orc_renormhost(_p(A), n, p, 1 if big else 0), "renormhost")
    return A


def bigB2(n, pp, Bq):
    Bq = _c(Bq)
    out = np.empty((n * pp) ** 2)
    _chk(lib().orc_bigB2(n, pp, _p(Bq), _p(out)), "bigB2")
    return out


def split4pd(x, p, use_ref=False):
    x = _c(x)
    out = np.empty(4 * x.size)
    fn = ref().ref_split4pd if use_ref else lib().orc_split4pd
    _chk(fn(x.size, p, _p(x), _p(out)), "split4pd")
    return out


def split8pd(x, p, use_ref=False):
    x = _c(x)
    out = np.empty(8 * x.size)
    fn = ref().ref_split8pd if use_ref else lib().orc_split8pd
    _chk(fn(x.size, p, _p(x), _p(out)), "split8pd")
    return out


def mixsplit2(x, use_ref=False):
    x = _c(x)
    out = np.empty(6 * x.size)
    fn = ref().ref_mixsplit2 if use_ref else lib().orc_mixsplit2
    _chk(fn(x.size, _p(x), _p(out)), "mixsplit2")
    return out


def exact_error(L, m, n, k, A, B, C, rows, cols, nthreads=None, lda=None, ldb=None,
                ldc=None):
    """Per sampled entry: (exact - C_hat, sum_k |a||b|, exact), all as doubles."""
    A, B, C = _c(A), _c(B), _c(C)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    ns = rows.size
    err, den, val = np.empty(ns), np.empty(ns), np.empty(ns)
    nt = nthreads or lib().orc_num_threads()
    rc = lib().orc_exact_error(L, m, n, k, _p(A), lda or k, _p(B), ldb or n, _p(C), ldc or n,
                               ns, rows.ctypes.data_as(_ip), cols.ctypes.data_as(_ip),
                               _p(err), _p(den), _p(val), nt)
    _chk(rc, "exact_error")
    return err, den, val


def normwise_error(L, m, n, k, A, B, C, nsamp=4096, seed=12345, nthreads=None):
    """||C_exact - C_hat||_F / || |A||B| ||_F over sampled entries (all entries
    when m*n <= nsamp), with the sample size."""
    if m * n <= nsamp:
        rr, cc = np.divmod(np.arange(m * n), n)
    else:
        rng = np.random.default_rng(seed)
        idx = rng.choice(m * n, size=nsamp, replace=False)
        rr, cc = np.divmod(idx, n)
    err, den, _ = exact_error(L, m, n, k, A, B, C, rr, cc, nthreads)
    return float(_norm2(err) / _norm2(den)), int(rr.size)


def _norm2(v):
    """2-norm without underflow of the squares (hexa-double errors are ~2^-800)."""
    s = float(np.max(np.abs(v))) if v.size else 0.0
    if s == 0.0:
        return 0.0
    return s * float(np.sqrt(np.sum((v / s) ** 2)))


def exact_entry(L, k, A, lda, B, ldb, r, c):
    """Exact value of entry (r, c) as a Python (numerator int, base exponent)."""
    A, B = _c(A), _c(B)
    words = np.zeros(80, dtype=np.uint64)
    base = ctypes.c_int()
    nw = ctypes.c_int()
    lib().orc_exact_entry_words(L, k, _p(A), lda, _p(B), ldb, r, c,
                                words.ctypes.data_as(_u64p), ctypes.byref(base),
                                ctypes.byref(nw))
    v = 0
    for i in range(nw.value - 1, -1, -1):
        v = (v << 64) | int(words[i])
    if v >> (64 * nw.value - 1):
        v -= 1 << (64 * nw.value)
    return v, base.value


def num_threads():
    return lib().orc_num_threads()
