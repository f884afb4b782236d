"""ctypes binding of libmdgemm.so (include/mdgemm.h) for tests and benchmarks.

Device memory and streams come from torch (float64 CUDA tensors on ROCm);
every call goes through the C ABI.  There is no fallback: if the library is
not built, or no GPU is present, calls raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdgemm.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mdgemm.h")

METHOD_VECTOR, METHOD_F64_SPLIT, METHOD_INT8_OZAKI, METHOD_BF16_OZAKI = 0, 1, 2, 3
METHODS = {"vector": 0, "f64": 1, "int8": 2, "bf16": 3, "int8_residue": 4, "f64_residue": 5}
TWOPROD_FMA, TWOPROD_DEKKER = 0, 1
GEN_DD, GEN_QD, GEN_REF = 0, 1, 2
# elementwise operations (mdg_op)
OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqr": 4, "sqrt": 5, "mul_d": 6, "inc_d": 7,
       "mul_pwr2": 8, "abs": 9, "neg": 10}
BINARY_OPS = ("add", "sub", "mul", "div")
DOUBLE_OPS = ("mul_d", "inc_d", "mul_pwr2")

STATUS = {0: "MDG_SUCCESS", 1: "MDG_ERROR_INVALID_VALUE", 2: "MDG_ERROR_NOT_SUPPORTED",
          3: "MDG_ERROR_HIP", 4: "MDG_ERROR_OUT_OF_MEMORY", 5: "MDG_ERROR_NON_FINITE",
          6: "MDG_ERROR_RANGE"}


class MdgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class GemmInfo(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("slices", ctypes.c_int), ("slice_bits", ctypes.c_int),
                ("diagonals", ctypes.c_int), ("pair_gemms", ctypes.c_longlong),
                ("total_bits", ctypes.c_int), ("spread_a", ctypes.c_int),
                ("spread_b", ctypes.c_int), ("flushed", ctypes.c_int),
                ("stage_ms", ctypes.c_float * 4), ("backend", ctypes.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["stage_ms"] = list(self.stage_ms)
        return d


class Problem(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_longlong),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_longlong),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_longlong)]


_lib = None
_vp, _i, _ll, _ull = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_ulonglong

_SIGS = {
    "mdg_last_error": ([], ctypes.c_char_p),
    "mdg_version": ([], _i),
    "mdg_init": ([], _i),
    "mdg_release_workspace": ([], _i),
    "mdg_set_profiling": ([_i], _i),
    "mdg_set_int8_backend": ([_i], _i),
    "mdg_vector_gemm": ([_i, _i, _i, _i, _vp, _ll, _vp, _ll, _vp, _ll, _i, _vp], _i),
    "mdg_gemm": ([_i, _i, _i, _i, _vp, _ll, _vp, _ll, _vp, _ll, _i, _i, ctypes.POINTER(GemmInfo), _vp], _i),
    "mdg_gemm_batched": ([_i, ctypes.POINTER(Problem), _i, _i, _i, _i, _vp], _i),
    "mdg_gemm_noalloc": ([_i, _i, _i, _i, _vp, _ll, _vp, _ll, _vp, _ll, _i, _i, ctypes.POINTER(GemmInfo), _vp], _i),
    "mdg_noalloc_workspace_bytes": ([], ctypes.c_size_t),
    "mdg_batched_last_flags": ([_vp, _i, ctypes.POINTER(ctypes.c_int)], _i),
    "mdg_set_batch_plan_cache": ([_i, _ll], _i),
    "mdg_batched_last_gemm_ms": ([_vp, ctypes.POINTER(ctypes.c_float)], _i),
    "mdg_split4pd": ([_vp, _ll, _i, _vp, _vp], _i),
    "mdg_split8pd": ([_vp, _ll, _i, _vp, _vp], _i),
    "mdg_mixsplit2": ([_vp, _ll, _vp, _vp], _i),
    "mdg_bigB2": ([_vp, _i, _i, _vp, _vp], _i),
    "mdg_matmulhost": ([_i, _i, _vp, _vp, _vp, _vp], _i),
    "mdg_renormhost": ([_vp, _i, _i, _i, _vp], _i),
    "mdg_dgemm_acc": ([_i, _i, _i, _vp, _vp, _vp, _vp], _i),
    "mdg_conv_gemm": ([_i, _i, _i, _i, _vp, _vp, _vp, _vp], _i),
    "mdg_squeeze": ([_i, _ll, _i, _vp, _vp, _vp], _i),
    "mdg_tc_gemm": ([_i, _i, _i, _i, _vp, _vp, _vp, _i, _vp], _i),
    "mdg_fill_random": ([_i, _i, _ll, _ull, _i, _i, _vp, _vp], _i),
    "mdg_elementwise": ([_i, _i, _ll, _vp, _vp, _vp, _i, _vp], _i),
    "mdg_to_planar": ([_i, _i, _i, _vp, _ll, _vp, _vp], _i),
    "mdg_from_planar": ([_i, _i, _i, _vp, _vp, _ll, _vp], _i),
    "mdg_transpose": ([_i, _i, _i, _vp, _ll, _vp, _ll, _vp], _i),
}


def lib():
    """Load libmdgemm.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def declared_symbols():
    """Entry points declared in include/mdgemm.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mdg_\w+)\s*\(", txt, re.M)))


def _check(rc):
    if rc != 0:
        raise MdgError(rc, lib().mdg_last_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


# --------------------------------------------------------------- wrappers --
def set_profiling(on):
    _check(lib().mdg_set_profiling(1 if on else 0))


def set_int8_backend(backend):
    """Slice-plane and residue GEMM engine: 0 or 2 this library's MFMA
    kernels (default: the band kernels), 1 rocBLAS, 3 hipBLASLt (INT8 only),
    4 the single-diagonal slice kernels (256 x 256 INT8, INT64-plane BF16)."""
    _check(lib().mdg_set_int8_backend(int(backend)))


def empty(n, device="cuda"):
    import torch
    return torch.empty(int(n), dtype=torch.float64, device=device)


def fill_random(kind, L, rows, cols, seed, emin=0, emax=0, stream=None):
    out = empty(rows * cols * L)
    _check(lib().mdg_fill_random(kind, L, rows * cols, seed, emin, emax, _ptr(out), _stream(stream)))
    return out


def vector_gemm(L, m, n, k, A, B, twoprod=TWOPROD_FMA, C=None, lda=None, ldb=None, ldc=None,
                stream=None):
    ldc = n if ldc is None else ldc
    if C is None:
        C = empty(m * ldc * L)
    _check(lib().mdg_vector_gemm(L, m, n, k, _ptr(A), k if lda is None else lda, _ptr(B),
                                 n if ldb is None else ldb, _ptr(C), ldc, twoprod, _stream(stream)))
    return C


def gemm(L, m, n, k, A, B, method="int8", target_bits=0, C=None, lda=None, ldb=None, ldc=None,
         stream=None):
    ldc = n if ldc is None else ldc
    if C is None:
        C = empty(m * ldc * L)
    info = GemmInfo()
    meth = METHODS[method] if isinstance(method, str) else method
    _check(lib().mdg_gemm(L, m, n, k, _ptr(A), k if lda is None else lda, _ptr(B),
                          n if ldb is None else ldb, _ptr(C), ldc, meth, target_bits,
                          ctypes.byref(info), _stream(stream)))
    return C, info.as_dict()


def gemm_noalloc(L, m, n, k, A, B, method="int8", target_bits=0, C=None, lda=None, ldb=None, ldc=None,
                 stream=None):
    """mdg_gemm_noalloc: mdg_gemm with its workspace in the library's static
    device array (C must be given or is allocated here by torch, outside the
    library)."""
    ldc = n if ldc is None else ldc
    if C is None:
        C = empty(m * ldc * L)
    info = GemmInfo()
    meth = METHODS[method] if isinstance(method, str) else method
    _check(lib().mdg_gemm_noalloc(L, m, n, k, _ptr(A), k if lda is None else lda, _ptr(B),
                                  n if ldb is None else ldb, _ptr(C), ldc, meth, target_bits,
                                  ctypes.byref(info), _stream(stream)))
    return C, info.as_dict()


def gemm_batched(problems, method="int8", target_bits=0, spread_bits=-1, persistent=1, stream=None):
    """problems: list of dicts with L, m, n, k, A, B, C (tensors) and optional lds."""
    arr = (Problem * len(problems))()
    for i, p in enumerate(problems):
        arr[i] = Problem(p["L"], p["m"], p["n"], p["k"], p["A"].data_ptr(), p.get("lda", p["k"]),
                         p["B"].data_ptr(), p.get("ldb", p["n"]), p["C"].data_ptr(),
                         p.get("ldc", p["n"]))
    meth = METHODS[method] if isinstance(method, str) else method
    if stream is None and persistent:
        # graph capture needs a non-null stream: run on a side stream ordered
        # after, and before, the caller's current stream
        import torch
        global _side_stream
        if _side_stream is None:
            _side_stream = torch.cuda.Stream()
        cur = torch.cuda.current_stream()
        _side_stream.wait_stream(cur)
        _set_last_stream(_side_stream)
        _check(lib().mdg_gemm_batched(len(problems), arr, meth, target_bits, spread_bits, persistent,
                                      ctypes.c_void_p(_side_stream.cuda_stream)))
        cur.wait_stream(_side_stream)
        return
    _set_last_stream(stream)
    _check(lib().mdg_gemm_batched(len(problems), arr, meth, target_bits, spread_bits, persistent,
                                  _stream(stream)))


_side_stream = None
_last_batch_stream = None


def _set_last_stream(stream):
    global _last_batch_stream
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    _last_batch_stream = stream


def batched_last_flags(count, stream=None):
    """(status, flags) of the last batched call on `stream` (default: the
    stream the last gemm_batched call ran on): flags[i] bit 0 =
    non-finite input, bit 1 = exponent spread above the allowance (C of
    that problem filled with NaN); status 0, 5 or 6 for the first flagged."""
    arr = (ctypes.c_int * max(count, 1))()
    if stream is None:
        stream = _last_batch_stream   # the stream gemm_batched last ran on
    rc = lib().mdg_batched_last_flags(_stream(stream), count, arr)
    if rc not in (0, 5, 6):
        _check(rc)
    return rc, [arr[i] for i in range(count)]


def batched_last_gemm_ms(stream=None):
    """GEMM-stage time of the last replay of a fused persistent batch built
    with profiling on (see mdg_batched_last_gemm_ms)."""
    ms = ctypes.c_float(0.0)
    if stream is None:
        stream = _last_batch_stream
    _check(lib().mdg_batched_last_gemm_ms(_stream(stream), ctypes.byref(ms)))
    return ms.value


def set_batch_plan_cache(max_plans, max_bytes):
    _check(lib().mdg_set_batch_plan_cache(max_plans, max_bytes))


def release_workspace():
    _check(lib().mdg_release_workspace())


def split4pd(x, p, stream=None):
    out = empty(4 * x.numel())
    _check(lib().mdg_split4pd(_ptr(x), x.numel(), p, _ptr(out), _stream(stream)))
    return out


def split8pd(x, p, stream=None):
    out = empty(8 * x.numel())
    _check(lib().mdg_split8pd(_ptr(x), x.numel(), p, _ptr(out), _stream(stream)))
    return out


def mixsplit2(x, stream=None):
    out = empty(6 * x.numel())
    _check(lib().mdg_mixsplit2(_ptr(x), x.numel(), _ptr(out), _stream(stream)))
    return out


def bigB2(Bq, n, pp, stream=None):
    out = empty((n * pp) ** 2)
    _check(lib().mdg_bigB2(_ptr(Bq), n, pp, _ptr(out), _stream(stream)))
    return out


def matmulhost(p, n, A, B, stream=None):
    """The reference's matmulhost (limb convolution in plain doubles, its loop
    order) on the GPU."""
    out = empty(n * n * p)
    _check(lib().mdg_matmulhost(p, n, _ptr(A), _ptr(B), _ptr(out), _stream(stream)))
    return out


def renormhost(A, n, p, big=False, stream=None):
    """The reference's renormhost (big=False) / renormhostbig (big=True), in
    place on the device tensor A."""
    _check(lib().mdg_renormhost(_ptr(A), n, p, 1 if big else 0, _stream(stream)))
    return A


def dgemm_acc(M, N, K, A, B, C, stream=None):
    _check(lib().mdg_dgemm_acc(M, N, K, _ptr(A), _ptr(B), _ptr(C), _stream(stream)))
    return C


def conv_gemm(m, n, k, pp, Aq, Bq, stream=None):
    out = empty(m * n * pp)
    _check(lib().mdg_conv_gemm(m, n, k, pp, _ptr(Aq), _ptr(Bq), _ptr(out), _stream(stream)))
    return out


def squeeze(L, pp, x, stream=None):
    nent = x.numel() // pp
    out = empty(nent * L)
    _check(lib().mdg_squeeze(L, nent, pp, _ptr(x), _ptr(out), _stream(stream)))
    return out


def tc_gemm(L, m, n, k, A, B, split_kind=4, stream=None):
    out = empty(m * n * L)
    _check(lib().mdg_tc_gemm(L, m, n, k, _ptr(A), _ptr(B), _ptr(out), split_kind, _stream(stream)))
    return out


def elementwise(op, L, a, b=None, twoprod=TWOPROD_FMA, out=None, stream=None):
    """c = op(a, b) entry by entry on L-limb interleaved arrays (mdg_elementwise).
    op: a name of OPS or its code; b: L-limb entries for add/sub/mul/div, one
    double per entry for mul_d/inc_d/mul_pwr2, unused otherwise."""
    code = OPS[op] if isinstance(op, str) else int(op)
    n = a.numel() // L
    if out is None:
        out = empty(n * L)
    _check(lib().mdg_elementwise(code, L, n, _ptr(a), _ptr(b), _ptr(out), twoprod, _stream(stream)))
    return out


def to_planar(L, rows, cols, x, ldx=None, stream=None):
    """Limb planes of an interleaved rows x cols matrix: out[l*rows*cols + r*cols + c]."""
    out = empty(rows * cols * L)
    _check(lib().mdg_to_planar(L, rows, cols, _ptr(x), cols if ldx is None else ldx, _ptr(out),
                               _stream(stream)))
    return out


def from_planar(L, rows, cols, planar, out=None, ldx=None, stream=None):
    ldx = cols if ldx is None else ldx
    if out is None:
        out = empty(rows * ldx * L)
    _check(lib().mdg_from_planar(L, rows, cols, _ptr(planar), _ptr(out), ldx, _stream(stream)))
    return out


def transpose(L, rows, cols, x, ldx=None, ldy=None, stream=None):
    """The cols x rows transpose of an interleaved rows x cols matrix."""
    ldy = rows if ldy is None else ldy
    out = empty(cols * ldy * L)
    _check(lib().mdg_transpose(L, rows, cols, _ptr(x), cols if ldx is None else ldx, _ptr(out), ldy,
                               _stream(stream)))
    return out
