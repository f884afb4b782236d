"""MI355X (gfx950) multiple-double GEMM: Python binding of libmdgemm.so.

The package directory holds the HIP sources (csrc/), the in-tree build
(build.py -> libmdgemm.so) and the ctypes binding (mdgemm.py).  The C ABI is
declared in include/mdgemm.h and the reference's C++ names in
include/mdgemm_compat.hpp.
"""
from .mdgemm import *  # noqa: F401,F403
from .mdgemm import lib, declared_symbols, MdgError, LIB_PATH  # noqa: F401
