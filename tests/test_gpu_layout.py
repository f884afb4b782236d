"""Layout utilities (mdg_to_planar, mdg_from_planar, mdg_transpose) against
numpy reshapes of the interleaved layout, bit for bit; pure data movement."""
import ctypes

import numpy as np
import pytest

from conftest import assert_bits_equal, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [2, 4, 8, 16])
@pytest.mark.parametrize("rows,cols,ld", [(1, 1, 1), (7, 300, 300), (33, 17, 20), (256, 513, 515)])
def test_planar_round_trip(mdg, orc, torch_cuda, L, rows, cols, ld):
    x = orc.generate(orc.GEN_QD, L, rows, ld, 5 * rows + cols, -30, 30)
    X = x.reshape(rows, ld, L)[:, :cols, :]
    planar = mdg.to_planar(L, rows, cols, to_dev(x), ldx=ld)
    torch_cuda.cuda.synchronize()
    want = np.ascontiguousarray(X.transpose(2, 0, 1)).ravel()
    assert_bits_equal(to_host(planar), want, "to_planar")
    # part() of the reference is one plane (C++/TCcuda.cu:34-40)
    if ld == cols:
        assert_bits_equal(to_host(planar)[rows * cols:2 * rows * cols], x[1::L], "part(x, p, 1)")
    back = to_dev(np.full(rows * ld * L, -7.0))
    mdg.from_planar(L, rows, cols, planar, out=back, ldx=ld)
    torch_cuda.cuda.synchronize()
    got = to_host(back).reshape(rows, ld, L)
    assert_bits_equal(got[:, :cols, :], X, "from_planar")
    assert np.all(got[:, cols:, :] == -7.0), "padding columns must stay untouched"


@pytest.mark.parametrize("L", [2, 4, 8, 16])
@pytest.mark.parametrize("rows,cols", [(1, 5), (16, 16), (31, 70), (300, 129)])
def test_transpose(mdg, orc, torch_cuda, L, rows, cols):
    ld = cols + 3
    x = orc.generate(orc.GEN_QD, L, rows, ld, 9 * rows + cols, -30, 30)
    y = mdg.transpose(L, rows, cols, to_dev(x), ldx=ld, ldy=rows + 1)
    torch_cuda.cuda.synchronize()
    X = x.reshape(rows, ld, L)[:, :cols, :]
    got = to_host(y).reshape(cols, rows + 1, L)[:, :rows, :]
    assert_bits_equal(got, X.transpose(1, 0, 2), "transpose")
    # A^T used as a GEMM operand: (A^T)^T A through the vector GEMM equals
    # the oracle's product of the transposed matrices
    if L <= 4 and rows <= 64:
        xt = np.ascontiguousarray(X.transpose(1, 0, 2)).ravel()
        At = mdg.transpose(L, rows, cols, to_dev(x), ldx=ld)
        Ad = to_dev(np.ascontiguousarray(X).ravel())
        C = mdg.vector_gemm(L, cols, cols, rows, At, Ad, twoprod=mdg.TWOPROD_DEKKER)
        torch_cuda.cuda.synchronize()
        want = orc.gemm(L, cols, cols, rows, xt, np.ascontiguousarray(X).ravel())
        assert_bits_equal(to_host(C), want, "A^T A")


def test_layout_argument_checks(mdg, torch_cuda):
    x = to_dev(np.ones(64))
    with pytest.raises(mdg.MdgError):
        mdg.transpose(2, 4, 8, x, ldx=4)          # row stride below cols
    with pytest.raises(mdg.MdgError):
        mdg.to_planar(3, 2, 2, x)
    # in place is rejected
    px = ctypes.c_void_p(x.data_ptr())
    assert mdg.lib().mdg_transpose(2, 4, 4, px, 4, px, 4, None) == 1
    assert mdg.transpose(2, 0, 5, x).numel() == 0
