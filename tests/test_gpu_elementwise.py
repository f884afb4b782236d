"""Elementwise p-double operations (mdg_elementwise: the reference's scalar
library off the GEMM path) against the oracle, bit for bit, for both two_prod
flavours; the oracle itself is pinned to the reference's code in
test_oracle.py."""
import numpy as np
import pytest

from conftest import assert_bits_equal_nan, to_dev, to_host

pytestmark = pytest.mark.gpu
LS = (2, 4, 8, 16)
OPS = ("add", "sub", "mul", "div", "sqr", "sqrt", "mul_d", "inc_d", "mul_pwr2", "abs", "neg")


def _inputs(orc, L, n, seed):
    a = orc.generate(orc.GEN_QD, L, n, 1, seed, -40, 40)
    b = orc.generate(orc.GEN_QD, L, n, 1, seed + 1, -40, 40)
    d = orc.generate(orc.GEN_QD, 1, n, 1, seed + 2, -40, 40)
    rng = np.random.default_rng(seed)
    a2, b2 = a.reshape(-1, L), b.reshape(-1, L)
    # whole zero entries (0/0, roots of zero), zero tail limbs (the
    # renormalisation's zero-skipping branches), small integers
    a2[:4] = 0.0
    b2[4:8] = 0.0
    a2[8:40, L // 2:] = 0.0
    b2[40:72, 1:] = 0.0
    a2[72:104] = 0.0
    a2[72:104, 0] = rng.integers(-9, 10, 32)
    b2[72:104] = 0.0
    b2[72:104, 0] = rng.integers(1, 10, 32)
    pw = np.ldexp(1.0, rng.integers(-30, 31, n)).astype(np.float64)
    # roots: mostly positive, some negative and zero entries
    sq = (a2 * np.where(a2[:, :1] < 0, -1.0, 1.0)).copy()
    sq[104:120] *= -1.0
    return a, b, d, pw, sq.ravel()


def _operands(op, a, b, d, pw, sq):
    if op in ("add", "sub", "mul", "div"):
        return a, b
    if op in ("mul_d", "inc_d"):
        return a, d
    if op == "mul_pwr2":
        return a, pw
    if op == "sqrt":
        return sq, None
    return a, None


@pytest.mark.parametrize("L", LS)
@pytest.mark.parametrize("op", OPS)
def test_elementwise_matches_oracle(mdg, orc, torch_cuda, L, op):
    n = 4099   # not a multiple of the 256-thread block
    a, b, d, pw, sq = _inputs(orc, L, n, 1000 + 7 * L)
    x, y = _operands(op, a, b, d, pw, sq)
    code = orc.__dict__["OP_" + op.upper()]
    for tp, mode in ((mdg.TWOPROD_FMA, orc.TP_FMA), (mdg.TWOPROD_DEKKER, orc.TP_DEKKER)):
        got = mdg.elementwise(op, L, to_dev(x), None if y is None else to_dev(y), twoprod=tp)
        torch_cuda.cuda.synchronize()
        want = orc.op(code, L, x, y, mode=mode)
        assert_bits_equal_nan(to_host(got), want, f"{op} L={L} twoprod={tp}")


@pytest.mark.parametrize("L", LS)
def test_elementwise_in_place_and_unaligned(mdg, orc, torch_cuda, L):
    torch = torch_cuda
    n = 1001
    a, b, _, _, _ = _inputs(orc, L, n, 77 + L)
    want = orc.op(orc.OP_DIV, L, a, b, mode=orc.TP_FMA)
    # result written over the first operand
    da, db = to_dev(a), to_dev(b)
    mdg.elementwise("div", L, da, db, out=da)
    torch.cuda.synchronize()
    assert_bits_equal_nan(to_host(da), want, "in place")
    # operands 8 bytes off a 16-byte boundary take the 8-byte access path
    buf_a = torch.zeros(n * L + 1, dtype=torch.float64, device="cuda")
    buf_b = torch.zeros(n * L + 1, dtype=torch.float64, device="cuda")
    buf_c = torch.zeros(n * L + 1, dtype=torch.float64, device="cuda")
    buf_a[1:] = to_dev(a)
    buf_b[1:] = to_dev(b)
    va, vb, vc = buf_a[1:], buf_b[1:], buf_c[1:]
    assert va.data_ptr() % 16 == 8
    mdg.elementwise("div", L, va, vb, out=vc)
    torch.cuda.synchronize()
    assert_bits_equal_nan(to_host(vc), want, "unaligned")
    assert to_host(buf_c)[0] == 0.0


def test_elementwise_argument_checks(mdg, torch_cuda):
    x = to_dev(np.ones(8))
    with pytest.raises(mdg.MdgError) as e:
        mdg.elementwise(11, 2, x)
    assert e.value.status == 1
    with pytest.raises(mdg.MdgError) as e:
        mdg.elementwise("add", 3, x, x)
    assert e.value.status == 1
    with pytest.raises(mdg.MdgError) as e:
        mdg.elementwise("add", 2, x, None)       # a binary op needs b
    assert e.value.status == 1
    with pytest.raises(mdg.MdgError) as e:
        mdg.elementwise("mul", 2, x, x, twoprod=2)
    assert e.value.status == 1
    empty = to_dev(np.zeros(0))
    assert mdg.elementwise("sqrt", 4, empty).numel() == 0
