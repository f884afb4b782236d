"""Write the golden fixtures of tests/golden/ from the reference's own code.

The reference's arithmetic and splitter (C++/*_double_functions.cu,
C++/floatybits.cu) are compiled from /root/reference by `make -C oracle ref`
into oracle/_ref/libmdref.so; this script feeds that library seeded inputs
and stores inputs and outputs as .npz data (no code).  Run it in a container
that has /root/reference:

    python tests/golden/gen_golden.py

Inputs come from the oracle's counter-based generators plus deliberately
awkward cases (zero limbs, small integers) that drive the renormalisation
ladder through its zero-skipping branches.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402


def op_inputs(L, cnt, seed):
    rng = np.random.default_rng(seed)
    xs = [
        O.generate(O.GEN_QD, L, cnt, 1, seed, -8, 8),
        O.generate(O.GEN_QD, L, cnt, 1, seed + 1, -60, 60),
        O.generate(O.GEN_REF, L, cnt, 1, seed + 2, 0, 0),
        O.generate(O.GEN_DD, L, cnt, 1, seed + 3),
    ]
    z = O.generate(O.GEN_QD, L, cnt, 1, seed + 4, -3, 3)
    z[rng.random(z.size) < 0.3] = 0.0
    xs.append(z)
    xs.append(rng.integers(-4, 5, size=cnt * L).astype(np.float64))
    return np.concatenate(xs)


def main():
    O.build(force=True)
    if not O.build_ref():
        raise SystemExit("oracle/_ref could not be built (needs /root/reference)")
    out = {}
    for L in (2, 4, 8, 16):
        a = op_inputs(L, 64, 100 + L)
        b = op_inputs(L, 64, 200 + L)
        y = op_inputs(1, 64 * L, 300 + L)[: a.size // L]
        out[f"L{L}_a"], out[f"L{L}_b"], out[f"L{L}_y"] = a, b, y
        out[f"L{L}_mul"] = O.op_mul(L, a, b, use_ref=True)
        out[f"L{L}_add"] = O.op_add(L, a, b, use_ref=True)
        out[f"L{L}_inc_d"] = O.op_inc_d(L, a, y, use_ref=True)
        # the scalar operations off the GEMM path
        sq = (a.reshape(-1, L) * np.where(a.reshape(-1, L)[:, :1] < 0, -1.0, 1.0)).ravel()
        pw = np.ldexp(1.0, np.arange(a.size // L) % 41 - 20)
        out[f"L{L}_sqrt_x"], out[f"L{L}_pw"] = sq, pw
        for name, code, x, bb in (("sub", O.OP_SUB, a, b), ("div", O.OP_DIV, a, b),
                                  ("sqr", O.OP_SQR, a, None), ("sqrt", O.OP_SQRT, sq, None),
                                  ("sqrt_signed", O.OP_SQRT, a, None),
                                  ("mul_d", O.OP_MUL_D, a, y), ("mul_pwr2", O.OP_MUL_PWR2, a, pw),
                                  ("abs", O.OP_ABS, a, None), ("neg", O.OP_NEG, a, None)):
            out[f"L{L}_op_{name}"] = O.op(code, L, x, bb, use_ref=True)
        if L > 2:
            f = op_inputs(L + 1, 64, 400 + L)
            out[f"L{L}_f"] = f
            out[f"L{L}_fast_renorm"] = O.op_fast_renorm(L, f, use_ref=True)
            out[f"L{L}_renorm"] = O.op_renorm(L, f, use_ref=True)
        # small GEMMs: signed spread inputs, and the reference's own generator
        for tag, kind, emin, emax in (("qd", O.GEN_QD, -8, 8), ("ref", O.GEN_REF, 0, 0)):
            m, n, k = 5, 7, 9
            A = O.generate(kind, L, m, k, 500 + L, emin, emax)
            B = O.generate(kind, L, k, n, 600 + L, emin, emax)
            out[f"L{L}_gemm_{tag}_A"], out[f"L{L}_gemm_{tag}_B"] = A, B
            out[f"L{L}_gemm_{tag}_C"] = O.gemm(L, m, n, k, A, B, use_ref=True)
        # splitter and squeeze
        for tag, kind in (("ref", O.GEN_REF), ("qd", O.GEN_QD)):
            v = O.generate(kind, L, 24, 1, 700 + L, -5, 5)
            out[f"L{L}_split_{tag}_x"] = v
            out[f"L{L}_split4pd_{tag}"] = O.split4pd(v, L, use_ref=True)
            out[f"L{L}_split8pd_{tag}"] = O.split8pd(v, L, use_ref=True)
        sq = op_inputs(1, 24 * 4 * L, 800 + L)
        out[f"L{L}_squeeze_x"] = sq
        out[f"L{L}_squeeze_y"] = O.squeeze(L, 4 * L, sq, use_ref=True)
    for tag, kind in (("ref", O.GEN_REF), ("qd", O.GEN_QD)):
        v = O.generate(kind, 2, 24, 1, 900, -5, 5)
        out[f"mix_{tag}_x"] = v
        out[f"mix_{tag}"] = O.mixsplit2(v, use_ref=True)
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
