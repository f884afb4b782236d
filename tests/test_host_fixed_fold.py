"""The slice GEMMs' fixed-point arithmetic, compiled for the host
(tests/cpp/host_fixed_fold.hip over csrc/slices.hpp and csrc/recon.hpp), no
GPU needed:

* to_fixed: N = sum over limbs of floor(x_l 2^shift) in two's complement
  over NW words, against Python integers;
* fixed_digits (any w) and fixed_digits8 (w = 8): the balanced base-2^w
  digits of N, slice 0 most significant, against a Python carry chain and
  against N itself (sum of digits times powers of 2^w);
* band_fold over INT64 band planes (bands of 4 and of 3 diagonals, chunked
  and unchunked) against diag_fold over the same diagonal sums: the same
  limbs bit for bit."""
import os
import random
import struct
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"
SRC = os.path.join(ROOT, "tests", "cpp", "host_fixed_fold.hip")
NW = {2: 8, 4: 12, 8: 20, 16: 36}


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("hff") / "host_fixed_fold")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O1", "-std=c++17", "-ffp-contract=off", SRC,
                    "-o", out], check=True, capture_output=True, timeout=600)
    return out


def _floor_scaled(x, shift):
    num, den = x.as_integer_ratio()   # den is a power of two
    num <<= max(shift, 0)
    den <<= max(-shift, 0)
    return num // den


def _balanced_digits(N, S, w):
    """slice t = S-1-i holds digit i (least significant first), digits
    0 .. S-2 in [-2^(w-1), 2^(w-1)), the top one the signed remainder"""
    d = [0] * S
    v = N
    for i in range(S - 1):
        r = v & ((1 << w) - 1)
        if r >= 1 << (w - 1):
            r -= 1 << w
        d[S - 1 - i] = r
        v = (v - r) >> w
    d[0] = v
    return d


def _case(rng, L, w):
    e = rng.randint(-40, 40)
    lead = rng.uniform(1.0, 2.0) * rng.choice((-1.0, 1.0)) * 2.0 ** e
    x = [lead]
    for _ in range(1, L):
        x.append(x[-1] * rng.uniform(-0.5, 0.5) * 2.0 ** -53)
    if rng.random() < 0.1:
        x[rng.randrange(L)] = 0.0
    if rng.random() < 0.05:
        x[-1] = rng.uniform(-1, 1) * 2.0 ** -1070    # subnormal limb
    E = e + 2 + (rng.randint(0, 100) if rng.random() < 0.3 else 0)
    S = max(1, rng.randint(53 * L // w, (32 * NW[L] - 1) // w))
    shift = S * w - 2 - E
    return x, shift, S


@pytest.mark.parametrize("L", [2, 4, 8, 16])
@pytest.mark.parametrize("w", [8, 13, 24])
def test_fixed_point_and_digits(exe, L, w):
    rng = random.Random(1000 * L + w)
    cases = [_case(rng, L, w) for _ in range(300)]
    lines = []
    for x, shift, S in cases:
        bits = " ".join("%016x" % struct.unpack("<Q", struct.pack("<d", v))[0] for v in x)
        lines.append(f"{L} {shift} {S} {w} {bits}")
    r = subprocess.run([exe, "fixed"], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout.strip().split("\n")
    assert len(out) == len(cases)
    for (x, shift, S), ln in zip(cases, out):
        parts = ln.split("|")
        words = [int(v, 16) for v in parts[0].split()]
        N = sum(_floor_scaled(v, shift) for v in x)          # per-limb floor, as documented
        mod = 1 << (32 * NW[L])
        assert sum(v << (32 * j) for j, v in enumerate(words)) == N % mod, (x, shift)
        want = _balanced_digits(N, S, w)
        got = [int(v) for v in parts[1].split()]
        assert got == want, (x, shift, S, w)
        assert sum(d << (w * (S - 1 - t)) for t, d in enumerate(got)) == N
        if w == 8:
            assert [int(v) for v in parts[2].split()] == want, (x, shift, S)


def test_band_fold_equals_diagonal_fold(exe):
    r = subprocess.run([exe, "fold", "7", "400"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "mismatch" not in r.stdout
    assert int(r.stdout.split("cases")[1]) >= 1600


def test_band_items_cover_every_plane_once(exe):
    # the band kernels' work items for slice counts 1..255, bands of 3 and 4,
    # depths up to 2047 k-blocks (to 32767 for a few slice counts) and four
    # chunk capacities: every plane the
    # epilogue writes lies inside the plane array and is written once, each
    # band's k-blocks are covered once, and each chunk's INT32 sums are exact
    r = subprocess.run([exe, "items"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert int(r.stdout.split("items")[1]) > 100000
