"""The 2D kernels divide the stencil sums by dx^2 and dz^2 with a reciprocal
known per launch (qdiv in csrc/vti2d_kernels.h): q0 = RN(x y),
r = fma(-q0, d, x), q = fma(r, y, q0) for |x| in [2^-100, 2^100), zero with
x's sign, the division elsewhere.  The host (vti_div_reciprocal) hands out
y = RN(1/d) only after checking every mantissa of one binade; the results
must equal IEEE float32 division, which the reference's kernels use.

These tests run on the CPU: the proof routine through the library, and the
device sequence re-evaluated with exact rational arithmetic (a correctly
rounded fma) against numpy's float32 division across the whole exponent
window, including the edges and signed zeros.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib(vti):
    L = vti.lib
    L.vti_div_reciprocal.restype = ctypes.c_float
    L.vti_div_reciprocal.argtypes = [ctypes.c_float]
    L.vti_div_mismatches.restype = ctypes.c_longlong
    L.vti_div_mismatches.argtypes = [ctypes.c_float, ctypes.c_float]
    return L


def rn32(v: Fraction) -> np.float32:
    """v rounded to the nearest float32, ties to even (normal results)."""
    if v == 0:
        return np.float32(0.0)
    s = -1 if v < 0 else 1
    v = abs(v)
    e = v.numerator.bit_length() - v.denominator.bit_length()
    if Fraction(2) ** e > v:
        e -= 1
    # v in [2^e, 2^(e+1)): 24 significant bits, ulp 2^(e-23)
    scaled = v / Fraction(2) ** (e - 23)
    n = scaled.numerator // scaled.denominator
    rem = scaled - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(s * float(Fraction(n) * Fraction(2) ** (e - 23)))


def qdiv_model(x: np.float32, d: np.float32, y: np.float32) -> np.float32:
    a = int(np.array(x).view(np.uint32)) & 0x7FFFFFFF
    if a != 0 and not (0x0D800000 <= a < 0x71800000):
        return np.float32(x / d)
    q0 = np.float32(x * y)
    r = rn32(Fraction(float(x)) - Fraction(float(q0)) * Fraction(float(d)))
    q = rn32(Fraction(float(r)) * Fraction(float(y)) + Fraction(float(q0)))
    qb = (int(np.array(q).view(np.uint32)) & 0x7FFFFFFF) | (int(np.array(x).view(np.uint32)) & 0x80000000)
    return np.array(qb, np.uint32).view(np.float32)[()]


@pytest.mark.parametrize("d", [100.0, 156.25, 625.0, 3.0, 0.3])
def test_reciprocal_is_proven_and_rounded(lib, d):
    y = lib.vti_div_reciprocal(d)
    assert np.float32(y) == np.float32(1.0) / np.float32(d)
    assert lib.vti_div_mismatches(d, y) == 0


def test_wrong_reciprocal_is_caught(lib):
    # a few ulps off is absorbed by the correction step (the check finds no
    # failing case); a reciprocal off in its 12th bit, or of another divisor,
    # leaves q0 too far from x/d and must fail
    y = np.float32(1.0) / np.float32(100.0)
    assert lib.vti_div_mismatches(100.0, float(np.nextafter(y, np.float32(1)))) == 0
    for bad in (y * np.float32(1 + 2.0 ** -12), y * np.float32(1 - 2.0 ** -12), np.float32(1.0) / np.float32(101.0)):
        assert lib.vti_div_mismatches(100.0, float(bad)) > 0


@pytest.mark.parametrize("d", [0.0, -100.0, 2.0 ** 21, 2.0 ** -21, float("nan"), float("inf")])
def test_divisors_without_proof_fall_back(lib, d):
    assert lib.vti_div_reciprocal(d) == 0.0


@pytest.mark.parametrize("d", [100.0, 156.25, 0.3])
def test_sequence_equals_division_across_window(lib, d):
    d = np.float32(d)
    y = np.float32(lib.vti_div_reciprocal(float(d)))
    assert y != 0
    rng = np.random.default_rng(7)
    exps = np.concatenate([rng.integers(-100, 100, 1500), [-100, -99, 99, 98, 0, 1, -1]])
    mant = rng.integers(0, 1 << 23, exps.size)
    bits = ((exps + 127).astype(np.uint32) << 23) | mant.astype(np.uint32)
    sign = rng.integers(0, 2, exps.size).astype(np.uint32) << 31
    xs = (bits | sign).view(np.float32)
    # the window's edges and what lies just outside, zeros of both signs
    extra = np.array([0x0D800000, 0x0D7FFFFF, 0x717FFFFF, 0x71800000, 0x00000001, 0x00400000, 0x0, 0x80000000,
                      0x8D800000, 0xF17FFFFF], np.uint32).view(np.float32)
    for x in np.concatenate([xs, extra]):
        got = qdiv_model(x, d, y)
        want = np.float32(x / d)
        assert np.array(got).view(np.uint32) == np.array(want).view(np.uint32), (float(x), float(d))
