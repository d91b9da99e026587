"""The frame kernels' quotient by a per-frame denominator (div_rn in
fuzzyblue_amd/csrc/render.hip): with y = RN(1/b), q = RN(a*y),
q' = RN(q + RN(a - b*q)*y) (two FMAs) equals the IEEE quotient RN(a/b) --
Markstein's theorem -- for the numerators it is used on (|a| >= 2^-80 or
a = 0; smaller ones take the division).  The residual a - b*q is a
multiple of about 2^(e_a - 46), so from 2^-80 on it is a normal number and
the FMAs are exact whether or not subnormals are flushed.  Checked here in exact rational
arithmetic with binary32 rounding, on random pairs in the ranges of the sky
lookups' coordinates (distances in km, dot products of unit vectors with
radii) and across the binary32 exponent range."""
from fractions import Fraction

import numpy as np


def rn32(x: Fraction) -> np.float32:
    """Round a rational to the nearest binary32 (ties to even), subnormals
    included."""
    if x == 0:
        return np.float32(0.0)
    neg = x < 0
    x = -x if neg else x
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    e = max(e, -126)  # below 2^-126 the quantum stays 2^-149
    assert e <= 127
    m = x * Fraction(2) ** (23 - e)  # in [2^23, 2^24) for normal x
    n = m.numerator // m.denominator
    rem = m - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    v = float(n) * 2.0 ** (e - 23)
    return np.float32(-v if neg else v)


def fma32(a, b, c) -> np.float32:
    return rn32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def div_rn(a: np.float32, b: np.float32) -> np.float32:
    y = np.float32(1.0) / b
    q = a * y
    return fma32(fma32(-q, b, a), y, q)


def check(a, b):
    bad = []
    for x, z in zip(a, b):
        want = x / z
        got = div_rn(x, z)
        if got.view(np.uint32) != want.view(np.uint32):
            bad.append((float(x), float(z), float(got), float(want)))
    return bad


def test_matches_ieee_division_in_the_lookup_ranges():
    rng = np.random.default_rng(3)
    n = 4000
    cases = [
        (rng.uniform(-1e4, 1e4, n), rng.uniform(1e-3, 1e4, n)),  # distances (km)
        (rng.uniform(-7e3, 7e3, n), rng.uniform(6e3, 7e3, n)),  # r*mu / r
        (rng.uniform(0.0, 200.0, n), rng.uniform(0.05, 200.0, n)),  # (d - d_min) / span
    ]
    for a, b in cases:
        a = a.astype(np.float32)
        b = b.astype(np.float32)
        assert check(a, b) == []


def test_matches_ieee_division_across_exponents():
    rng = np.random.default_rng(4)
    n = 4000
    a = (rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-80, 60, n)).astype(np.float32)
    a *= np.where(rng.random(n) < 0.5, -1, 1).astype(np.float32)
    b = (rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-40, 40, n)).astype(np.float32)
    keep = (np.abs(a / b) > 2.0 ** -100) & (np.abs(a / b) < 2.0 ** 100)
    assert keep.sum() > n // 2
    assert check(a[keep], b[keep]) == []


def test_tiny_numerators_are_why_the_fallback_exists():
    """Far below 2^-80 the residual a - b*q is not representable even as a
    subnormal and the two-FMA quotient differs from the IEEE one: the
    kernel takes the division for such numerators.  (Pins that the
    fallback is needed.)"""
    rng = np.random.default_rng(5)
    n = 3000
    a = (rng.uniform(1.0, 2.0, n) * 2.0 ** -125).astype(np.float32)
    b = (rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-110, -20, n)).astype(np.float32)
    keep = (np.abs(a / b) > 2.0 ** -120) & (np.abs(a / b) < 2.0 ** 100)
    assert len(check(a[keep], b[keep])) > 0
