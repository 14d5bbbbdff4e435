"""The engine's sin/cos (shapes_amd/csrc/shp_trig.h) against mathematics
and against the host libm that GHC's sinDouble#/cosDouble# would call
(rotate22, shapes/src/Physics/Linear.hs:353-357)."""
import math
import os
import sys
from fractions import Fraction

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "shapes_amd", "csrc"))
from gen_trig_constants import pi_fraction  # noqa: E402

PI = pi_fraction()


FRAC = 400
PIO2_FIX = (PI.numerator << FRAC) // (2 * PI.denominator)  # floor(pi/2 * 2^400)


def exact_sincos(x: float):
    """sin/cos of a double to ~2^-390 absolute: exact reduction against a
    400-bit pi/2, then Taylor series in 400-bit integer fixed point.  The
    result is returned as an exact Fraction of the fixed-point value."""
    m, e = Fraction(x).numerator, Fraction(x).denominator
    xf = (m << FRAC) // e if m >= 0 else -((-m << FRAC) // e)
    k = (2 * xf + PIO2_FIX) // (2 * PIO2_FIX)  # nearest integer
    r = xf - k * PIO2_FIX
    one = 1 << FRAC
    r2 = (r * r) >> FRAC
    s = 0
    t = r
    for mm in range(60):
        s += t
        t = -((t * r2) >> FRAC) // ((2 * mm + 2) * (2 * mm + 3))
    c = 0
    t = one
    for mm in range(60):
        c += t
        t = -((t * r2) >> FRAC) // ((2 * mm + 1) * (2 * mm + 2))
    S = Fraction(s, one)
    Cc = Fraction(c, one)
    return [(S, Cc), (Cc, -S), (-S, -Cc), (-Cc, S)][k % 4]


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    xs = list(rng.uniform(-4, 4, n)) + list(rng.uniform(-200, 200, n // 4))
    xs += [math.pi / 2, -math.pi / 2, math.pi, 3 * math.pi / 4, 2.0 ** -20,
           0.7853981633974483, 1e6 + 0.5]
    return xs


def test_sincos_correctly_rounded(orc):
    for x in _samples(1500, 1):
        s, c = orc.sincos(x, orc.TRIG_ENGINE)
        es, ec = exact_sincos(x)
        assert s == float(es), x.hex()
        assert c == float(ec), x.hex()


def test_sincos_special_values(orc):
    assert orc.sincos(0.0) == (0.0, 1.0)
    s, c = orc.sincos(-0.0)
    assert math.copysign(1.0, s) == -1.0 and c == 1.0
    for bad in (math.inf, -math.inf, math.nan, 2.0 ** 51):
        s, c = orc.sincos(bad)
        assert math.isnan(s) and math.isnan(c)


def test_libm_census(orc):
    """Where the host libm and the engine disagree, the engine is the
    correctly rounded one (glibc's sin/cos are not correctly rounded in
    about 0.1-0.3% of arguments)."""
    rng = np.random.default_rng(5)
    xs = rng.uniform(-10, 10, 20000)
    mism = 0
    for x in xs:
        s, c = orc.sincos(float(x), orc.TRIG_ENGINE)
        ls, lc = orc.sincos(float(x), orc.TRIG_LIBM)
        if (s, c) != (ls, lc):
            mism += 1
            es, ec = exact_sincos(float(x))
            assert (s, c) == (float(es), float(ec))
    assert mism < 0.01 * len(xs)


def test_sincos_tiny_arguments(orc):
    """|x| < 2^-27: sin x rounds to x and cos x to 1."""
    for x in (1e-300, -5e-324, 2.0 ** -28, -3e-9):
        assert orc.sincos(x) == (x, 1.0)
