"""The FP64 division and square-root sequences of the stage kernels
(alsvinn_amd/csrc/physics.hpp: rcp_refined, dv, dvp, dsqrt), replayed on the CPU
with exact fused multiply-adds (rational arithmetic, one rounding each) from
a perturbed reciprocal / reciprocal square root estimate, give the IEEE
results bit for bit, including the sign of zero quotients, for operands in
the ranges of valid states. The GPU parity tests check the same end to end
on the hardware's estimates."""
import math
import struct
from fractions import Fraction

import numpy as np
import pytest


def _rn(q):
    return float(q)  # Fraction -> nearest double, ties to even


def _fma(a, b, c):
    if a == 0 or b == 0:
        # signed-zero rules of a*b + c when the product is zero
        prod = math.copysign(0.0, a) * math.copysign(1.0, b) if a == 0 else \
            math.copysign(0.0, b) * math.copysign(1.0, a)
        return prod + c
    r = Fraction(a) * Fraction(b) + Fraction(c)
    if r == 0:
        return 0.0
    return _rn(r)


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _dvp(a, b, est):
    """dvp(): denominators positive in every valid state."""
    r = est
    t = _fma(-b, r, 1.0)
    r = _fma(r, t, r)
    t = _fma(-b, r, 1.0)
    r = _fma(r, t, r)
    q = a * r
    e = _fma(b, q, -a)
    return _fma(-e, r, q)


def _div(a, b, est):
    """dv(): either sign; the result takes the sign of a*r."""
    r = est
    t = _fma(-b, r, 1.0)
    r = _fma(r, t, r)
    t = _fma(-b, r, 1.0)
    r = _fma(r, t, r)
    q = a * r
    e = _fma(b, q, -a)
    x = _fma(-e, r, q)
    return math.copysign(x, q)


def _sqrt(x, est):
    # dsqrt(): the +-inf estimate of a zero operand is replaced by 2^1023
    y0 = 2.0 ** 1023 if x == 0.0 else est
    g = x * y0
    h = y0 * 0.5
    r = _fma(-h, g, 0.5)
    g = _fma(g, r, g)
    h = _fma(h, r, h)
    d = _fma(-g, g, x)
    g = _fma(d, h, g)
    d = _fma(-g, g, x)
    return _fma(d, h, g)


def test_division_sequence_is_ieee_for_ordinary_operands():
    rng = np.random.default_rng(1)
    n = 4000
    # magnitudes 2^-60 .. 2^60 (far inside the no-rescale range), both signs
    a = rng.standard_normal(n) * np.exp2(rng.integers(-60, 61, n))
    b = rng.standard_normal(n) * np.exp2(rng.integers(-60, 61, n))
    a[:20] = 0.0
    a[20:40] = -0.0
    a[40:60] = b[40:60] * 3  # exact quotients
    est_err = rng.uniform(-2.0 ** -20, 2.0 ** -20, n)
    for x, y, d in zip(a, b, est_err):
        est = float(Fraction(1) / Fraction(float(y))) * (1.0 + d)
        got = _div(float(x), float(y), est)
        assert _bits(got) == _bits(float(x) / float(y)), (x, y)


def test_sqrt_sequence_is_ieee_for_ordinary_operands():
    rng = np.random.default_rng(2)
    n = 4000
    x = np.abs(rng.standard_normal(n)) * np.exp2(rng.integers(-200, 201, n))
    x[:10] = 0.0
    est_err = rng.uniform(-2.0 ** -20, 2.0 ** -20, n)
    for v, d in zip(x, est_err):
        v = float(v)
        est = (1.0 / math.sqrt(v) if v > 0 else math.inf) * (1.0 + d)
        got = _sqrt(v, est)
        assert _bits(got) == _bits(math.sqrt(v)), v


@pytest.mark.parametrize("zero", [0.0, -0.0])
def test_sqrt_sequence_of_signed_zero(zero):
    """A zero operand runs the whole sequence from the substituted estimate
    2^1023 (no select on the result) and gives the IEEE square root, the
    zero itself with its sign."""
    got = _sqrt(zero, math.copysign(math.inf, zero))
    assert _bits(got) == _bits(math.sqrt(zero)) == _bits(zero)


def test_shared_reciprocal_gives_each_quotient():
    """Quotients with a common denominator share the refined reciprocal:
    each still equals its own IEEE quotient."""
    rng = np.random.default_rng(3)
    for _ in range(500):
        y = float(rng.uniform(0.1, 10))
        est = (1.0 / y) * (1.0 + float(rng.uniform(-1e-7, 1e-7)))
        xs = rng.standard_normal(3) * 5
        for x in xs:
            assert _bits(_div(float(x), y, est)) == _bits(float(x) / y)


@pytest.mark.parametrize("x,y", [(-0.0, 2.0), (0.0, -3.0), (-0.0, -5.0), (0.0, 7.0)])
def test_signed_zero_quotients(x, y):
    assert _bits(_div(x, y, 1.0 / y)) == _bits(x / y)
    if y > 0:
        assert _bits(_dvp(x, y, 1.0 / y)) == _bits(x / y)


def test_positive_denominator_variant():
    """dvp() on positive denominators (with zero numerators of both signs)
    equals IEEE division; it is used only where the denominator is a
    density, a pressure or a sum of positive weights."""
    rng = np.random.default_rng(4)
    n = 2000
    a = rng.standard_normal(n) * np.exp2(rng.integers(-40, 41, n))
    b = np.abs(rng.standard_normal(n)) * np.exp2(rng.integers(-40, 41, n)) + 1e-300
    a[:10] = 0.0
    a[10:20] = -0.0
    for x, y in zip(a, b):
        est = (1.0 / float(y)) * (1.0 + float(rng.uniform(-2.0 ** -20, 2.0 ** -20)))
        assert _bits(_dvp(float(x), float(y), est)) == _bits(float(x) / float(y))
