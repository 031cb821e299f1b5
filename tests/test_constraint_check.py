"""The division-free constraint test of the Euler stage kernels
(alsvinn_amd/csrc/physics.hpp: EulerEq::surely_obeys) is sufficient for the
reference's check (CPUCellComputer.cpp:79-125 with Euler::obeysConstraints:
every conserved value, every velocity and the pressure finite, rho > 0,
p > 0), evaluated as the kernels evaluate it (make(): u = m/rho,
ie = E - (0.5*m.m)/rho, p = (gamma - 1)*ie, IEEE divisions). The kernels
run the full check on every cell the test does not clear, so the outcome is
the reference's on every cell; these tests check the implication on states
drawn at the test's margin and bounds, where a smaller margin, a missing
bound or a wrong comparison would let a failing state through.
"""
import math

import numpy as np
import pytest


def _obeys(q, gamma):
    rho, E = q[0], q[-1]
    m = q[1:-1]
    dot = m[0] * m[0]
    for mk in m[1:]:
        dot = dot + mk * mk
    with np.errstate(all="ignore"):
        ie = E - (0.5 * dot) / rho if rho != 0 else math.nan
        u = [mk / rho if rho != 0 else math.nan for mk in m]
        p = (gamma - 1.0) * ie
    ok = all(math.isfinite(x) for x in q) and all(math.isfinite(x) for x in u)
    return ok and math.isfinite(p) and rho > 0 and p > 0


def _surely(q, gamma):
    rho, E = q[0], q[-1]
    m = q[1:-1]
    dot = m[0] * m[0]
    for mk in m[1:]:
        dot = dot + mk * mk
    with np.errstate(all="ignore"):
        K = 0.5 * dot
        rE = rho * E
        return (gamma > 1.0 + 2.0 ** -40 and gamma < 2.0 ** 100 and rho > 2.0 ** -100
                and E > 2.0 ** -100 and rE < 2.0 ** 200 and rE > K * (1.0 + 2.0 ** -20))


def _states(rng, nsd, n):
    """States whose energy lies near the kinetic energy (relative offsets
    from -2^-10 to 2^-10, and down to 2^-60 on either side), with
    magnitudes up to the test's bounds and beyond."""
    out = []
    for _ in range(n):
        rho = float(2.0 ** rng.uniform(-110, 110))
        m = [float(rng.standard_normal() * 2.0 ** rng.uniform(-110, 110)) for _ in range(nsd)]
        dot = m[0] * m[0]
        for mk in m[1:]:
            dot = dot + mk * mk
        kin = 0.5 * dot / rho
        sgn = rng.choice([-1.0, 1.0])
        d = sgn * 2.0 ** rng.uniform(-60, -10)
        E = kin * (1.0 + d) if kin > 0 else float(2.0 ** rng.uniform(-110, 10))
        if not math.isfinite(E):
            continue
        out.append([rho] + m + [E])
    return out


@pytest.mark.parametrize("nsd", [1, 2, 3])
def test_division_free_test_implies_the_full_check(nsd):
    rng = np.random.default_rng(10 + nsd)
    gamma = 1.4
    cleared = rejected = 0
    for q in _states(rng, nsd, 60000):
        if _surely(q, gamma):
            cleared += 1
            assert _obeys(q, gamma), q
        elif _obeys(q, gamma):
            rejected += 1  # valid, left to the full check
    # the sample exercises both sides of the margin
    assert cleared > 1000 and rejected > 1000


def test_ordinary_states_are_cleared():
    """The Kelvin-Helmholtz states (rho 1..2, |u| <= 1, p = 2.5) never need
    the full check."""
    rng = np.random.default_rng(3)
    for _ in range(2000):
        rho = float(rng.uniform(1, 2))
        u = rng.uniform(-1, 1, 2)
        p = 2.5 * float(rng.uniform(0.5, 1.5))
        E = p / 0.4 + 0.5 * rho * float(u @ u)
        assert _surely([rho, rho * u[0], rho * u[1], E], 1.4)


@pytest.mark.parametrize("q", [
    [math.inf, 0.0, 1.0], [1.0, math.nan, 1.0], [1.0, 0.0, math.inf], [-1.0, 0.0, 1.0],
    [0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [2.0 ** -101, 0.0, 1.0], [1.0, 2.0 ** 101, 2.0 ** 200],
    [2.0 ** 150, 0.0, 2.0 ** 60],
])
def test_special_values_are_left_to_the_full_check(q):
    assert not _surely(q, 1.4)


@pytest.mark.parametrize("gamma", [1.0, 1.0 + 2.0 ** -41, 2.0 ** 100, math.nan])
def test_gamma_outside_the_bounds_clears_nothing(gamma):
    assert not _surely([1.0, 0.5, 3.0], gamma)


def test_a_smaller_margin_would_be_wrong():
    """With a margin of 2^-60 instead of 2^-20 the test would clear states
    whose pressure rounds to zero or below: the margin is load-bearing."""
    rng = np.random.default_rng(5)
    bad = 0
    for q in _states(rng, 2, 60000):
        rho, E = q[0], q[-1]
        K = 0.5 * (q[1] * q[1] + q[2] * q[2])
        loose = rho > 2.0 ** -100 and E > 2.0 ** -100 and rho * E < 2.0 ** 200 and \
            rho * E > K * (1.0 + 2.0 ** -60)
        if loose and not _obeys(q, 1.4):
            bad += 1
    assert bad > 0
