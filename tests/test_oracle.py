"""Pins the CPU oracle (no GPU): the reference's own known-answer tests,
closed forms, conservation/symmetry invariants, convergence rates and an
exact Riemann solution. Each test names the reference test or the
mathematics that fixes the expected value.
"""
import math

import numpy as np
import pytest

from tests import oracle_lib as O

GAMMA = 1.4


def euler_state(rho, vel, p, gamma=GAMMA):
    vel = list(vel)
    dot = sum(v * v for v in vel)
    return np.array([rho] + [rho * v for v in vel] + [p / (gamma - 1) + 0.5 * rho * dot])


# ---------------------------------------------------------------- sampler

def test_mt19937_64_known_answer():
    """C++ standard [rand.predef]: the 10000th invocation of a
    default-constructed mt19937_64 produces 9981545732273789042."""
    assert O.mt19937_64_nth(10000) == 9981545732273789042


def test_sampler_sample_major_draw_order():
    """STLMersenne.cpp:38-85: component c of sample s is the (s*D+c)-th draw;
    a run starting at sample_start skips forward to exactly that draw."""
    full = O.stl_mersenne_uniform(6, 5)
    tail = O.stl_mersenne_uniform(3, 5, sample_start=3)
    assert np.array_equal(full[3:], tail)
    flat = O.stl_mersenne_uniform(1, 30)[0]
    assert np.array_equal(full.ravel(), flat)
    scaled = O.stl_mersenne_uniform(6, 5, lower=-2.0, upper=3.0)
    assert np.array_equal(scaled, full * 5.0 + (-2.0))


def test_sampler_moments_converge():
    """MeanVarTest.cpp:85-149: sample mean/variance of U[0,1] draws converge
    to 1/2 and 1/12 at rate ~ M^-1/2."""
    x = O.stl_mersenne_uniform(1, 200000)[0]
    assert abs(x.mean() - 0.5) < 5 / np.sqrt(12 * len(x))
    assert abs(x.var() - 1 / 12) < 5e-3
    assert x.min() >= 0 and x.max() < 1


# ---------------------------------------------------------------- fluxes

def np_point_flux(q, d, gamma=GAMMA):
    """Euler flux through a face of normal e_d, written from the primitive
    variables in extended precision (Toro, Riemann Solvers, 3rd ed., eq. 3.60):
    (rho u_d, rho u u_d + p e_d, (E + p) u_d) with p = (gamma-1)(E - rho|u|^2/2).
    Independent of the oracle's Euler::pointFlux, which it pins."""
    ld = np.longdouble
    q = np.asarray(q, dtype=ld)
    rho, E = q[0], q[-1]
    u = q[1:-1] / rho
    p = (ld(gamma) - 1) * (E - ld(0.5) * rho * np.dot(u, u))
    F = np.empty(len(q), dtype=ld)
    F[0] = rho * u[d]
    F[1:-1] = rho * u * u[d]
    F[1 + d] += p
    F[-1] = (E + p) * u[d]
    return F


def _flux_scale(*arrays):
    return max(float(np.max(np.abs(np.asarray(a, dtype=float)))) for a in arrays)


@pytest.mark.parametrize("nsd", [1, 2, 3])
def test_point_flux_textbook(nsd):
    """Euler.hpp:172-181 computePointFlux: the oracle's point flux equals the
    textbook Euler flux (np_point_flux) in every direction, transverse
    momentum components included."""
    rng = np.random.default_rng(100 + nsd)
    for _ in range(20):
        rho, p = rng.uniform(0.2, 3.0, 2)
        u = rng.uniform(-2, 2, nsd)
        q = euler_state(rho, list(u), p)
        for d in range(nsd):
            f = O.point_flux(nsd, d, GAMMA, q)
            expect = np_point_flux(q, d)
            np.testing.assert_allclose(f, expect.astype(float), rtol=1e-13,
                                       atol=1e-14 * _flux_scale(expect))


@pytest.mark.parametrize("nsd", [1, 2, 3])
@pytest.mark.parametrize("flux", [O.FLUX_HLL, O.FLUX_HLL3])
def test_flux_consistency(nsd, flux):
    """HLLTest.cpp:24-71: F(U, U) equals the point flux f(U) in every
    direction (for states where the point-flux branch is taken and for
    subsonic states where the HLL average of identical states reduces to f);
    f is the independent textbook flux."""
    states = [euler_state(1.0, [2.0, 3.0, 4.0][:nsd], 6.0),
              euler_state(0.7, [0.1, -0.2, 0.05][:nsd], 1.3),
              euler_state(2.0, [-3.5, 0.5, 1.0][:nsd], 0.4)]
    for q in states:
        for d in range(nsd):
            F, ws = O.face_flux(nsd, flux, d, GAMMA, q, q)
            f = np_point_flux(q, d)
            u = q[1 + d] / q[0]
            p = (GAMMA - 1) * (q[-1] - 0.5 * np.dot(q[1:-1], q[1:-1]) / q[0])
            c = np.sqrt(GAMMA * p / q[0])
            np.testing.assert_allclose(F, f.astype(float), rtol=1e-13,
                                       atol=1e-14 * _flux_scale(f))
            assert ws == pytest.approx(abs(u) + c, rel=1e-14)


def hllc_toro(qL, qR, d, gamma=GAMMA):
    """HLLC flux of Toro, Riemann Solvers, 3rd ed., section 10.4, evaluated
    in extended precision, with the reference's wave-speed estimates
    (HLL3.hpp:118-146: Roe-type averages with weights sqrt(rho), the fast
    speeds widened by 0.5*max(0, uL_d - uR_d)).

    Contact speed S* from Toro eq. 10.37; star states from eq. 10.39; star
    fluxes F*K = FK + SK (U*K - UK), eq. 10.38; the four-region selection of
    eq. 10.26. The one departure from the textbook is the reference's: the
    tangential velocity in both star states is that of the HLL intermediate
    state U_hll = (SR UR - SL UL + FL - FR)/(SR - SL) (Toro eq. 10.13) instead
    of each side's own; HLL3.hpp:80-82 forms it as (fR.m - fL.m - sR mR +
    sL mL)/aa. Returns (F, ws, info) with the branch facts the tests assert."""
    ld = np.longdouble
    g = ld(gamma)
    qL = np.asarray(qL, dtype=ld)
    qR = np.asarray(qR, dtype=ld)

    def prim(q):
        rho = q[0]
        u = q[1:-1] / rho
        p = (g - 1) * (q[-1] - ld(0.5) * rho * np.dot(u, u))
        return rho, u, p

    rL, uL, pL = prim(qL)
    rR, uR, pR = prim(qR)
    wL, wR = np.sqrt(rL), np.sqrt(rR)
    ubar = (wL * uL + wR * uR) / (wL + wR)
    pbar = (wL * pL + wR * pR) / (wL + wR)
    cbar = np.sqrt(g * pbar / ((rL + rR) / 2))
    corr = ld(0.5) * max(ld(0), uL[d] - uR[d])
    SL = min(uL[d] + corr - np.sqrt(g * pL / rL), ubar[d] - cbar)
    SR = max(uR[d] - corr + np.sqrt(g * pR / rR), ubar[d] + cbar)
    FL, FR = np_point_flux(qL, d, gamma), np_point_flux(qR, d, gamma)
    Sstar = ((pR - pL + rL * uL[d] * (SL - uL[d]) - rR * uR[d] * (SR - uR[d]))
             / (rL * (SL - uL[d]) - rR * (SR - uR[d])))
    Uhll = (SR * qR - SL * qL + FL - FR) / (SR - SL)
    vt = Uhll[1:-1] / Uhll[0]

    def star(q, rho, u, p, S):
        k = rho * (S - u[d]) / (S - Sstar)
        vel = vt.copy()
        vel[d] = Sstar
        U = np.empty(len(q), dtype=ld)
        U[0] = k
        U[1:-1] = k * vel
        U[-1] = k * (q[-1] / rho + (Sstar - u[d]) * (Sstar + p / (rho * (S - u[d]))))
        return U

    UsL = star(qL, rL, uL, pL, SL)
    UsR = star(qR, rR, uR, pR, SR)
    # the two star states average to the HLL state (Toro eq. 10.34's
    # integral condition): a check of this function, not of the oracle
    np.testing.assert_allclose((SR - Sstar) * UsR + (Sstar - SL) * UsL,
                               (SR - SL) * Uhll, rtol=1e-15, atol=1e-15)
    if SL >= 0:
        F = FL
    elif Sstar >= 0:
        F = FL + SL * (UsL - qL)
    elif SR > 0:
        F = FR + SR * (UsR - qR)
    else:
        F = FR
    info = {"subsonic": SL < 0 < SR, "sstar_nonneg": Sstar >= 0, "corr": corr,
            "scale": _flux_scale(FL, FR, SL * qL, SR * qR)}
    return F, max(abs(SL), abs(SR)), info


# (rho, u_normal, u_t1, u_t2, p) of the left and right states; for face
# direction d the normal velocity goes to component d and the tangential
# ones fill the other components in order
HLL3_PAIRS = [
    ((1.0, 0.3, 0.2, -0.1, 1.0), (0.125, -0.1, -0.3, 0.4, 0.1)),     # uL > uR, S* > 0
    ((0.125, 0.1, 0.5, 0.2, 0.1), (1.0, -0.3, -0.2, 0.3, 1.0)),      # uL > uR, S* < 0
    ((1.0, -0.4, 0.7, 0.1, 0.8), (1.2, 0.5, -0.6, -0.2, 1.1)),       # uL < uR, corr = 0
    ((2.0, 0.5, -0.5, 0.5, 2.5), (1.0, -0.5, 0.5, -0.5, 2.5)),       # shear + compression
    ((0.6, 0.2, 0.0, -0.8, 0.4), (1.7, 0.25, 0.9, 0.0, 1.9)),        # pR > pL, S* < 0
    ((1.5, 0.9, 0.3, 0.3, 1.2), (0.9, 0.1, -0.1, 0.6, 0.6)),         # strong compression, S* > 0
    ((0.8, -0.6, 0.4, -0.3, 0.9), (0.5, -0.9, 0.2, 0.7, 0.3)),       # leftward flow, uL > uR
]


def _directed_state(spec, nsd, d):
    rho, un, t1, t2, p = spec
    vel = [0.0] * nsd
    vel[d] = un
    tang = iter([t1, t2])
    for k in range(nsd):
        if k != d:
            vel[k] = next(tang)
    return euler_state(rho, vel, p)


@pytest.mark.parametrize("nsd", [1, 2, 3])
def test_hll3_hllc_textbook(nsd):
    """HLL3.hpp:52-148 on distinct subsonic states equals the HLLC flux of
    Toro section 10.4 (hllc_toro, extended precision) in every direction of
    1D, 2D and 3D, with both signs of the contact speed, the wave-speed
    correction active (uL_d > uR_d) and inactive, and non-zero tangential
    momentum in every direction. A sign error in the correction, the left
    star quantities in the right star branch or max(0, uR - uL) moves the
    flux by O(0.1), far outside the tolerance."""
    seen = {"pos": 0, "neg": 0, "corr": 0, "nocorr": 0}
    for d in range(nsd):
        for specL, specR in HLL3_PAIRS:
            qL = _directed_state(specL, nsd, d)
            qR = _directed_state(specR, nsd, d)
            F, ws = O.face_flux(nsd, O.FLUX_HLL3, d, GAMMA, qL, qR)
            expect, ws_expect, info = hllc_toro(qL, qR, d)
            assert info["subsonic"]
            seen["pos" if info["sstar_nonneg"] else "neg"] += 1
            seen["corr" if info["corr"] > 0 else "nocorr"] += 1
            np.testing.assert_allclose(F, expect.astype(float), rtol=1e-12,
                                       atol=1e-13 * info["scale"])
            assert ws == pytest.approx(float(ws_expect), rel=1e-14)
    assert min(seen.values()) >= 2 * nsd, seen


@pytest.mark.parametrize("nsd", [1, 2, 3])
@pytest.mark.parametrize("flux", [O.FLUX_HLL, O.FLUX_HLL3])
def test_flux_mirror_symmetry(nsd, flux):
    """Reflection invariance of the Euler equations: mirroring the axis d
    (normal momentum negated) and swapping the sides negates the mass,
    tangential-momentum and energy fluxes and keeps the normal-momentum flux
    and the wave speed. The two star branches of HLL3 map onto each other,
    so this checks the right-hand branch against the left-hand one."""
    for d in range(nsd):
        for specL, specR in HLL3_PAIRS:
            qL = _directed_state(specL, nsd, d)
            qR = _directed_state(specR, nsd, d)
            mL, mR = qL.copy(), qR.copy()
            mL[1 + d] = -mL[1 + d]
            mR[1 + d] = -mR[1 + d]
            F, ws = O.face_flux(nsd, flux, d, GAMMA, qL, qR)
            Fm, wsm = O.face_flux(nsd, flux, d, GAMMA, mR, mL)
            sign = -np.ones(len(F))
            sign[1 + d] = 1
            scale = _flux_scale(F, Fm, ws * qL, ws * qR)
            np.testing.assert_allclose(F, sign * Fm, rtol=1e-12, atol=1e-13 * scale)
            assert ws == pytest.approx(wsm, rel=1e-14)


@pytest.mark.parametrize("flux", [O.FLUX_HLL, O.FLUX_HLL3])
def test_flux_rotation_invariance(flux):
    """Rotational invariance: the flux in direction d of a 3D state pair is
    the direction-0 flux of the pair with axes d and 0 exchanged, with the
    momentum components of the result exchanged back."""
    for specL, specR in HLL3_PAIRS:
        qL0 = _directed_state(specL, 3, 0)
        qR0 = _directed_state(specR, 3, 0)
        F0, ws0 = O.face_flux(3, flux, 0, GAMMA, qL0, qR0)
        for d in (1, 2):
            perm = [0, 1, 2, 3, 4]
            perm[1], perm[1 + d] = perm[1 + d], perm[1]
            F, ws = O.face_flux(3, flux, d, GAMMA, qL0[perm], qR0[perm])
            scale = _flux_scale(F0, ws0 * qL0, ws0 * qR0)
            np.testing.assert_allclose(F[perm], F0, rtol=1e-13, atol=1e-14 * scale)
            assert ws == ws0


def test_godunov_closed_form():
    """burgers/Godunov.hpp:40-60: F = max(f(max(uL,0)), f(min(uR,0))) with
    f(u) = u^2/2; wave speed max(|uL|, |uR|)."""
    for uL, uR in [(1.0, 2.0), (-1.0, 0.5), (2.0, -3.0), (-0.5, -0.25), (0.0, 0.0)]:
        F, ws = O.face_flux(O.EQ_BURGERS, O.FLUX_GODUNOV, 0, GAMMA, [uL], [uR])
        expect = max(max(uL, 0) ** 2 / 2, min(uR, 0) ** 2 / 2)
        assert F[0] == expect
        assert ws == max(abs(uL), abs(uR))


def test_hll_supersonic_upwinding():
    """For a supersonic flow to the right (sL > 0) HLL/HLL3 return f(left);
    to the left (sR < 0) f(right)."""
    L = euler_state(1.0, [5.0], 1.0)
    R = euler_state(0.5, [4.0], 0.8)
    for flux in (O.FLUX_HLL, O.FLUX_HLL3):
        F, _ = O.face_flux(1, flux, 0, GAMMA, L, R)
        np.testing.assert_allclose(F, np_point_flux(L, 0).astype(float), rtol=1e-14)
        L2 = euler_state(1.0, [-5.0], 1.0)
        R2 = euler_state(0.5, [-6.0], 0.8)
        F2, _ = O.face_flux(1, flux, 0, GAMMA, L2, R2)
        np.testing.assert_allclose(F2, np_point_flux(R2, 0).astype(float), rtol=1e-14)


def test_hll_textbook_formula():
    """Subsonic HLL flux equals the Harten-Lax-van Leer formula (Toro 10.21)
    with the reference's Roe-type wave-speed estimates, evaluated in
    extended precision."""
    L = euler_state(1.0, [0.3], 1.0)
    R = euler_state(0.125, [-0.1], 0.1)
    F, ws = O.face_flux(1, O.FLUX_HLL, 0, GAMMA, L, R)
    ld = np.longdouble
    rl, rr = ld(L[0]), ld(R[0])
    ul, ur = ld(L[1]) / rl, ld(R[1]) / rr
    pl = (ld(GAMMA) - 1) * (ld(L[2]) - ld(0.5) * ld(L[1]) ** 2 / rl)
    pr = (ld(GAMMA) - 1) * (ld(R[2]) - ld(0.5) * ld(R[1]) ** 2 / rr)
    wl, wr = np.sqrt(rl), np.sqrt(rr)
    ub = (wl * ul + wr * ur) / (wl + wr)
    pb = (wl * pl + wr * pr) / (wl + wr)
    cb = np.sqrt(ld(GAMMA) * pb / ((rl + rr) / 2))
    sl = min(ul - np.sqrt(ld(GAMMA) * pl / rl), ub - cb)
    sr = max(ur + np.sqrt(ld(GAMMA) * pr / rr), ub + cb)
    fl = np.array([rl * ul, rl * ul * ul + pl, (ld(L[2]) + pl) * ul])
    fr = np.array([rr * ur, rr * ur * ur + pr, (ld(R[2]) + pr) * ur])
    expect = (sr * fl - sl * fr + sr * sl * (R.astype(ld) - L.astype(ld))) / (sr - sl)
    np.testing.assert_allclose(F, expect.astype(float), rtol=1e-13)
    assert ws == pytest.approx(float(max(abs(sl), abs(sr))), rel=1e-14)


# ---------------------------------------------------------------- reconstruction

def _line(n, rho_overrides):
    q = np.zeros((5, n))
    q[0] = 1.0
    q[1:4] = 1.0
    q[4] = 10.0
    for i, v in rho_overrides.items():
        q[0, i] = v
    return q


def test_weno2_known_answer():
    """WenoTest.cpp:228-299: rho = {.,2,0,1,.} around cell 2 gives
    right = w0*0.5 + w1*(-1), left = w0~*(-0.5) + w1~*1 with
    w ~ d/(beta+eps)^2, d = (2/3, 1/3), beta = (1, 4) (tolerance 1e-8 as there)."""
    q = _line(14, {3: 2.0, 4: 0.0, 5: 1.0})  # 2 ghost cells: interior index 2 -> 4
    left, right = O.reconstruct_1d(O.RECON_WENO2, q)
    eps = 1e-8
    d0, d1 = 2 / 3, 1 / 3
    b0, b1 = 1.0, 4.0
    a0, a1 = d0 / (b0 + eps) ** 2, d1 / (b1 + eps) ** 2
    at0, at1 = d1 / (b0 + eps) ** 2, d0 / (b1 + eps) ** 2
    assert right[0, 4] == pytest.approx((a0 * 0.5 + a1 * -1) / (a0 + a1), abs=1e-8)
    assert left[0, 4] == pytest.approx((at0 * -0.5 + at1 * 1) / (at0 + at1), abs=1e-8)


def test_eno2_known_answer():
    """EnoTest.cpp:218-252: the same pattern gives right = 1/2, left = -1/2 exactly."""
    q = _line(14, {3: 2.0, 4: 0.0, 5: 1.0})
    left, right = O.reconstruct_1d(O.RECON_ENO2, q)
    assert right[0, 4] == 0.5
    assert left[0, 4] == -0.5


def _smooth_state(setup, seed):
    """An admissible smooth state (Euler: rho, p > 0) with the boundary's ghosts."""
    shape = O.volume_shape(setup)
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.linspace(0, 1, m) for m in shape[1:]], indexing="ij")

    def field(lo, hi):
        k = rng.integers(1, 3, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        f = np.prod([np.sin(2 * np.pi * k[a] * grids[a] + ph[a]) for a in range(3)], axis=0)
        return lo + (hi - lo) * 0.5 * (f + 1)

    u = np.zeros(shape)
    if setup.equation == O.EQ_BURGERS:
        u[0] = field(-1, 1)
    else:
        nsd = setup.equation
        rho = field(0.7, 1.8)
        vel = np.stack([field(-0.5, 0.5) for _ in range(nsd)])
        p = field(0.6, 2.0)
        u[0] = rho
        u[1:1 + nsd] = rho * vel
        u[-1] = p / (GAMMA - 1) + 0.5 * rho * (vel ** 2).sum(axis=0)
    return O.apply_boundary(setup, u)


REGION_SETUPS = [
    (O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.BC_NEUMANN, (23, 1, 1)),
    (O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.BC_PERIODIC, (17, 13, 1)),
    (O.EQ_EULER3, O.FLUX_HLL3, O.RECON_ENO3, O.BC_PERIODIC, (11, 9, 10)),
    (O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.BC_PERIODIC, (31, 1, 1)),
]


@pytest.mark.parametrize("setup_args", REGION_SETUPS, ids=["euler1", "euler2", "euler3", "burgers"])
def test_compute_flux_regions_compose_the_system_rhs(setup_args):
    """The oracle's computeFlux on a region (NumericalFluxCPU.cpp:170-221):
    the inner region and the 2*d side regions of ConservedSystem.cpp:31-103,
    applied one after the other to one output, give the whole system RHS bit
    for bit (each cell exactly once: a cell written twice would be summed
    twice, a missed one would keep the prior value), the wave speeds' max
    over the regions is the system's, and a region alone leaves every cell
    outside it at its prior value."""
    eq, flux, recon, bc, n = setup_args
    setup = O.make_setup(eq, flux, recon, O.INT_RK3, bc, n, (0, 0, 0), (1, 1, 1))
    u = _smooth_state(setup, 3)
    ref, ws_ref = O.system_rhs(setup, u)
    out = np.zeros_like(u)
    ws = np.zeros(3)
    regs = O.conserved_system_regions(setup)
    for start, end in regs:
        out, w = O.compute_flux(setup, u, start, end, out=out)
        ws = np.maximum(ws, w)
    assert np.array_equal(ref.view(np.uint64), out.view(np.uint64))
    assert np.array_equal(ws_ref, ws)
    whole, w_whole = O.compute_flux(setup, u)
    assert np.array_equal(ref.view(np.uint64), whole.view(np.uint64))
    # one side region alone: the cells it covers hold the RHS, the rest the prior
    prior = np.random.default_rng(4).standard_normal(u.shape)
    start, end = regs[-1]
    side, _ = O.compute_flux(setup, u, start, end, out=prior)
    changed = side != prior
    assert changed.any() and (~changed).any()
    assert np.array_equal(side[changed], ref[changed])


@pytest.mark.parametrize("recon", [O.RECON_WENO2, O.RECON_ENO2, O.RECON_NONE, O.RECON_ENO3])
def test_volume_reconstruction_equals_line_reconstruction(recon):
    """The oracle's performReconstruction on a 2D volume (each reconstruction's
    own index ranges) gives, on every interior row and interior cell, the
    values of the 1D line reconstruction that the known-answer tests above
    pin, in both directions."""
    setup = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, recon, O.INT_RK3, O.BC_PERIODIC, (14, 11, 1),
                         (0, 0, 0), (1, 1, 0))
    g = O.ghosts(setup)
    u = _smooth_state(setup, 8)
    for d in (0, 1):
        left, right = O.reconstruct(setup, u, d)
        lines = np.moveaxis(u, 3 - d, -1)[:, 0]  # (nvar, rows, cells) along d
        ll = np.moveaxis(left, 3 - d, -1)[:, 0]
        rr = np.moveaxis(right, 3 - d, -1)[:, 0]
        for r in range(g, lines.shape[1] - g):
            l1, r1 = O.reconstruct_1d(recon, np.ascontiguousarray(lines[:, r]))
            sl = slice(g, lines.shape[2] - g)
            assert np.array_equal(ll[:, r, sl], l1[:, sl]), (d, r)
            assert np.array_equal(rr[:, r, sl], r1[:, sl]), (d, r)


@pytest.mark.parametrize("recon", [O.RECON_NONE, O.RECON_WENO2, O.RECON_ENO2])
def test_reconstruction_of_constant(recon):
    """A constant state is reproduced (to a few ulp: the WENO2 weights are
    not exact fractions)."""
    q = np.full((3, 20), 1.7)
    left, right = O.reconstruct_1d(recon, q)
    g = 1 if recon == O.RECON_NONE else 2
    sl = slice(g - 1, 20 - (g - 1))
    np.testing.assert_allclose(left[:, sl], 1.7, rtol=4e-16)
    np.testing.assert_allclose(right[:, sl], 1.7, rtol=4e-16)


@pytest.mark.parametrize("recon,rate", [(O.RECON_WENO2, 1.9), (O.RECON_ENO2, 1.9),
                                        (O.RECON_NONE, 0.9), (O.RECON_MC, 1.99),
                                        (O.RECON_MM, 1.9)])
def test_reconstruction_convergence(recon, rate):
    """ReconstructionConvergenceTest.cpp:288-316: face values of cell averages
    of a smooth function converge at rate >= 1.9 (weno2, eno2), 1.99 (mc, the
    reference's double-precision bound), 0.9 (none); mm (not in the reference
    test) at >= 1.9."""
    errs = []
    Ns = [64, 128, 256, 512]
    for N in Ns:
        g = 1 if recon == O.RECON_NONE else 2
        h = 1.0 / N
        x0 = (np.arange(N + 2 * g) - g) * h
        # cell averages of sin(2 pi x) + 0.5 sin(4 pi x) + 2 (positive)
        F = lambda x: -np.cos(2 * np.pi * x) / (2 * np.pi) - 0.25 * np.cos(4 * np.pi * x) / (2 * np.pi) + 2 * x
        avg = (F(x0 + h) - F(x0)) / h
        f = lambda x: np.sin(2 * np.pi * x) + 0.5 * np.sin(4 * np.pi * x) + 2
        q = np.vstack([avg, avg])
        left, right = O.reconstruct_1d(recon, q)
        sl = slice(g, g + N)
        err = np.mean(np.abs(left[0, sl] - f(x0[sl]))) + np.mean(np.abs(right[0, sl] - f(x0[sl] + h)))
        errs.append(err)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] >= rate, rates


@pytest.mark.parametrize("order,recon,rate", [(3, O.RECON_ENO3, 2.95), (4, O.RECON_ENO4, 3.9)])
def test_eno3_eno4_convergence(order, recon, rate):
    """ReconstructionConvergenceTest.cpp:160-240 for eno3 with its double
    precision L1 bound 2.95 (line 314); eno4 (not in the reference test)
    at >= 3.9. Cell averages of sin(2 pi x) + 2 on a periodic line."""
    errs = []
    for k in range(5, 10):
        N = 1 << k
        h = 1.0 / N
        g = order
        x0 = (np.arange(N + 2 * g) - g) * h
        avg = (-np.cos(2 * np.pi * (x0 + h)) + np.cos(2 * np.pi * x0)) / (2 * np.pi * h) + 2
        left, right = O.reconstruct_1d(recon, np.vstack([avg, avg]))
        f = lambda x: np.sin(2 * np.pi * x) + 2
        sl = slice(g, g + N)
        errs.append(np.mean(np.abs(left[0, sl] - f(x0[sl]))) +
                    np.mean(np.abs(right[0, sl] - f(x0[sl] + h))))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] >= rate, rates


def test_weno3_fifth_order():
    """WENOCPU<3> with the linear weights 3/10, 3/5, 1/10 combines three
    third-order stencils into a fifth-order face value on smooth data (cell
    averages of sin(2 pi x) + 2; not in the reference's convergence test)."""
    errs = []
    for k in range(5, 10):
        N = 1 << k
        h = 1.0 / N
        x0 = (np.arange(N + 6) - 3) * h
        avg = (-np.cos(2 * np.pi * (x0 + h)) + np.cos(2 * np.pi * x0)) / (2 * np.pi * h) + 2
        left, right = O.reconstruct_1d(O.RECON_WENO3, np.vstack([avg, avg]))
        f = lambda x: np.sin(2 * np.pi * x) + 2
        sl = slice(3, 3 + N)
        errs.append(np.mean(np.abs(left[0, sl] - f(x0[sl]))) +
                    np.mean(np.abs(right[0, sl] - f(x0[sl] + h))))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] >= 4.9, rates


def test_weno3_mirror_symmetry():
    """The left face uses the mirrored linear weights and coefficient rows:
    reconstructing the reversed line swaps the left and right face values."""
    rng = np.random.default_rng(17)
    q = rng.random((2, 30)) + 1
    left, right = O.reconstruct_1d(O.RECON_WENO3, q)
    lr, rr = O.reconstruct_1d(O.RECON_WENO3, q[:, ::-1].copy())
    sl = slice(2, 28)
    np.testing.assert_allclose(left[:, sl], rr[:, ::-1][:, sl], rtol=1e-14)
    np.testing.assert_allclose(right[:, sl], lr[:, ::-1][:, sl], rtol=1e-14)


@pytest.mark.parametrize("order,recon", [(2, O.RECON_ENO2), (3, O.RECON_ENO3), (4, O.RECON_ENO4),
                                         (3, O.RECON_WENO3)])
def test_eno_reproduces_polynomials(order, recon):
    """Every candidate stencil of ENO-k reconstructs the face values of a
    polynomial of degree k-1 exactly from its cell averages, so the result is
    exact whichever stencil the differences select (rounding only); WENO3's
    convex combination of the ENO3 candidates likewise."""
    rng = np.random.default_rng(order)
    coef = rng.standard_normal(order)
    P = np.polynomial.Polynomial(coef)
    I = P.integ()
    n, h = 24, 0.1
    x0 = np.arange(n) * h
    avg = (I(x0 + h) - I(x0)) / h
    left, right = O.reconstruct_1d(recon, np.vstack([avg, avg]))
    sl = slice(order - 1, n - order + 1)
    np.testing.assert_allclose(left[0, sl], P(x0[sl]), rtol=0, atol=1e-12)
    np.testing.assert_allclose(right[0, sl], P(x0[sl] + h), rtol=0, atol=1e-12)


def test_eno3_compares_the_same_positions_at_every_level():
    """ENOCPU.cpp:143-158: at level 1 the reference compares the differences
    at i-1 and i again (not at i-2 and i-1 after a left shift). Cells
    (0, 1, 1.5, 3, 3.5) around i: level 0 |0.5| < |1.5| shifts once, level 1
    |1| < |-1| does not, so the stencil is (1, 1.5, 3) with rows 1 and 2:
    left = 1/3 + 5/6*1.5 - 1/6*3, right = -1/6 + 5/6*1.5 + 1/3*3 (the
    textbook choice, cells (0, 1, 1.5), would give 4/3 and 19/12)."""
    q = np.array([0.0, 0.0, 0.0, 1.0, 1.5, 3.0, 3.5, 3.5, 3.5])
    left, right = O.reconstruct_1d(O.RECON_ENO3, np.vstack([q]))
    assert left[0, 4] == pytest.approx(1 / 3 + 5 / 6 * 1.5 - 1 / 6 * 3, rel=1e-15)
    assert right[0, 4] == pytest.approx(-1 / 6 + 5 / 6 * 1.5 + 1 / 3 * 3, rel=1e-15)


def _euler3_cells(rho, u, p, gamma=GAMMA):
    """Conserved euler3 cells (5, n) from primitive rows (Euler::computeConserved)."""
    rho = np.asarray(rho, dtype=float)
    n = rho.size
    u = np.asarray(u, dtype=float)
    u = np.broadcast_to(u if u.ndim == 2 else u.reshape(-1, 1), (3, n))
    p = np.broadcast_to(np.asarray(p, dtype=float), (n,))
    q = np.zeros((5, n))
    q[0] = rho
    q[1:4] = rho * u
    q[4] = p / (gamma - 1) + 0.5 * rho * (u * u).sum(axis=0)
    return q


def test_wenof2_convergence_reference_case():
    """ReconstructionConvergenceTest.cpp:160-240 with its wenof2 bound 1.99
    (double precision): euler3 cells with rho the cell averages of
    sin(2 pi x) + 2, p = 2.5, u = 0; L1 error of the density face values."""
    errs = []
    for k in range(6, 11):
        N = 1 << k
        h = 1.0 / N
        x0 = (np.arange(N + 4) - 2) * h
        avg = (-np.cos(2 * np.pi * (x0 + h)) + np.cos(2 * np.pi * x0)) / (2 * np.pi * h) + 2
        left, right = O.reconstruct_1d(O.RECON_WENOF2, _euler3_cells(avg, 0.0, 2.5))
        f = lambda x: np.sin(2 * np.pi * x) + 2
        sl = slice(2, 2 + N)
        errs.append(np.mean(np.abs(left[0, sl] - f(x0[sl]))) +
                    np.mean(np.abs(right[0, sl] - f(x0[sl] + h))))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] >= 1.99, rates


def test_wenof2_at_rest_is_weno2_on_density():
    """With u = 0 and constant p the velocity factor is 1/2 and the density
    faces are WENO2's (the same weights; M -/+ dP/2 equals WENO2's
    combination algebraically: a few ulp apart); the pressure faces stay p,
    so E = p/(gamma-1) exactly and the momenta are zero."""
    rng = np.random.default_rng(4)
    rho = 1.0 + rng.random(40)
    q = _euler3_cells(rho, 0.0, 0.75)
    l5, r5 = O.reconstruct_1d(O.RECON_WENOF2, q)
    l2, r2 = O.reconstruct_1d(O.RECON_WENO2, q)
    sl = slice(2, 38)
    np.testing.assert_allclose(l5[0, sl], l2[0, sl], rtol=4e-15)
    np.testing.assert_allclose(r5[0, sl], r2[0, sl], rtol=4e-15)
    assert np.all(l5[1:4, sl] == 0) and np.all(r5[1:4, sl] == 0)
    assert np.all(l5[4, sl] == 0.75 / (GAMMA - 1)) and np.all(r5[4, sl] == 0.75 / (GAMMA - 1))


def test_wenof2_clamps_density_and_pressure_slopes():
    """A one-cell dip (10, 1, 10) in rho (or p) gives slopes -3 / +3 (weights
    2/3, 1/3 from equal smoothness), clamped to -/+1.9 times the middle value:
    both faces are 1 + 0.5 * 1.9."""
    q = _euler3_cells([10.0, 1.0, 10.0], 0.0, 1.0)
    left, right = O.reconstruct_1d(O.RECON_WENOF2, np.hstack([q[:, :1]] * 2 + [q] + [q[:, 2:]] * 2))
    assert left[0, 3] == 1 + 0.5 * 1.9 and right[0, 3] == 1 + 0.5 * 1.9
    q = _euler3_cells(1.0 + np.zeros(3), 0.0, [10.0, 1.0, 10.0])
    left, right = O.reconstruct_1d(O.RECON_WENOF2, np.hstack([q[:, :1]] * 2 + [q] + [q[:, 2:]] * 2))
    pf = 1 + 0.5 * 1.9
    assert left[4, 3] == pf / (GAMMA - 1) and right[4, 3] == pf / (GAMMA - 1)


def test_wenof2_velocity_scaling_caps_kinetic_energy():
    """rho = 1, u = (-1, 0, 1) in x, p = 1e-3: the velocity slope 1 would add
    kinetic energy LL = 1/8 > R = p/(gamma-1); the slope is scaled by
    0.5*sqrt(R/LL), so each face has kinetic energy exactly R (closed form):
    E_face = 2R, u_face = -/+ sqrt(2R). With p = 1 (LL < R) the scale is 1/2:
    u_face = -/+ 0.5."""
    for p, uf in ((1e-3, math.sqrt(2 * 1e-3 / (GAMMA - 1))), (1.0, 0.5)):
        q = _euler3_cells(np.ones(3), np.array([[-1.0, 0.0, 1.0], [0, 0, 0], [0, 0, 0]]), p)
        left, right = O.reconstruct_1d(O.RECON_WENOF2,
                                       np.hstack([q[:, :1]] * 2 + [q] + [q[:, 2:]] * 2))
        assert left[1, 3] == pytest.approx(-uf, rel=1e-14)
        assert right[1, 3] == pytest.approx(uf, rel=1e-14)
        R = p / (GAMMA - 1)
        assert right[4, 3] == pytest.approx(R + 0.5 * uf * uf, rel=1e-14)
        if p < 0.1:
            assert right[4, 3] == pytest.approx(2 * R, rel=1e-14)


def test_wenof2_euler1_equals_euler3_without_transverse_velocity():
    """The primitive slopes are per component: euler1 and euler3 cells with
    zero transverse velocity reconstruct the same rho, mx and E faces."""
    rng = np.random.default_rng(9)
    rho = 1 + rng.random(30)
    ux = rng.standard_normal(30)
    p = 1 + rng.random(30)
    q3 = _euler3_cells(rho, np.vstack([ux, 0 * ux, 0 * ux]), p)
    q1 = q3[[0, 1, 4]]
    l3, r3 = O.reconstruct_1d(O.RECON_WENOF2, q3)
    l1, r1 = O.reconstruct_1d(O.RECON_WENOF2, q1)
    sl = slice(2, 28)
    for a, b in ((l1, l3), (r1, r3)):
        assert np.array_equal(a[:, sl], b[[0, 1, 4]][:, sl])


@pytest.mark.parametrize("recon", [O.RECON_MC, O.RECON_MM])
def test_limiters_closed_forms(recon):
    """MC.hpp / MM.hpp with minmod.hpp: linear data give the exact face values
    m -/+ h/2; at a local extremum the slope is zero (left = right = middle);
    hand-computed values for a steep-then-shallow stencil (0, 1, 1.5):
    MC slope minmod(1, 0.75, 2) = 0.75, MM slope minmod(0.5, 1) = 0.5."""
    q = np.array([[0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0]]) * 0.25 + 3.0  # linear, h = 0.25
    left, right = O.reconstruct_1d(recon, q)
    for i in range(1, 6):
        assert left[0, i] == q[0, i] - 0.125 and right[0, i] == q[0, i] + 0.125
    q = np.array([[0.0, 0.0, 2.0, 5.0, 2.0, 0.0, 0.0]])  # extremum at index 3
    left, right = O.reconstruct_1d(recon, q)
    assert left[0, 3] == 5.0 and right[0, 3] == 5.0
    q = np.array([[-1.0, -1.0, 0.0, 1.0, 1.5, 1.5, 1.5]])
    left, right = O.reconstruct_1d(recon, q)
    slope = 0.75 if recon == O.RECON_MC else 0.5
    assert left[0, 3] == 1.0 - slope / 2 and right[0, 3] == 1.0 + slope / 2


@pytest.mark.parametrize("flux", [O.FLUX_CENTRAL, O.FLUX_RUSANOV, O.FLUX_ROE, O.FLUX_GODUNOV])
def test_burgers_fluxes_consistent(flux):
    """F(u, u) = f(u) = u^2/2 for every Burgers flux, and the wave speed is |u|."""
    for u in (-2.5, -0.3, 0.0, 0.7, 3.0):
        F, ws = O.face_flux(O.EQ_BURGERS, flux, 0, 1.4, [u], [u])
        assert F[0] == u * u / 2 and ws == abs(u)


def test_burgers_fluxes_known_answers():
    """Hand-evaluated: Rusanov(1, -1) = 0.5*(0.5 + 0.5) - 0.5*1*(-2) = 1.5;
    Roe upwinds on the sign of uL + uR: Roe(1, -1) = f(1) = 0.5,
    Roe(-1, 1) = f(max(-1, 0)) = 0, Roe(-2, -1) = f(min(-1, 0)) = 0.5;
    Central(1, 3) = 0.5*(0.5 + 4.5) = 2.5 with wave speed 3."""
    B = O.EQ_BURGERS
    assert O.face_flux(B, O.FLUX_RUSANOV, 0, 1.4, [1.0], [-1.0])[0][0] == 1.5
    assert O.face_flux(B, O.FLUX_ROE, 0, 1.4, [1.0], [-1.0])[0][0] == 0.5
    assert O.face_flux(B, O.FLUX_ROE, 0, 1.4, [-1.0], [1.0])[0][0] == 0.0
    assert O.face_flux(B, O.FLUX_ROE, 0, 1.4, [-2.0], [-1.0])[0][0] == 0.5
    F, ws = O.face_flux(B, O.FLUX_CENTRAL, 0, 1.4, [1.0], [3.0])
    assert F[0] == 2.5 and ws == 3.0


@pytest.mark.parametrize("nsd", [1, 2, 3])
def test_euler_central_flux(nsd):
    """Central.hpp: F = 0.5*(f(L) + f(R)) with f the textbook Euler flux
    (np_point_flux), wave speed max(|u_d| + sqrt(gamma p/rho)); F(U, U) = f(U)."""
    rng = np.random.default_rng(nsd)
    g = 1.4
    ld = np.longdouble
    for d in range(nsd):
        for _ in range(5):
            rho, p = rng.uniform(0.5, 2, 2)
            u = rng.uniform(-1, 1, (2, nsd))
            qL = euler_state(1.3 * rho, list(u[0]), p)
            qR = euler_state(rho, list(u[1]), 0.7 * p)
            F, ws = O.face_flux(nsd, O.FLUX_CENTRAL, d, g, qL, qR)
            fL = np_point_flux(qL, d, g)
            fR = np_point_flux(qR, d, g)
            expect = (ld(0.5) * (fL + fR)).astype(float)
            np.testing.assert_allclose(F, expect, rtol=1e-13, atol=1e-14 * _flux_scale(fL, fR))
            cL = abs(u[0][d]) + np.sqrt(g * p / (1.3 * rho))
            cR = abs(u[1][d]) + np.sqrt(g * (0.7 * p) / rho)
            assert ws == pytest.approx(max(cL, cR), rel=1e-14)
            F2, _ = O.face_flux(nsd, O.FLUX_CENTRAL, d, g, qL, qL)
            np.testing.assert_allclose(F2, fL.astype(float), rtol=1e-13,
                                       atol=1e-14 * _flux_scale(fL))


# ---------------------------------------------------------------- operator

@pytest.mark.parametrize("eq,flux,recon,n", [
    (O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, (24, 1, 1)),
    (O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, (16, 12, 1)),
    (O.EQ_EULER3, O.FLUX_HLL3, O.RECON_ENO2, (8, 7, 6)),
    (O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, (30, 1, 1)),
    (O.EQ_BURGERS, O.FLUX_CENTRAL, O.RECON_MC, (30, 1, 1)),
    (O.EQ_BURGERS, O.FLUX_RUSANOV, O.RECON_MM, (30, 1, 1)),
    (O.EQ_BURGERS, O.FLUX_ROE, O.RECON_NONE, (30, 1, 1)),
    (O.EQ_EULER2, O.FLUX_CENTRAL, O.RECON_MC, (12, 10, 1)),
    (O.EQ_EULER1, O.FLUX_HLL, O.RECON_MM, (24, 1, 1)),
])
def test_net_flux_zero_for_constant_state(eq, flux, recon, n):
    """numericalfluxtest.cpp:65-126: the divergence of a constant state is zero."""
    s = O.make_setup(eq, flux, recon, O.INT_FE, O.BC_PERIODIC, n, (0, 0, 0), (1, 1, 1))
    shape = O.volume_shape(s)
    u = np.zeros(shape)
    if eq == O.EQ_BURGERS:
        u[0] = 0.7
    else:
        st = euler_state(1.3, [0.2, -0.1, 0.3][:eq], 2.0)
        for k in range(shape[0]):
            u[k] = st[k]
    out, ws = O.system_rhs(s, u)
    g = O.ghosts(s)
    sl = tuple([slice(None)] + [slice(g, -g) if m > 1 else slice(None) for m in (n[2], n[1], n[0])])
    assert np.all(out[sl] == 0.0)
    assert ws[0] > 0


def _fill_distinct(s):
    shape = O.volume_shape(s)
    u = np.full(shape, np.nan)
    g = O.ghosts(s)
    nx, ny, nz = s.nx, s.ny, s.nz
    gy = g if ny > 1 else 0
    gz = g if nz > 1 else 0
    vals = np.arange(shape[0] * nx * ny * nz, dtype=float).reshape(shape[0], nz, ny, nx) + 1
    u[:, gz:gz + nz, gy:gy + ny, g:g + nx] = vals
    return u, vals, (g, gy, gz)


@pytest.mark.parametrize("bc", [O.BC_NEUMANN, O.BC_PERIODIC])
@pytest.mark.parametrize("n,recon", [((9, 1, 1), O.RECON_WENO2), ((7, 6, 1), O.RECON_WENO2),
                                     ((5, 4, 3), O.RECON_NONE)])
def test_boundary_ghost_values(bc, n, recon):
    """BoundaryTest.cpp: periodic ghost = interior at (p-G) mod N; Neumann
    ghost = mirror image across the boundary face; corners compose per axis."""
    s = O.make_setup(O.EQ_BURGERS, O.FLUX_GODUNOV, recon, O.INT_FE, bc, n, (0, 0, 0), (1, 1, 1))
    u, vals, gs = _fill_distinct(s)
    out = O.apply_boundary(s, u)

    def src(p, N, G):
        if G == 0:
            return 0
        if bc == O.BC_PERIODIC:
            return (p - G) % N
        i = p - G
        if i < 0:
            return -i - 1
        if i >= N:
            return 2 * N - 1 - i
        return i

    g, gy, gz = gs
    for z in range(out.shape[1]):
        for y in range(out.shape[2]):
            for x in range(out.shape[3]):
                assert out[0, z, y, x] == vals[0, src(z, n[2], gz), src(y, n[1], gy), src(x, n[0], g)]


def test_timestep_closed_form_and_adjusters():
    """Integrator.cpp:20-46: dt = cfl / sum_d(ws_d/dx_d); FixedInterval
    adjusters clamp to saved*interval - t (FixedIntervalWriter.cpp:45-53)."""
    ws = [2.0, 3.0, 0.0]
    dx = [0.5, 0.25, 1.0]
    dt = O.compute_timestep(ws, dx, 0.475, 0.0)
    assert dt == 0.475 / (2.0 / 0.5 + 3.0 / 0.25)
    assert O.compute_timestep(ws, dx, 0.475, 0.99, [1.0], [1]) == pytest.approx(0.01, abs=1e-15)
    assert O.compute_timestep(ws, dx, 0.475, 0.5, [0.25], [3]) == dt


# ---------------------------------------------------------------- whole solves

def exact_riemann(x, t, L, R, gamma=GAMMA):
    """Exact solution of the 1D Euler Riemann problem (Toro, ch. 4): density."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)

    def f(p, r, pk, c):
        if p > pk:
            A, B = 2 / ((gamma + 1) * r), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * c / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    lo, hi = 1e-8, 10.0
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if f(pm, rl, pl, cl) + f(pm, rr, pr, cr) + ur - ul > 0:
            hi = pm
        else:
            lo = pm
    p = 0.5 * (lo + hi)
    um = 0.5 * (ul + ur) + 0.5 * (f(p, rr, pr, cr) - f(p, rl, pl, cl))
    rho = np.empty_like(x)
    g1 = (gamma - 1) / (gamma + 1)
    for i, xi in enumerate(x):
        s = xi / t
        if s < um:  # left of contact
            if p > pl:
                S = ul - cl * np.sqrt((gamma + 1) / (2 * gamma) * p / pl + (gamma - 1) / (2 * gamma))
                rho[i] = rl if s < S else rl * (p / pl + g1) / (g1 * p / pl + 1)
            else:
                shl, stl = ul - cl, um - cl * (p / pl) ** ((gamma - 1) / (2 * gamma))
                if s < shl:
                    rho[i] = rl
                elif s > stl:
                    rho[i] = rl * (p / pl) ** (1 / gamma)
                else:
                    rho[i] = rl * (2 / (gamma + 1) + g1 / cl * (ul - s)) ** (2 / (gamma - 1))
        else:
            if p > pr:
                S = ur + cr * np.sqrt((gamma + 1) / (2 * gamma) * p / pr + (gamma - 1) / (2 * gamma))
                rho[i] = rr if s > S else rr * (p / pr + g1) / (g1 * p / pr + 1)
            else:
                shr, str_ = ur + cr, um + cr * (p / pr) ** ((gamma - 1) / (2 * gamma))
                if s > shr:
                    rho[i] = rr
                elif s < str_:
                    rho[i] = rr * (p / pr) ** (1 / gamma)
                else:
                    rho[i] = rr * (2 / (gamma + 1) - g1 / cr * (ur - s)) ** (2 / (gamma - 1))
    return rho


def test_sod_converges_to_exact_riemann_solution():
    """Sod (HLL, no reconstruction, forward Euler, CFL 0.9, T = 0.2): the L1
    density error against the exact Riemann solution decreases with N
    (first-order method: rate ~0.5-1 at discontinuities), and the run ends
    exactly at T."""
    from alsvinn_amd import initial_data as ic
    errs = []
    for N in (256, 512, 1024):
        s = O.make_setup(O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.INT_FE, O.BC_NEUMANN,
                         (N, 1, 1), (-1, 0, 0), (1, 0, 0), cfl=0.9, end_time=0.2)
        rc, out, t, steps, dts = O.run(s, ic.sod(N))
        assert rc == 0 and t == 0.2
        x = ic.midpoints(N, -1.0, 1.0)
        rho = out[0, 0, 0, 1:-1]
        errs.append(np.mean(np.abs(rho - exact_riemann(x, 0.2, (1, 0, 1), (0.125, 0, 0.1)))) * 2)
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.012
    assert np.log2(errs[1] / errs[2]) > 0.4


@pytest.mark.parametrize("integ,order", [(O.INT_FE, 1), (O.INT_RK2, 2), (O.INT_RK3, 3),
                                         (O.INT_RK4, 4)])
def test_integrator_temporal_order(integ, order):
    """IntegratorConvergenceTest.cpp:158-170 (FE/RK2/RK3/RK4 >= 0.9/1.9/2.9/3.9): on a
    fixed grid, the difference to a run with a 64x smaller CFL number shrinks
    like cfl^order."""
    from alsvinn_amd import initial_data as ic
    N = 64
    x = ic.midpoints(N, 0.0, 1.0)
    rho = 1 + 0.2 * np.sin(2 * np.pi * x)
    u0 = np.zeros((3, 1, 1, N + 4))
    st = ic.conserved_euler(rho, [0.5 + 0 * x], 1.0 + 0 * x, GAMMA)
    for k in range(3):
        u0[k, 0, 0, 2:-2] = st[k]

    def run(cfl):
        s = O.make_setup(O.EQ_EULER1, O.FLUX_HLL3, O.RECON_WENO2, integ, O.BC_PERIODIC,
                         (N, 1, 1), (0, 0, 0), (1, 0, 0), cfl=cfl, end_time=0.1)
        rc, out, t, steps, _ = O.run(s, u0)
        assert rc == 0 and t == pytest.approx(0.1, abs=1e-15)
        return out[0, 0, 0, 2:-2]

    ref = run(0.4 / 64)
    e1 = np.abs(run(0.4) - ref).mean()
    e2 = np.abs(run(0.2) - ref).mean()
    assert np.log2(e1 / e2) > order - 0.25


def test_kh_conserves_totals_and_x_invariance():
    """Periodic KH: sum(rho), sum(m), sum(E) are conserved to roundoff; with
    all perturbation amplitudes zero the flow stays x-invariant."""
    from alsvinn_amd import initial_data as ic
    n = 32
    X = O.stl_mersenne_uniform(1, 40)[0]
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                     (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=1.0, max_steps=30)
    u0 = ic.kelvin_helmholtz_2d(n, X)
    rc, out, t, steps, _ = O.run(s, u0)
    assert rc == 0 and steps == 30
    tot0 = u0[:, 0, 2:-2, 2:-2].sum(axis=(1, 2))
    tot1 = out[:, 0, 2:-2, 2:-2].sum(axis=(1, 2))
    np.testing.assert_allclose(tot1, tot0, rtol=1e-12, atol=1e-12 * n * n)
    Xz = X.copy()
    Xz[:20] = 0.0  # a = 0: no perturbation
    rc, out, t, steps, _ = O.run(s, ic.kelvin_helmholtz_2d(n, Xz))
    inner = out[:, 0, 2:-2, 2:-2]
    assert np.all(inner == inner[:, :, :1])


def test_constraint_violation_is_reported():
    """Simulator.cpp:180-189: a state with negative pressure THROWs (status 2)."""
    n = 16
    s = O.make_setup(O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.INT_FE, O.BC_NEUMANN,
                     (n, 1, 1), (0, 0, 0), (1, 0, 0), cfl=0.9, end_time=1.0)
    u = np.zeros(O.volume_shape(s))
    u[0] = 1.0
    u[2] = 2.5
    u[2, 0, 0, 4:9] = -100.0  # negative energy -> negative pressure
    assert O.obeys_constraints(s, u) == 0
    rc, *_ = O.run(s, u)
    assert rc == 2


# ---------------------------------------------------------------- statistics

def test_meanvar_closed_form():
    """MeanVariance.cpp:30-67: samples c + k, k = 0..M-1 -> mean c + (M-1)/2,
    variance (1/M normalisation) (M^2-1)/12."""
    M = 9
    base = np.linspace(-1, 1, 12).reshape(3, 4)
    samples = np.stack([base + k for k in range(M)])
    mean, var = O.meanvar(samples)
    np.testing.assert_allclose(mean, base + (M - 1) / 2, rtol=1e-14)
    np.testing.assert_allclose(var, np.full_like(base, (M * M - 1) / 12), rtol=1e-10)


def test_meanvar_within_an_ulp_of_runtime_pow():
    """The reference accumulates the second moment with libm's pow(u, 2) at
    run time (HostMemory::addPower) and subtracts pow(mean, 2.); the oracle
    squares. math.pow is libm's pow: the means agree exactly and the
    variances within 2 ulp of E[u^2] plus 1 ulp of mean^2."""
    rng = np.random.default_rng(11)
    M, n = 5, 400
    samples = rng.standard_normal((M, n)) * rng.uniform(0.1, 10, n)
    mean, var = O.meanvar(samples)
    s1 = [0.0] * n
    s2 = [0.0] * n
    for k in range(M):
        for i in range(n):
            s1[i] += samples[k, i]
            s2[i] += math.pow(samples[k, i], 2.0)
    m_ref = np.array([a / M for a in s1])
    e2 = np.array([b / M for b in s2])
    v_ref = np.array([e2[i] - math.pow(m_ref[i], 2.0) for i in range(n)])
    assert np.array_equal(mean, m_ref)
    assert np.all(np.abs(var - v_ref) <= 2 * np.spacing(e2) + np.spacing(m_ref ** 2))


def test_onepoint_moment_closed_forms():
    """OnePointMoment.cpp:30-45: samples c + k, k = 0..M-1 -> m1 = c + (M-1)/2
    (equal to meanvar's mean bit for bit), m2 = c^2 + c(M-1) + (M-1)(2M-1)/6,
    m3 = c^3 + 3c^2(M-1)/2 + c(M-1)(2M-1)/2 + M(M-1)^2/4."""
    M = 7
    c = np.linspace(-2, 3, 30).reshape(5, 6)
    samples = np.stack([c + k for k in range(M)])
    m1 = O.onepoint_moment(samples, 1)
    assert np.array_equal(m1, O.meanvar(samples)[0])
    np.testing.assert_allclose(O.onepoint_moment(samples, 2),
                               c ** 2 + c * (M - 1) + (M - 1) * (2 * M - 1) / 6, rtol=1e-13)
    np.testing.assert_allclose(O.onepoint_moment(samples, 3),
                               c ** 3 + 3 * c ** 2 * (M - 1) / 2 + c * (M - 1) * (2 * M - 1) / 2
                               + M * (M - 1) ** 2 / 4, rtol=1e-12, atol=1e-12)


def test_onepoint_moment_within_an_ulp_of_runtime_pow():
    """p = 2 squares, p = 3 calls libm's pow as the reference does: each term
    is within 1 ulp of math.pow (the same libm) for p = 2, equal for p = 3."""
    rng = np.random.default_rng(5)
    u = rng.standard_normal((1, 2000)) * 3
    for p, ulps in ((2, 1), (3, 0)):
        got = O.onepoint_moment(u, p)
        ref = np.array([math.pow(x, float(p)) for x in u[0]])
        assert np.all(np.abs(got - ref) <= ulps * np.spacing(np.abs(ref)))


def _with_ghosts(inner, g, dim, fill):
    """(tz, ty, tx) volume with `fill` in the ghost layers."""
    if dim == 1:
        out = np.full((1, 1, inner.shape[-1] + 2 * g), fill)
        out[0, 0, g:-g] = inner
    else:
        out = np.full((1, inner.shape[0] + 2 * g, inner.shape[1] + 2 * g), fill)
        out[0, g:-g, g:-g] = inner
    return out


@pytest.mark.parametrize("p", [1, 2, 3])
def test_total_variation_linear_closed_form(p):
    """HostMemory::getTotalVariation on u = a*i + b*j over the interior: every
    term is sqrt(a^2 + b^2)^p and there are (nx-1)(ny-1) of them (the first
    interior row and column only serve as neighbours); in 1D (nx-1)|a|^p.
    Ghost layers hold 1e9 and must not enter."""
    nx, ny, g, a, b = 12, 9, 2, 0.75, -1.5
    j, i = np.mgrid[0:ny, 0:nx]
    u2 = _with_ghosts(a * i + b * j, g, 2, 1e9)
    assert O.total_variation(u2, (g, g, 0), p) == pytest.approx(
        (nx - 1) * (ny - 1) * math.sqrt(a * a + b * b) ** p, rel=1e-13)
    u1 = _with_ghosts(a * np.arange(nx, dtype=float), g, 1, 1e9)
    assert O.total_variation(u1, (g, 0, 0), p) == pytest.approx((nx - 1) * abs(a) ** p, rel=1e-13)


def test_total_variation_steps_and_constants():
    """A unit step across x = s contributes one unit term per row from the
    second row on (ny-1 terms); across y = s one per column from the second
    column on (nx-1); a constant has zero variation; 3D is rejected."""
    nx, ny, g, s = 10, 7, 1, 4
    j, i = np.mgrid[0:ny, 0:nx]
    for p in (1, 2, 3):
        assert O.total_variation(_with_ghosts((i >= s) * 1.0, g, 2, -5.0), (g, g, 0), p) == ny - 1
        assert O.total_variation(_with_ghosts((j >= s) * 1.0, g, 2, -5.0), (g, g, 0), p) == nx - 1
    assert O.total_variation(_with_ghosts(np.full((ny, nx), 2.5), g, 2, 7.0), (g, g, 0), 1) == 0
    assert O.total_variation(np.zeros((4, 4, 4)), (0, 0, 0), 1) == -1


def test_bounded_variation_mean_over_samples():
    """BoundedVariation.cpp:38-53: the per-variable variations of the samples
    summed in order and divided by M: samples k*u give TV(u) * (M-1)/2 for
    p = 1."""
    nx, ny, g, M = 8, 6, 2, 5
    rng = np.random.default_rng(3)
    u = _with_ghosts(rng.random((ny, nx)), g, 2, 0.0)
    samples = np.stack([np.stack([k * u, -k * u]) for k in range(M)])
    bv = O.bounded_variation(samples, (g, g, 0), 1)
    tv = O.total_variation(u, (g, g, 0), 1)
    np.testing.assert_allclose(bv, [tv * (M - 1) / 2] * 2, rtol=1e-13)


def _structure_linear_expect(N, h, p, periodic):
    """StructureTestBoundary.cpp:255-290: u_i = i in 1D."""
    tot = 0.0
    for i in range(N):
        for j in (-h, h):
            if 0 <= i + j < N:
                tot += abs(-j) ** p
            elif i + j < 0:
                tot += abs(i - (N + (i + j))) ** p if periodic else abs(i) ** p
            else:
                tot += abs(i - (i + j - N)) ** p if periodic else abs(i - (N - 1)) ** p
    return tot / N


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("p", [1.0, 2.0, 3.0])
def test_structure_cube_linear_1d(periodic, p):
    N, H, g = 16, 8, 2
    u = np.zeros((1, 1, 1, N + 2 * g))
    u[0, 0, 0, g:-g] = np.arange(N)
    out = O.structure_cube(u, (N, 1, 1), g, H, p, periodic)
    for h in range(1, H):
        assert out[0, h] == pytest.approx(_structure_linear_expect(N, h, p, periodic), rel=4e-16 * 4)


def test_structure_constant_is_zero():
    u = np.full((2, 1, 10, 10), 3.25)
    assert np.all(O.structure_cube(u, (6, 6, 1), 2, 4, 2.0, True) == 0)
    assert np.all(O.structure_basic(u, (6, 6, 1), 2, 4, 2.0, 1, False) == 0)


def test_structure_cube_shell_counts():
    """The radius-h shell of the 2D (3D) cube holds 8h (24h^2+2) points: with
    u = 1 at one cell and 0 elsewhere on a large periodic grid, S_1(h) =
    2*shell(h)/N."""
    for n, dims in (((21, 21, 1), 2), ((11, 11, 11), 3)):
        g = 2
        shape = (1, n[2] + (2 * g if dims == 3 else 0), n[1] + 2 * g, n[0] + 2 * g)
        u = np.zeros(shape)
        zc = g + n[2] // 2 if dims == 3 else 0
        u[0, zc, g + n[1] // 2, g + n[0] // 2] = 1.0
        out = O.structure_cube(u, n, g, 4, 1.0, True)
        N = n[0] * n[1] * n[2]
        for h in range(1, 4):
            shell = 8 * h if dims == 2 else 24 * h * h + 2
            assert out[0, h] == pytest.approx(2 * shell / N, rel=1e-14)
