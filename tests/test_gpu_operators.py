"""The operators on a reference Volume (per-variable device memories) against
the oracle, bit for bit: NumericalFlux::computeFlux over the regions the
reference's ConservedSystem uses (inner + 2 per direction) and the whole
interior, Reconstruction::performReconstruction per direction with the
reference's index ranges, Boundary::applyBoundaryConditions, and the
statistics fed one sample at a time (equal to the ensemble entry points and
to the oracle).

Device memory is hipMalloc'd per variable (with unrelated allocations in
between, so the variables are not evenly spaced); nothing is copied through
the host between the operator calls.
"""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _case(name):
    print("CASE", name, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev(hip):
    from tests import hip_runtime
    hip_runtime.rt()
    return hip_runtime


class DevVolume:
    """A volume as separate hipMalloc allocations per variable."""

    def __init__(self, dev, u, n, g, precision=0):
        from alsvinn_amd import operators as ops
        dt = np.float64 if precision == 0 else np.float32
        self.keep = []
        self.t = []
        for v in range(u.shape[0]):
            self.keep.append(dev.DeviceArray(17 + 29 * v))
            a = dev.DeviceArray(u[v].size, dt)
            a.upload(u[v])
            self.t.append(a)
        self.shape = u.shape
        self.vol = ops.make_volume(u.shape[0], n, g, [x.address for x in self.t], precision)

    def host(self):
        return np.stack([x.download().astype(np.float64) for x in self.t]).reshape(self.shape)


def _bits_equal(ref, got, what):
    diff = np.flatnonzero(ref.view(np.uint64) != got.view(np.uint64))
    if diff.size:
        idx = np.unravel_index(diff[0], ref.shape)
        pytest.fail("%s: %d cells differ; first at %s: ref=%r got=%r"
                    % (what, diff.size, idx, ref[idx], got[idx]))


def _ghosts3(setup, g):
    return O.ghosts3(setup)


def _valid_state(setup, rng):
    """A random admissible state, smooth enough that every reconstruction
    (up to ENO4) keeps positive face densities and pressures (random modes
    plus 1e-3 noise), with the boundary's ghost cells."""
    shape = O.volume_shape(setup)
    grids = np.meshgrid(*[np.linspace(0, 1, m) for m in shape[1:]], indexing="ij")

    def field(lo, hi):
        f = np.zeros(shape[1:])
        for _ in range(3):
            k = rng.integers(1, 4, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            f += np.prod([np.sin(2 * np.pi * k[a] * grids[a] + ph[a]) for a in range(3)], axis=0)
        f = f / 3 + 1e-3 * rng.standard_normal(shape[1:])
        return lo + (hi - lo) * 0.5 * (f + 1)

    u = np.zeros(shape)
    if setup.equation == O.EQ_BURGERS:
        u[0] = field(-1, 1)
    else:
        nsd = setup.equation
        rho = field(0.6, 2.0)
        vel = np.stack([field(-0.6, 0.6) for _ in range(nsd)])
        p = field(0.5, 2.5)
        u[0] = rho
        u[1:1 + nsd] = rho * vel
        u[-1] = p / (GAMMA - 1) + 0.5 * rho * (vel ** 2).sum(axis=0)
    return O.apply_boundary(setup, u)


def _regions(setup, g):
    """The whole interior, then the computeFlux regions of
    ConservedSystem.cpp:31-103."""
    return [((0, 0, 0), (0, 0, 0))] + O.conserved_system_regions(setup)


FLUX_CASES = [
    ("sod euler1 hll none", O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.BC_NEUMANN, (70, 1, 1)),
    ("euler1 hll3 eno3 (wide)", O.EQ_EULER1, O.FLUX_HLL3, O.RECON_ENO3, O.BC_PERIODIC, (90, 1, 1)),
    ("kh euler2 hll3 weno2", O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.BC_PERIODIC, (75, 41, 1)),
    ("euler2 hll wenof2", O.EQ_EULER2, O.FLUX_HLL, O.RECON_WENOF2, O.BC_NEUMANN, (64, 20, 1)),
    ("euler2 hll3 mc", O.EQ_EULER2, O.FLUX_HLL3, O.RECON_MC, O.BC_PERIODIC, (33, 30, 1)),
    ("euler2 central eno4 (wide)", O.EQ_EULER2, O.FLUX_CENTRAL, O.RECON_ENO4, O.BC_NEUMANN, (40, 26, 1)),
    ("euler2 hll3 weno3 (wide)", O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO3, O.BC_PERIODIC, (45, 23, 1)),
    ("euler3 hll3 weno2", O.EQ_EULER3, O.FLUX_HLL3, O.RECON_WENO2, O.BC_PERIODIC, (36, 13, 11)),
    ("euler3 hll eno2 neumann", O.EQ_EULER3, O.FLUX_HLL, O.RECON_ENO2, O.BC_NEUMANN, (20, 9, 14)),
    ("euler3 hll3 eno3 (wide)", O.EQ_EULER3, O.FLUX_HLL3, O.RECON_ENO3, O.BC_PERIODIC, (21, 10, 9)),
    ("burgers godunov eno2", O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.BC_PERIODIC, (150, 1, 1)),
    ("burgers 2d rusanov mm", O.EQ_BURGERS, O.FLUX_RUSANOV, O.RECON_MM, O.BC_NEUMANN, (50, 31, 1)),
]


@pytest.mark.parametrize("case", FLUX_CASES, ids=[c[0] for c in FLUX_CASES])
def test_compute_flux_regions_bit_exact(hip, dev, case):
    """computeFlux on per-variable device memories for the whole interior
    and each ConservedSystem region: the region's cells bit-identical to the
    oracle's NumericalFlux::computeFlux, every other cell of the output
    untouched, the wave speeds bit-identical."""
    name, eq, flux, recon, bc, n = case
    _case("computeFlux " + name)
    from alsvinn_amd import operators as ops
    setup = O.make_setup(eq, flux, recon, O.INT_RK3, bc, n, (0, 0, 0), (1, 1, 1))
    g = O.ghosts(setup)
    rng = np.random.default_rng([ord(ch) for ch in name])
    u = _valid_state(setup, rng)
    scheme = hip.make_scheme(eq, flux, recon, hip.INT_RK3, bc, gamma=GAMMA)
    U = DevVolume(dev, u, n, _ghosts3(setup, g))
    with ops.NumericalFlux(scheme, n) as nf:
        assert nf.get_number_of_ghost_cells() == g
        for start, end in _regions(setup, g):
            prior = rng.standard_normal(u.shape)
            OUT = DevVolume(dev, prior, n, _ghosts3(setup, g))
            ws = nf.compute_flux(U.vol, OUT.vol, start, end)
            ref, ws_ref = O.compute_flux(setup, u, start, end, out=prior)
            assert np.all(np.isfinite(ref)), "test state too rough for %s" % name
            _bits_equal(ref, OUT.host(), "%s region %s %s" % (name, start, end))
            _bits_equal(ws_ref, ws, "%s wave speeds %s %s" % (name, start, end))
    _bits_equal(u, U.host(), "input unchanged")


def test_conserved_system_from_regions_equals_ensemble_rhs(hip, dev):
    """The reference's ConservedSystem (inner region, then both sides of each
    direction, wave speeds the max over regions) composed from the Volume
    operator equals the ensemble's fused operator and the oracle's system
    RHS, bit for bit, FP64 and FP32."""
    _case("ConservedSystem from computeFlux regions")
    from alsvinn_amd import operators as ops
    n = (96, 70, 1)
    setup = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC, n,
                         (0, 0, 0), (1, 1, 0))
    g = O.ghosts(setup)
    u = _valid_state(setup, np.random.default_rng(5))
    ref, ws_ref = O.system_rhs(setup, u)
    for prec in (hip.FP64, hip.FP32):
        scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                                 gamma=GAMMA, precision=prec)
        U = DevVolume(dev, u, n, (g, g, 0), prec)
        OUT = DevVolume(dev, np.zeros_like(u), n, (g, g, 0), prec)
        with ops.NumericalFlux(scheme, n) as nf:
            regs = _regions(setup, g)[1:]
            ws = nf.compute_flux(U.vol, OUT.vol, *regs[0])
            side = np.zeros(3)
            for start, end in regs[1:]:
                side = np.maximum(side, nf.compute_flux(U.vol, OUT.vol, start, end))
            ws = np.maximum(ws, side)
        got = OUT.host()
        with hip.Ensemble(scheme, [hip.make_sample(n, (0, 0, 0), (1, 1, 0), 1.0)]) as e:
            e.upload(0, u)
            fused, ws_fused = e.rhs(0)
        sl = (slice(None), slice(None), slice(g, -g), slice(g, -g))
        _bits_equal(np.ascontiguousarray(fused[sl]), np.ascontiguousarray(got[sl]),
                    "regions vs fused prec=%d" % prec)
        _bits_equal(np.asarray(ws_fused, dtype=np.float64), ws, "wave speeds vs fused")
        if prec == hip.FP64:
            _bits_equal(ref, got, "regions vs oracle ConservedSystem")
            _bits_equal(ws_ref, ws, "wave speeds vs oracle")


RECON_CASES = [
    ("none", O.EQ_EULER2, O.RECON_NONE, (30, 17, 1)),
    ("weno2", O.EQ_EULER2, O.RECON_WENO2, (30, 17, 1)),
    ("wenof2", O.EQ_EULER3, O.RECON_WENOF2, (14, 9, 8)),
    ("eno2", O.EQ_BURGERS, O.RECON_ENO2, (120, 1, 1)),
    ("eno3", O.EQ_EULER2, O.RECON_ENO3, (25, 21, 1)),
    ("eno4", O.EQ_EULER1, O.RECON_ENO4, (60, 1, 1)),
    ("weno3", O.EQ_EULER3, O.RECON_WENO3, (15, 12, 9)),
    ("mc", O.EQ_EULER2, O.RECON_MC, (22, 19, 1)),
    ("mm", O.EQ_BURGERS, O.RECON_MM, (26, 24, 1)),
]


@pytest.mark.parametrize("case", RECON_CASES, ids=[c[0] for c in RECON_CASES])
def test_reconstruction_bit_exact(hip, dev, case):
    """performReconstruction in every direction, for the whole range and a
    side region, bit-identical to the oracle's reconstruction (its index
    ranges included: cells outside them keep their prior contents)."""
    name, eq, recon, n = case
    _case("reconstruction " + name)
    from alsvinn_amd import operators as ops
    flux = O.FLUX_GODUNOV if eq == O.EQ_BURGERS else O.FLUX_HLL3
    setup = O.make_setup(eq, flux, recon, O.INT_RK3, O.BC_PERIODIC, n, (0, 0, 0), (1, 1, 1))
    g = O.ghosts(setup)
    gv = _ghosts3(setup, g)
    rng = np.random.default_rng(11)
    u = _valid_state(setup, rng)
    scheme = hip.make_scheme(eq, flux, recon, hip.INT_RK3, hip.BC_PERIODIC, gamma=GAMMA)
    U = DevVolume(dev, u, n, gv)
    dim = 1 + (n[1] > 1) + (n[2] > 1)
    with ops.NumericalFlux(scheme, n) as nf:
        for d in range(dim):
            end = [0, 0, 0]
            end[d] = -(n[d] // 2)
            for start, stop in [((0, 0, 0), (0, 0, 0)), ((0, 0, 0), tuple(end))]:
                pl = rng.standard_normal(u.shape)
                pr = rng.standard_normal(u.shape)
                L = DevVolume(dev, pl, n, gv)
                R = DevVolume(dev, pr, n, gv)
                nf.perform_reconstruction(U.vol, d, L.vol, R.vol, start, stop)
                rl, rr = O.reconstruct(setup, u, d, start, stop, left=pl, right=pr)
                _bits_equal(rl, L.host(), "%s left d=%d %s" % (name, d, stop))
                _bits_equal(rr, R.host(), "%s right d=%d %s" % (name, d, stop))
        with pytest.raises(hip.HipError):
            nf.perform_reconstruction(U.vol, 0, U.vol, U.vol, indicator_variable=1)


@pytest.mark.parametrize("bc", [O.BC_NEUMANN, O.BC_PERIODIC])
@pytest.mark.parametrize("n,g", [((40, 1, 1), 2), ((23, 17, 1), 2), ((11, 9, 7), 3)])
def test_boundary_bit_exact(hip, dev, bc, n, g):
    """applyBoundaryConditions on per-variable memories: every ghost cell
    (corners included) equal to the oracle's boundary; interior untouched."""
    _case("boundary bc=%d n=%s g=%d" % (bc, n, g))
    from alsvinn_amd import operators as ops
    recon = O.RECON_ENO3 if g == 3 else O.RECON_WENO2
    eq = O.EQ_EULER1 + (n[1] > 1) + (n[2] > 1)
    setup = O.make_setup(eq, O.FLUX_HLL3, recon, O.INT_RK3, bc, n, (0, 0, 0), (1, 1, 1))
    assert O.ghosts(setup) == g
    u = np.random.default_rng(3).standard_normal(O.volume_shape(setup))
    V = DevVolume(dev, u, n, _ghosts3(setup, g))
    ops.apply_boundary_conditions(bc, V.vol)
    _bits_equal(O.apply_boundary(setup, u), V.host(), "boundary")


def _stat_samples(hip, rng, M, n, scheme):
    with hip.Ensemble(scheme, [hip.make_sample(n, (0, 0, 0), (1, 1, 1), 1.0) for _ in range(M)]) as e:
        for i in range(M):
            e.upload(i, rng.standard_normal(e.volume_shape(i)) + 2.0)
        e.apply_boundary()
        return [e.download(i) for i in range(M)]


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_statistics_per_sample_equal_ensemble_and_oracle(hip, dev, dim):
    """Statistics fed one sample at a time from per-variable device memories
    (compute per sample, then combine and finalize) equal the ensemble entry
    points over the same samples bit for bit, and the oracle: meanvar and
    onepointmoment (p = 1, 2) bit for bit, p = 3 within 4 ulp per term (the
    correctly rounded power against libm's pow), the structure functions
    within 1e-12 relative (parallel sums), bv (p = 1) bit for bit, p = 3
    within 1e-14. The wide
    structure_cube (the untiled kernel: H = 40 in 1D/2D) uses H = 10 in 3D,
    where 40 shells make ~4e8 sequential additions in the oracle, whose own
    rounding then reaches ~1e-11."""
    _case("per-sample statistics dim=%d" % dim)
    from alsvinn_amd import operators as ops
    rng = np.random.default_rng(20 + dim)
    M = 5
    if dim == 1:
        scheme = hip.make_scheme(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2, O.BC_PERIODIC)
        n = (131, 1, 1)
    elif dim == 2:
        scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
        n = (37, 29, 1)
    else:
        scheme = hip.make_scheme(O.EQ_EULER3, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_NEUMANN)
        n = (12, 9, 7)
    bc = scheme.boundary
    vols = _stat_samples(hip, rng, M, n, scheme)
    gv = (2, 2 if dim > 1 else 0, 2 if dim > 2 else 0)
    dev = [DevVolume(dev, v, n, gv) for v in vols]
    shape = dev[0].vol
    hw = 40 if dim < 3 else 10
    kinds = [("meanvar", hip._lib.STAT_MEANVAR, {}),
             ("m2", hip._lib.STAT_ONEPOINTMOMENT, {"p": 2.0}),
             ("m3", hip._lib.STAT_ONEPOINTMOMENT, {"p": 3.0}),
             ("cube8", hip._lib.STAT_STRUCTURE_CUBE, {"p": 2.0, "number_of_h": 8}),
             ("cubew", hip._lib.STAT_STRUCTURE_CUBE, {"p": 1.0, "number_of_h": hw}),
             ("basic", hip._lib.STAT_STRUCTURE_BASIC, {"p": 2.0, "number_of_h": 9,
                                                       "direction": dim - 1})]
    if dim < 3:
        kinds.append(("bv1", hip._lib.STAT_BV, {"p": 1.0}))
        kinds.append(("bv3", hip._lib.STAT_BV, {"p": 3.0}))
    got = {}
    for name, kind, kw in kinds:
        with ops.Statistics(kind, shape, boundary=bc, **kw) as st:
            for _ in range(2):  # a reset starts a new snapshot
                st.reset()
                for d in dev:
                    st.compute_statistics(d.vol)
                st.combine_statistics(M)
                st.finalize_statistics()
                got[name] = [st.download(i) for i in range(st.outputs)]
    with hip.Ensemble(scheme, [hip.make_sample(n, (0, 0, 0), (1, 1, 1), 1.0) for _ in range(M)]) as e:
        for i, v in enumerate(vols):
            e.upload(i, v)
        ens = {"meanvar": e.meanvar(), "m2": e.onepoint_moment(2), "m3": e.onepoint_moment(3),
               "cube8": e.structure_cube(8, 2.0), "cubew": e.structure_cube(hw, 1.0),
               "basic": e.structure_basic(9, dim - 1, 2.0)}
        if dim < 3:
            ens["bv1"] = e.bounded_variation(1)
            ens["bv3"] = e.bounded_variation(3)
    mean, var = ens["meanvar"]
    _bits_equal(mean.ravel(), got["meanvar"][0], "mean vs ensemble")
    _bits_equal(var.ravel(), got["meanvar"][1], "variance vs ensemble")
    rm, rv = O.meanvar(np.stack(vols))
    _bits_equal(rm.ravel(), got["meanvar"][0], "mean vs oracle")
    _bits_equal(rv.ravel(), got["meanvar"][1], "variance vs oracle")
    _bits_equal(ens["m2"].ravel(), got["m2"][0], "m2 vs ensemble")
    _bits_equal(O.onepoint_moment(np.stack(vols), 2).ravel(), got["m2"][0], "m2 vs oracle")
    _bits_equal(ens["m3"].ravel(), got["m3"][0], "m3 vs ensemble")
    scale = O.onepoint_moment(np.abs(np.stack(vols)), 3).ravel()
    assert np.all(np.abs(got["m3"][0] - O.onepoint_moment(np.stack(vols), 3).ravel())
                  <= 4 * np.spacing(scale)), "m3 vs oracle"
    periodic = bc == O.BC_PERIODIC
    for key, H, p in (("cube8", 8, 2.0), ("cubew", hw, 1.0)):
        _bits_equal(ens[key].ravel(), got[key][0], key + " vs ensemble")
        ref = sum(O.structure_cube(v, n, 2, H, p, periodic) for v in vols) / M
        np.testing.assert_allclose(got[key][0], ref.ravel(), rtol=1e-12, atol=0)
    _bits_equal(ens["basic"].ravel(), got["basic"][0], "basic vs ensemble")
    refb = sum(O.structure_basic(v, n, 2, 9, 2.0, dim - 1, periodic) for v in vols) / M
    np.testing.assert_allclose(got["basic"][0], refb.ravel(), rtol=1e-12, atol=0)
    if dim < 3:
        _bits_equal(ens["bv1"].ravel(), got["bv1"][0], "bv vs ensemble")
        _bits_equal(O.bounded_variation(np.stack(vols), gv, 1), got["bv1"][0], "bv vs oracle")
        _bits_equal(ens["bv3"].ravel(), got["bv3"][0], "bv3 vs ensemble")
        np.testing.assert_allclose(got["bv3"][0], O.bounded_variation(np.stack(vols), gv, 3),
                                   rtol=1e-14, atol=0)


def test_operator_argument_errors(hip, dev):
    """Shape, ghost, region and aliasing errors are reported as the
    reference's THROW would be (a HipError with a message), before any
    kernel runs."""
    _case("operator argument errors")
    from alsvinn_amd import operators as ops
    n = (32, 16, 1)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    u = np.ones((4, 20, 36))
    U = DevVolume(dev, u, n, (2, 2, 0))
    OUT = DevVolume(dev, u, n, (2, 2, 0))
    thin = DevVolume(dev, np.ones((4, 18, 34)), n, (1, 1, 0))
    with ops.NumericalFlux(scheme, n) as nf:
        for bad in [lambda: nf.compute_flux(U.vol, U.vol),
                    lambda: nf.compute_flux(thin.vol, OUT.vol),
                    lambda: nf.compute_flux(U.vol, OUT.vol, (-1, 0, 0), (0, 0, 0)),
                    lambda: nf.compute_flux(U.vol, OUT.vol, (0, 0, 1), (0, 0, 0)),
                    lambda: nf.perform_reconstruction(U.vol, 2, OUT.vol, OUT.vol)]:
            with pytest.raises(hip.HipError):
                bad()
    with pytest.raises(hip.HipError):
        ops.NumericalFlux(scheme, (1, 16, 1))
    with pytest.raises(hip.HipError):
        ops.Statistics(hip._lib.STAT_BV, ops.make_volume(5, (8, 8, 8), (2, 2, 2), [0] * 5))
