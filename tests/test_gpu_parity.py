"""Parity of the hip platform against the CPU oracle (FP64, bit-exact).

Every test runs the HIP kernels through the C ABI (alsvinn_amd -> libalsvinn_hip.so)
and the oracle (oracle/liboracle.so) on the same initial data, then compares
the conserved variables cell for cell (interior and ghost cells), the dt
sequence and the number of steps. Integer/index work and the CFL max are
exact; the FP64 arithmetic follows the reference's operation order, so the
expected difference is zero (the test tolerance is 0 ulp).
"""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("CASE", name, flush=True)
    sys.stdout.flush()


def _oracle_setup(eq, flux, recon, integ, bc, n, origin, top, cfl, T, saves=(1,), max_steps=-1):
    return O.make_setup(eq, flux, recon, integ, bc, n, origin, top, cfl=cfl, end_time=T,
                        saves=saves, max_steps=max_steps)


def _run_both(hip, eq, flux, recon, integ, bc, n, origin, top, cfl, T, u0, max_steps=-1,
              saves=(1,)):
    s = _oracle_setup(eq, flux, recon, integ, bc, n, origin, top, cfl, T, saves, max_steps)
    rc, ref, t_ref, steps_ref, dts_ref = O.run(s, u0)
    assert rc == 0
    scheme = hip.make_scheme(eq, flux, recon, integ, bc, gamma=1.4, cfl=cfl)
    sample = hip.make_sample(n, origin, top, T, saves=saves, max_steps=max_steps)
    with hip.Ensemble(scheme, [sample]) as e:
        e.record_dt(max(1, steps_ref + 1))
        e.upload(0, u0)
        e.apply_boundary()
        e.run(steps_per_sync=8)
        got = e.download(0)
        info = e.info(0)
        dts = e.dt_history(0)
    return ref, got, (t_ref, steps_ref, dts_ref), (info, dts)


def _assert_bit_exact(ref, got):
    assert ref.shape == got.shape
    diff = np.flatnonzero(ref.view(np.uint64) != got.view(np.uint64))
    if diff.size:
        i = diff[0]
        idx = np.unravel_index(i, ref.shape)
        rel = np.abs(ref - got).sum() / max(np.abs(ref).sum(), 1e-300)
        pytest.fail("%d cells differ; first at %s: ref=%r got=%r; rel L1=%.3e"
                    % (diff.size, idx, ref[idx], got[idx], rel))


def test_sod_fe_hll_bit_exact(hip):
    _print_case("sod 1024 hll none fe")
    from alsvinn_amd import initial_data as ic
    u0 = ic.sod(1024, ghosts=1)
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.INT_FE, O.BC_NEUMANN, (1024, 1, 1),
        (-1, 0, 0), (1, 0, 0), 0.9, 0.2, u0)
    assert info["steps"] == steps_ref
    assert info["t"] == t_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


@pytest.mark.parametrize("recon,flux,integ", [
    (O.RECON_WENO2, O.FLUX_HLL3, O.INT_RK3),
    (O.RECON_NONE, O.FLUX_HLL, O.INT_RK2),
    (O.RECON_ENO2, O.FLUX_HLL3, O.INT_FE),
])
def test_euler1_schemes_bit_exact(hip, recon, flux, integ):
    _print_case("euler1 sod 200 recon=%d flux=%d integ=%d" % (recon, flux, integ))
    from alsvinn_amd import initial_data as ic
    g = 1 if recon == O.RECON_NONE else 2
    u0 = ic.sod(200, ghosts=g)
    cfl = 0.9 if recon == O.RECON_NONE else 0.475
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER1, flux, recon, integ, O.BC_NEUMANN, (200, 1, 1), (-1, 0, 0), (1, 0, 0),
        cfl, 0.1, u0)
    assert info["steps"] == steps_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_kh_weno2_hll3_rk3_bit_exact(hip):
    _print_case("kh 64x64 weno2 hll3 rk3 periodic 40 steps")
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    X = sampling.stl_mersenne_uniform(1, 40)[0]
    u0 = ic.kelvin_helmholtz_2d(64, X, ghosts=2)
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC, (64, 64, 1),
        (0, 0, 0), (1, 1, 0), 0.475, 2.0, u0, max_steps=40)
    assert info["steps"] == steps_ref == 40
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_riemann2d_neumann_bit_exact(hip):
    _print_case("riemann4 80x72? no: 96x96 none hll fe neumann")
    from alsvinn_amd import initial_data as ic
    params = ic.riemann_parameters(np.linspace(0.05, 0.95, 16))
    u0 = ic.riemann_four_quadrant(96, params, ghosts=1)
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER2, O.FLUX_HLL, O.RECON_NONE, O.INT_FE, O.BC_NEUMANN, (96, 96, 1),
        (0, 0, 0), (1, 1, 0), 0.9, 0.05, u0)
    assert info["steps"] == steps_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_burgers_eno2_godunov_rk2_bit_exact(hip):
    _print_case("burgers 4096 eno2 godunov rk2 periodic")
    from alsvinn_amd import initial_data as ic
    u0 = ic.burgers_fbm(4096, 0.5, seed=7, ghosts=2)
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2, O.BC_PERIODIC,
        (4096, 1, 1), (0, 0, 0), (1, 0, 0), 0.475, 0.05, u0)
    assert info["steps"] == steps_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_rhs_operator_matches_conserved_system(hip):
    """One ConservedSystem::operator() evaluation on a non-trivial state:
    interior RHS and the per-direction max wave speeds, bit for bit."""
    _print_case("rhs operator kh 48x40")
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    X = sampling.stl_mersenne_uniform(2, 40)[1]
    n = (48, 40, 1)
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC, n,
                     (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=1.0)
    # KH on a non-square grid: build from a 48x48 field cut to 40 rows
    u_sq = ic.kelvin_helmholtz_2d(48, X, ghosts=2)
    u0 = np.ascontiguousarray(u_sq[:, :, :44, :])
    rng = np.random.default_rng(3)
    u0[0] *= 1.0 + 0.05 * rng.random(u0[0].shape)
    u0[1] += 0.1 * rng.standard_normal(u0[1].shape)
    u0 = O.apply_boundary(s, u0)
    out_ref, ws_ref = O.system_rhs(s, u0)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    with hip.Ensemble(scheme, [hip.make_sample(n, (0, 0, 0), (1, 1, 0), 1.0)]) as e:
        e.upload(0, u0)
        out, ws = e.rhs(0)
    g = 2
    _assert_bit_exact(out_ref[:, :, g:-g, g:-g], out[:, :, g:-g, g:-g])
    assert ws[0] == ws_ref[0] and ws[1] == ws_ref[1]


def test_ensemble_of_samples_equals_single_runs(hip):
    """Samples advanced together give the same bits as each run alone."""
    _print_case("ensemble 3 kh 32x32 vs single")
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    Xs = sampling.stl_mersenne_uniform(3, 40)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    samples = [hip.make_sample((32, 32, 1), (0, 0, 0), (1, 1, 0), 0.3) for _ in range(3)]
    u0s = [ic.kelvin_helmholtz_2d(32, X) for X in Xs]
    with hip.Ensemble(scheme, samples) as e:
        for i, u in enumerate(u0s):
            e.upload(i, u)
        e.apply_boundary()
        e.run(4)
        batch = [e.download(i) for i in range(3)]
        steps = [e.info(i)["steps"] for i in range(3)]
    for i in range(3):
        s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                         (32, 32, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=0.3)
        rc, ref, t, st, _ = O.run(s, u0s[i])
        assert rc == 0 and st == steps[i]
        _assert_bit_exact(ref, batch[i])


@pytest.mark.parametrize("n", [(76, 24, 1), (16, 16, 1)])
def test_packed_strips_bit_exact(hip, n):
    """Grids whose last strip is narrow: several samples' strips share one
    wavefront (76 = 60 + 16 columns; 16 columns = one narrow strip). Each
    sample must still match its own oracle run bit for bit."""
    _print_case("packed strips %s" % (n,))
    from alsvinn_amd import initial_data as ic
    rng = np.random.default_rng(11)
    M = 5
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    samples = [hip.make_sample(n, (0, 0, 0), (1, 1, 0), 0.2) for _ in range(M)]
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC, n,
                     (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=0.2)
    shape = O.volume_shape(s)
    u0s = []
    for i in range(M):
        u = np.zeros(shape)
        g = 2
        rho = 1.0 + 0.5 * rng.random((n[1], n[0]))
        vx = 0.3 * rng.standard_normal((n[1], n[0]))
        vy = 0.3 * rng.standard_normal((n[1], n[0]))
        p = 1.0 + 0.5 * rng.random((n[1], n[0]))
        cons = ic.conserved_euler(rho, [vx, vy], p, 1.4)
        for k in range(4):
            u[k, 0, g:-g, g:-g] = cons[k]
        u0s.append(u)
    with hip.Ensemble(scheme, samples) as e:
        for i, u in enumerate(u0s):
            e.upload(i, u)
        e.apply_boundary()
        e.run(5)
        got = [e.download(i) for i in range(M)]
        steps = [e.info(i)["steps"] for i in range(M)]
    for i in range(M):
        rc, ref, t, st, _ = O.run(s, u0s[i])
        assert rc == 0 and st == steps[i]
        _assert_bit_exact(ref, got[i])


@pytest.mark.parametrize("recon,flux,integ", [
    (O.RECON_WENO2, O.FLUX_HLL3, O.INT_RK3),
    (O.RECON_NONE, O.FLUX_HLL, O.INT_FE),
])
def test_kh3d_bit_exact(hip, recon, flux, integ):
    """3D Kelvin-Helmholtz (10x10 modes in x and z), 20^3, periodic."""
    _print_case("kh3d 20^3 recon=%d" % recon)
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    g = 1 if recon == O.RECON_NONE else 2
    n = 20
    X = sampling.stl_mersenne_uniform(1, 400)[0]
    u0 = ic.kelvin_helmholtz_3d(n, X, ghosts=g)
    cfl = 0.9 if recon == O.RECON_NONE else 0.475
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER3, flux, recon, integ, O.BC_PERIODIC, (n, n, n), (0, 0, 0), (1, 1, 1),
        cfl, 2.0, u0, max_steps=6)
    assert info["steps"] == steps_ref == 6
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_euler3_neumann_bit_exact(hip):
    """3D with Neumann boundaries on a non-cubic grid (ghost images on all axes)."""
    _print_case("euler3 neumann 22x12x9")
    rng = np.random.default_rng(5)
    n = (22, 12, 9)
    from alsvinn_amd import initial_data as ic
    s = O.make_setup(O.EQ_EULER3, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_NEUMANN, n,
                     (0, 0, 0), (1, 1, 1), cfl=0.475, end_time=1.0, max_steps=4)
    u = np.zeros(O.volume_shape(s))
    shp = (n[2], n[1], n[0])
    cons = ic.conserved_euler(1 + rng.random(shp), [0.3 * rng.standard_normal(shp) for _ in range(3)],
                              1 + rng.random(shp), 1.4)
    for k in range(5):
        u[k, 2:-2, 2:-2, 2:-2] = cons[k]
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER3, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_NEUMANN, n, (0, 0, 0),
        (1, 1, 1), 0.475, 1.0, u, max_steps=4)
    assert info["steps"] == 4
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_fp32_state_error_vs_fp64_oracle(hip):
    """FP32 state (FP32 reconstruction and fluxes, FP64 CFL): the relative L1
    error of rho against the FP64 oracle after 50 steps of KH 64^2 stays below
    1e-4 (the float rounding level amplified by the flow), and the dt sequence
    agrees to 1e-5 relative."""
    _print_case("fp32 kh 64x64 50 steps")
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    X = sampling.stl_mersenne_uniform(1, 40)[0]
    n = 64
    u0 = ic.kelvin_helmholtz_2d(n, X)
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                     (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=2.0, max_steps=50)
    rc, ref, t, steps, dts_ref = O.run(s, u0)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                             precision=hip.FP32)
    with hip.Ensemble(scheme, [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), 2.0,
                                               max_steps=50)]) as e:
        e.record_dt(60)
        e.upload(0, u0)
        e.apply_boundary()
        e.run(10)
        got = e.download(0)
        dts = e.dt_history(0)
    g = 2
    r = ref[0, 0, g:-g, g:-g]
    q = got[0, 0, g:-g, g:-g]
    l1 = np.abs(r - q).sum() / np.abs(r).sum()
    print("fp32 rho rel L1 vs fp64 oracle: %.3e" % l1)
    assert l1 < 1e-4
    assert np.allclose(dts, dts_ref, rtol=1e-5, atol=0)


def test_caller_owned_state_advanced_in_place(hip):
    """A state buffer owned by the caller (hipMalloc'd through the HIP runtime
    the library links) is advanced in place and gives the same bits as the
    oracle and as the library's own download; the buffer stays the caller's."""
    _print_case("caller-owned state KH 32^2 x 2")
    import ctypes
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    rt = ctypes.CDLL("libamdhip64.so.7")  # the runtime instance libalsvinn_hip.so loaded
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    n, T = 32, 0.05
    X = sampling.stl_mersenne_uniform(2, 40)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    samples = [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), T) for _ in range(2)]
    u0 = [ic.kelvin_helmholtz_2d(n, X[i]) for i in range(2)]
    total = hip.state_elements(scheme, samples)
    assert total == 2 * u0[0].size
    host = np.ascontiguousarray(np.concatenate([u.ravel() for u in u0]))
    nbytes = host.nbytes
    ptr = ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(ptr), nbytes) == 0
    try:
        assert rt.hipMemcpy(ptr, host.ctypes.data, nbytes, 1) == 0  # host to device
        with hip.Ensemble(scheme, samples, state=ptr.value) as e:
            e.apply_boundary()
            e.run(8)
            via_api = [e.download(i) for i in range(2)]
        got = np.empty_like(host)
        assert rt.hipMemcpy(got.ctypes.data, ptr, nbytes, 2) == 0  # device to host
    finally:
        rt.hipFree(ptr)
    got = got.reshape(2, *u0[0].shape)
    for i in range(2):
        s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                         (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=T)
        rc, ref, t, steps, _ = O.run(s, u0[i])
        assert rc == 0
        assert np.array_equal(ref.view(np.uint64), got[i].view(np.uint64))
        assert np.array_equal(ref.view(np.uint64), via_api[i].view(np.uint64))


@pytest.mark.parametrize("flux,recon,integ", [
    (O.FLUX_CENTRAL, O.RECON_MC, O.INT_RK2),
    (O.FLUX_RUSANOV, O.RECON_MM, O.INT_RK3),
    (O.FLUX_ROE, O.RECON_NONE, O.INT_FE),
    (O.FLUX_GODUNOV, O.RECON_MC, O.INT_RK2),
    (O.FLUX_ROE, O.RECON_MM, O.INT_RK3),
])
def test_burgers_fluxes_and_limiters_bit_exact(hip, flux, recon, integ):
    """The central, rusanov and roe fluxes (Central.hpp, Rusanov.hpp,
    burgers/Roe.hpp) and the mc / mm reconstructions (MC.hpp, MM.hpp) on a
    periodic 1D Burgers solve with shocks, against the oracle."""
    _print_case("burgers 1024 flux=%d recon=%d integ=%d" % (flux, recon, integ))
    from alsvinn_amd import initial_data as ic
    g = 1 if recon == O.RECON_NONE else 2
    u0 = ic.burgers_fbm(1024, 0.5, seed=11, ghosts=g)
    cfl = 0.9 if recon == O.RECON_NONE else 0.475
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_BURGERS, flux, recon, integ, O.BC_PERIODIC, (1024, 1, 1), (0, 0, 0),
        (1, 0, 0), cfl, 0.05, u0)
    assert info["steps"] == steps_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


@pytest.mark.parametrize("flux,recon,integ", [
    (O.FLUX_HLL, O.RECON_MC, O.INT_RK2),
    (O.FLUX_HLL3, O.RECON_MM, O.INT_RK3),
    (O.FLUX_CENTRAL, O.RECON_MC, O.INT_RK3),
])
def test_euler2_limiters_and_central_bit_exact(hip, flux, recon, integ):
    """mc / mm reconstructions and the central flux on the 2D Kelvin-Helmholtz
    state (periodic, 12 steps; the central flux has no dissipation, so only a
    few steps), against the oracle."""
    _print_case("kh 48^2 flux=%d recon=%d integ=%d" % (flux, recon, integ))
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    X = sampling.stl_mersenne_uniform(1, 40)[0]
    u0 = ic.kelvin_helmholtz_2d(48, X)
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER2, flux, recon, integ, O.BC_PERIODIC, (48, 48, 1), (0, 0, 0), (1, 1, 0),
        0.475, 2.0, u0, max_steps=12)
    assert info["steps"] == steps_ref == 12
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


def test_euler3_mc_grouped_kernel_bit_exact(hip):
    """mc reconstruction through the 3D workgroup kernel (Neumann, 10 rows so
    that the last group of four rows is partial), against the oracle."""
    _print_case("euler3 mc hll3 14x10x9 neumann")
    rng = np.random.default_rng(21)
    n = (14, 10, 9)
    s = O.make_setup(O.EQ_EULER3, O.FLUX_HLL3, O.RECON_MC, O.INT_RK2, O.BC_NEUMANN, n,
                     (0, 0, 0), (1, 1, 1), cfl=0.475, end_time=1.0, max_steps=6)
    shape = O.volume_shape(s)
    u0 = np.zeros(shape)
    u0[0] = 1.0 + 0.5 * rng.random(shape[1:])
    for k in (1, 2, 3):
        u0[k] = 0.3 * rng.standard_normal(shape[1:])
    u0[4] = 3.0 + rng.random(shape[1:])
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(
        hip, O.EQ_EULER3, O.FLUX_HLL3, O.RECON_MC, O.INT_RK2, O.BC_NEUMANN, n, (0, 0, 0),
        (1, 1, 1), 0.475, 1.0, u0, max_steps=6)
    assert info["steps"] == steps_ref == 6
    _assert_bit_exact(ref, got)


@pytest.mark.parametrize("eq", [O.EQ_EULER1, O.EQ_EULER2, O.EQ_EULER3])
def test_wenof2_bit_exact(hip, eq):
    """wenof2 (primitive-variable WENO2 with the density/pressure clamp and the
    velocity scaling) on Sod (1D, HLL3, SSP-RK3), the Kelvin-Helmholtz state
    (2D, periodic, HLL, SSP-RK3) and random 3D data (Neumann, HLL3, SSP-RK2),
    against the oracle."""
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    if eq == O.EQ_EULER1:
        _print_case("wenof2 euler1 sod 200")
        args = (O.FLUX_HLL3, O.RECON_WENOF2, O.INT_RK3, O.BC_NEUMANN, (200, 1, 1), (-1, 0, 0),
                (1, 0, 0), 0.475, 0.1, ic.sod(200, ghosts=2))
        steps = None
    elif eq == O.EQ_EULER2:
        _print_case("wenof2 euler2 kh 48^2")
        X = sampling.stl_mersenne_uniform(1, 40)[0]
        args = (O.FLUX_HLL, O.RECON_WENOF2, O.INT_RK3, O.BC_PERIODIC, (48, 48, 1), (0, 0, 0),
                (1, 1, 0), 0.475, 2.0, ic.kelvin_helmholtz_2d(48, X))
        steps = 20
    else:
        _print_case("wenof2 euler3 14x10x9 neumann")
        rng = np.random.default_rng(23)
        n = (14, 10, 9)
        s = O.make_setup(O.EQ_EULER3, O.FLUX_HLL3, O.RECON_WENOF2, O.INT_RK2, O.BC_NEUMANN, n,
                         (0, 0, 0), (1, 1, 1), cfl=0.475, end_time=1.0, max_steps=6)
        shape = O.volume_shape(s)
        u0 = np.zeros(shape)
        u0[0] = 1.0 + 0.5 * rng.random(shape[1:])
        for k in (1, 2, 3):
            u0[k] = 0.3 * rng.standard_normal(shape[1:])
        u0[4] = 3.0 + rng.random(shape[1:])
        args = (O.FLUX_HLL3, O.RECON_WENOF2, O.INT_RK2, O.BC_NEUMANN, n, (0, 0, 0), (1, 1, 1),
                0.475, 1.0, u0)
        steps = 6
    kw = {} if steps is None else {"max_steps": steps}
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(hip, eq, *args, **kw)
    assert info["steps"] == steps_ref
    if steps is not None:
        assert steps_ref == steps
    else:
        assert info["t"] == t_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


@pytest.mark.parametrize("recon", [O.RECON_ENO3, O.RECON_ENO4, O.RECON_WENO3])
@pytest.mark.parametrize("case", ["sod", "kh2d", "euler3", "burgers2d"])
def test_wide_stencil_reconstructions_bit_exact(hip, recon, case):
    """eno3 / eno4 / weno3 (3, 4, 3 ghost layers) through the per-cell wide-stencil
    kernel: Sod (1D, HLL3, SSP-RK3), Kelvin-Helmholtz (2D periodic, HLL3,
    SSP-RK3), random 3D data (periodic, HLL, SSP-RK2) and 2D Burgers (Neumann,
    Godunov, forward Euler), against the oracle."""
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    g = 4 if recon == O.RECON_ENO4 else 3
    _print_case("recon %d %s" % (recon, case))
    rng = np.random.default_rng(31 + g)
    steps = 10
    if case == "sod":
        eq = O.EQ_EULER1
        args = (O.FLUX_HLL3, recon, O.INT_RK3, O.BC_NEUMANN, (150, 1, 1), (-1, 0, 0), (1, 0, 0),
                0.475, 0.1, ic.sod(150, ghosts=g))
        steps = None
    elif case == "kh2d":
        eq = O.EQ_EULER2
        X = sampling.stl_mersenne_uniform(1, 40)[0]
        args = (O.FLUX_HLL3, recon, O.INT_RK3, O.BC_PERIODIC, (40, 40, 1), (0, 0, 0), (1, 1, 0),
                0.475, 2.0, ic.kelvin_helmholtz_2d(40, X, ghosts=g))
    elif case == "euler3":
        eq = O.EQ_EULER3
        n = (13, 10, 9)
        s = O.make_setup(eq, O.FLUX_HLL, recon, O.INT_RK2, O.BC_PERIODIC, n, (0, 0, 0),
                         (1, 1, 1), cfl=0.475, end_time=1.0, max_steps=6)
        shape = O.volume_shape(s)
        # random but mild data: ENO4 without limiting drives rougher fields to
        # negative pressure (the oracle stops as the reference does)
        u0 = np.zeros(shape)
        u0[0] = 1.0 + 0.2 * rng.random(shape[1:])
        for k in (1, 2, 3):
            u0[k] = 0.1 * rng.standard_normal(shape[1:])
        u0[4] = 3.0 + 0.3 * rng.random(shape[1:])
        args = (O.FLUX_HLL, recon, O.INT_RK2, O.BC_PERIODIC, n, (0, 0, 0), (1, 1, 1), 0.475,
                1.0, u0)
        steps = 6
    else:
        eq = O.EQ_BURGERS
        n = (30, 22, 1)
        s = O.make_setup(eq, O.FLUX_GODUNOV, recon, O.INT_FE, O.BC_NEUMANN, n, (0, 0, 0),
                         (1, 1, 0), cfl=0.475, end_time=1.0, max_steps=10)
        u0 = rng.standard_normal(O.volume_shape(s))
        args = (O.FLUX_GODUNOV, recon, O.INT_FE, O.BC_NEUMANN, n, (0, 0, 0), (1, 1, 0), 0.475,
                1.0, u0)
    kw = {} if steps is None else {"max_steps": steps}
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(hip, eq, *args, **kw)
    assert info["steps"] == steps_ref
    if steps is not None:
        assert steps_ref == steps
    else:
        assert info["t"] == t_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)


@pytest.mark.parametrize("eq,recon,bc,sizes", [
    (O.EQ_BURGERS, O.RECON_WENO2, O.BC_PERIODIC, [(2, 2, 1), (3, 5, 1), (61, 7, 1), (62, 3, 1),
                                                  (125, 2, 1)]),
    (O.EQ_EULER3, O.RECON_MC, O.BC_NEUMANN, [(2, 2, 2), (3, 2, 5), (61, 3, 2)]),
    (O.EQ_EULER1, O.RECON_ENO4, O.BC_PERIODIC, [(4, 1, 1), (5, 1, 1), (57, 1, 1), (200, 1, 1)]),
    (O.EQ_EULER2, O.RECON_WENO3, O.BC_PERIODIC, [(3, 3, 1), (40, 9, 1), (59, 4, 1), (121, 6, 1),
                                                 (59, 5, 1)]),
    (O.EQ_BURGERS, O.RECON_ENO3, O.BC_NEUMANN, [(3, 4, 1), (57, 6, 1), (58, 3, 1), (130, 5, 1)]),
])
def test_minimal_and_ragged_grids_mixed_end_times(hip, eq, recon, bc, sizes):
    """One ensemble of samples on the smallest grids the ghost layers allow,
    on grids whose strips are narrow or split across wavefronts, each with
    its own end time: every sample equals its own oracle run bit for bit, with
    the same number of steps and the same final time."""
    _print_case("edge eq=%d recon=%d sizes=%s" % (eq, recon, sizes))
    rng = np.random.default_rng(41 + eq)
    flux = O.FLUX_GODUNOV if eq == O.EQ_BURGERS else O.FLUX_HLL3
    integ = O.INT_RK3
    dim = 3 if eq == O.EQ_EULER3 else (1 if eq == O.EQ_EULER1 else 2)
    top = (1, 1 if dim >= 2 else 0, 1 if dim == 3 else 0)
    ends = [0.1 + 0.05 * k for k in range(len(sizes))]
    setups, u0s = [], []
    for n, T in zip(sizes, ends):
        s = O.make_setup(eq, flux, recon, integ, bc, n, (0, 0, 0), top, cfl=0.475, end_time=T)
        shape = O.volume_shape(s)
        if eq == O.EQ_BURGERS:
            u0 = 0.5 + rng.random(shape)
        else:
            u0 = np.zeros(shape)
            u0[0] = 1.0 + 0.2 * rng.random(shape[1:])
            for k in range(1, shape[0] - 1):
                u0[k] = 0.1 * rng.standard_normal(shape[1:])
            u0[-1] = 3.0 + 0.3 * rng.random(shape[1:])
        setups.append(s)
        u0s.append(u0)
    scheme = hip.make_scheme(eq, flux, recon, integ, bc, gamma=1.4, cfl=0.475)
    samples = [hip.make_sample(n, (0, 0, 0), top, T) for n, T in zip(sizes, ends)]
    with hip.Ensemble(scheme, samples) as e:
        for i, u in enumerate(u0s):
            e.upload(i, u)
        e.apply_boundary()
        e.run(4)
        got = [e.download(i) for i in range(len(sizes))]
        infos = [e.info(i) for i in range(len(sizes))]
    for i, s in enumerate(setups):
        rc, ref, t, steps, _ = O.run(s, u0s[i])
        assert rc == 0
        assert infos[i]["steps"] == steps and infos[i]["t"] == t == ends[i], (i, sizes[i])
        _assert_bit_exact(ref, got[i])


@pytest.mark.parametrize("recon", [O.RECON_WENO2, O.RECON_ENO3])
def test_schedule_does_not_change_results(hip, recon):
    """Every band height and register budget the autotuner can choose gives
    the same bits: three 2D Euler samples (two 48^2, one ragged 70 x 33),
    HLL3 + SSP-RK3, eight steps under each schedule, against the oracle. The
    reported waves per SIMD follow the budget: 1/2/3 on the marching kernel
    (WENO2), 2/3/4 on the wide-stencil kernel (ENO3)."""
    _print_case("schedules recon=%d" % recon)
    rng = np.random.default_rng(61)
    sizes = [(48, 48, 1), (48, 48, 1), (70, 33, 1)]
    setups, u0s, refs = [], [], []
    for n in sizes:
        s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, recon, O.INT_RK3, O.BC_PERIODIC, n, (0, 0, 0),
                         (1, 1, 0), cfl=0.475, end_time=2.0, max_steps=8)
        shape = O.volume_shape(s)
        u0 = np.zeros(shape)
        u0[0] = 1.0 + 0.2 * rng.random(shape[1:])
        u0[1] = 0.1 * rng.standard_normal(shape[1:])
        u0[2] = 0.1 * rng.standard_normal(shape[1:])
        u0[3] = 3.0 + 0.3 * rng.random(shape[1:])
        rc, ref, t, steps, _ = O.run(s, u0)
        assert rc == 0 and steps == 8
        setups.append(s)
        u0s.append(u0)
        refs.append(ref)
    waves = [2, 3, 4] if recon == O.RECON_ENO3 else [1, 2, 3]
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, recon, O.INT_RK3, O.BC_PERIODIC,
                             gamma=1.4, cfl=0.475)
    for rows in (0, 2, 16):
        for variant in (0, 1, 2):
            samples = [hip.make_sample(n, (0, 0, 0), (1, 1, 0), 2.0, max_steps=8) for n in sizes]
            with hip.Ensemble(scheme, samples) as e:
                for i, u in enumerate(u0s):
                    e.upload(i, u)
                e.apply_boundary()
                e.set_schedule(rows, variant)
                sch = e.schedule()
                assert sch["stage_waves_per_simd"] == [waves[variant]] * 3, sch
                e.run(4)
                got = [e.download(i) for i in range(len(sizes))]
                infos = [e.info(i) for i in range(len(sizes))]
            for i in range(len(sizes)):
                assert infos[i]["steps"] == 8, (rows, variant, i)
                _assert_bit_exact(refs[i], got[i])


@pytest.mark.parametrize("case", ["sod", "kh2d", "euler3", "burgers_eno3"])
def test_rk4_bit_exact(hip, case):
    """The classical RK4 integrator (RungeKutta4.cpp:50-96): three stage
    evaluations without the wave-speed reduction, the half and full stage
    updates and the five-term final combination, on the marching kernels (1D,
    2D, 3D) and the wide-stencil kernel, against the oracle."""
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    _print_case("rk4 %s" % case)
    rng = np.random.default_rng(51)
    steps = 12
    if case == "sod":
        eq = O.EQ_EULER1
        args = (O.FLUX_HLL, O.RECON_NONE, O.INT_RK4, O.BC_NEUMANN, (300, 1, 1), (-1, 0, 0),
                (1, 0, 0), 0.9, 0.2, ic.sod(300, ghosts=1))
        steps = None
    elif case == "kh2d":
        eq = O.EQ_EULER2
        X = sampling.stl_mersenne_uniform(1, 40)[0]
        args = (O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK4, O.BC_PERIODIC, (48, 48, 1), (0, 0, 0),
                (1, 1, 0), 0.475, 2.0, ic.kelvin_helmholtz_2d(48, X))
    elif case == "euler3":
        eq = O.EQ_EULER3
        n = (14, 10, 9)
        s = O.make_setup(eq, O.FLUX_HLL3, O.RECON_MC, O.INT_RK4, O.BC_PERIODIC, n, (0, 0, 0),
                         (1, 1, 1), cfl=0.475, end_time=1.0, max_steps=6)
        shape = O.volume_shape(s)
        u0 = np.zeros(shape)
        u0[0] = 1.0 + 0.2 * rng.random(shape[1:])
        for k in (1, 2, 3):
            u0[k] = 0.1 * rng.standard_normal(shape[1:])
        u0[4] = 3.0 + 0.3 * rng.random(shape[1:])
        args = (O.FLUX_HLL3, O.RECON_MC, O.INT_RK4, O.BC_PERIODIC, n, (0, 0, 0), (1, 1, 1), 0.475,
                1.0, u0)
        steps = 6
    else:
        eq = O.EQ_BURGERS
        n = (40, 30, 1)
        s = O.make_setup(eq, O.FLUX_GODUNOV, O.RECON_ENO3, O.INT_RK4, O.BC_PERIODIC, n, (0, 0, 0),
                         (1, 1, 0), cfl=0.475, end_time=1.0, max_steps=12)
        u0 = rng.standard_normal(O.volume_shape(s))
        args = (O.FLUX_GODUNOV, O.RECON_ENO3, O.INT_RK4, O.BC_PERIODIC, n, (0, 0, 0), (1, 1, 0),
                0.475, 1.0, u0)
    kw = {} if steps is None else {"max_steps": steps}
    ref, got, (t_ref, steps_ref, dts_ref), (info, dts) = _run_both(hip, eq, *args, **kw)
    assert info["steps"] == steps_ref
    if steps is not None:
        assert steps_ref == steps
    else:
        assert info["t"] == t_ref
    assert np.array_equal(dts, dts_ref)
    _assert_bit_exact(ref, got)
