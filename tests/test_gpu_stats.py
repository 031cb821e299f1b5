"""Statistics, mixed-size ensembles and full-size invariants on the GPU.

meanvar (reference formula), onepointmoment (p = 1, 2) and bv (p = 1, 2) are
compared bit for bit with the oracle's sequential loops; the Welford variant,
the structure functions (parallel reductions) and p = 3 (correctly rounded
power against libm's pow) within a stated tolerance.
"""
import hashlib
import sys
from fractions import Fraction

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

RTOL_REDUCTION = 1e-12  # parallel vs sequential summation of O(1e5) terms


def _case(name):
    print("CASE", name, flush=True)
    sys.stdout.flush()


def _kh_ensemble(hip, sizes, T=0.05, seed_start=0):
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    X = sampling.stl_mersenne_uniform(len(sizes), 40, sample_start=seed_start)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    samples = [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), T) for n in sizes]
    u0 = [ic.kelvin_helmholtz_2d(n, X[i]) for i, n in enumerate(sizes)]
    return scheme, samples, u0


def _oracle_final(n, u0, T):
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                     (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=T)
    rc, out, t, steps, _ = O.run(s, u0)
    assert rc == 0 and t == T
    return out, steps


def test_mixed_size_ensemble_and_meanvar_bit_exact(hip):
    """An ensemble mixing 16^2, 24^2 and 40^2 samples: every sample equals its
    own oracle run, and meanvar per size group equals the oracle's sequential
    accumulation (MeanVariance.cpp:30-67) bit for bit."""
    _case("mixed sizes 16/24/40 x 3")
    sizes = [16, 24, 40, 16, 24, 40, 16, 24, 40]
    T = 0.05
    scheme, samples, u0 = _kh_ensemble(hip, sizes, T)
    with hip.Ensemble(scheme, samples) as e:
        for i, u in enumerate(u0):
            e.upload(i, u)
        e.apply_boundary()
        e.run(8)
        finals = [e.download(i) for i in range(len(sizes))]
        infos = [e.info(i) for i in range(len(sizes))]
        stats = {n: e.meanvar([i for i, m in enumerate(sizes) if m == n]) for n in (16, 24, 40)}
        welf = {n: e.meanvar([i for i, m in enumerate(sizes) if m == n], welford=True)
                for n in (16, 24, 40)}
    refs = []
    for i, n in enumerate(sizes):
        out, steps = _oracle_final(n, u0[i], T)
        assert infos[i]["steps"] == steps and infos[i]["t"] == T
        assert np.array_equal(out.view(np.uint64), finals[i].view(np.uint64)), "sample %d" % i
        refs.append(out)
    for n in (16, 24, 40):
        group = [refs[i] for i, m in enumerate(sizes) if m == n]
        mean_ref, var_ref = O.meanvar(np.stack(group))
        mean, var = stats[n]
        assert np.array_equal(mean.view(np.uint64), mean_ref.view(np.uint64))
        assert np.array_equal(var.view(np.uint64), var_ref.view(np.uint64))
        wm, wv = welf[n]
        np.testing.assert_allclose(wm, mean_ref, rtol=1e-13, atol=1e-15)
        scale = np.maximum(np.abs(mean_ref) ** 2, 1e-30)
        assert np.all(np.abs(wv - var_ref) <= 1e-12 * scale)


@pytest.mark.parametrize("bc", [O.BC_PERIODIC, O.BC_NEUMANN])
def test_structure_functions_match_oracle(hip, bc):
    """structure_cube (h = 1..H-1) and structure_basic (h = 0..H-1, x and y)
    over 4 samples against the oracle's per-sample loops, divided by M."""
    _case("structure functions bc=%d" % bc)
    from alsvinn_amd import initial_data as ic
    rng = np.random.default_rng(2)
    n, H, M = 20, 7, 4
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, bc)
    samples = [hip.make_sample((n, n + 4, 1), (0, 0, 0), (1, 1, 0), 1.0) for _ in range(M)]
    vols = []
    with hip.Ensemble(scheme, samples) as e:
        shape = e.volume_shape(0)
        for i in range(M):
            u = rng.random(shape) + 1.0
            e.upload(i, u)
        e.apply_boundary()
        vols = [e.download(i) for i in range(M)]
        cube = e.structure_cube(H, 2.0)
        cube3 = e.structure_cube(H, 3.0)
        basic_x = e.structure_basic(H, 0, 2.0)
        basic_y = e.structure_basic(H, 1, 1.0)
    per = lambda f: sum(f(v) for v in vols) / M
    ref_cube = per(lambda v: O.structure_cube(v, (n, n + 4, 1), 2, H, 2.0, bc == O.BC_PERIODIC))
    ref_cube3 = per(lambda v: O.structure_cube(v, (n, n + 4, 1), 2, H, 3.0, bc == O.BC_PERIODIC))
    ref_bx = per(lambda v: O.structure_basic(v, (n, n + 4, 1), 2, H, 2.0, 0, bc == O.BC_PERIODIC))
    ref_by = per(lambda v: O.structure_basic(v, (n, n + 4, 1), 2, H, 1.0, 1, bc == O.BC_PERIODIC))
    np.testing.assert_allclose(cube, ref_cube, rtol=RTOL_REDUCTION, atol=0)
    np.testing.assert_allclose(cube3, ref_cube3, rtol=RTOL_REDUCTION, atol=0)
    np.testing.assert_allclose(basic_x, ref_bx, rtol=RTOL_REDUCTION, atol=0)
    np.testing.assert_allclose(basic_y, ref_by, rtol=RTOL_REDUCTION, atol=0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("bc", [O.BC_PERIODIC, O.BC_NEUMANN])
def test_structure_cube_2d_tiles_and_wide_halos(hip, bc, prec):
    """2D structure_cube at the edges of its LDS-tiled kernel: a halo wider
    than the grid (H = 12 on 9 x 7: several periodic wraps or Neumann clamps
    per side), partial tiles at the largest H of the statistics
    configurations (H = 33 on 33 x 17), H = 40 (periodic: the 32 x 32-tile
    kernel's 110 x 110 window; Neumann: the per-cell kernel) and H = 56 (the
    largest window the periodic kernel stages, 163104 B of LDS with the wave
    sums), for FP64 and FP32 states, against the oracle on the downloaded
    volumes; structure_basic in x and y on the same volumes (shifts up to
    H-1, past the grid's width on 9 x 7)."""
    _case("structure cube tiles bc=%d %s" % (bc, prec))
    rng = np.random.default_rng(11)
    precision = hip.FP64 if prec == "fp64" else hip.FP32
    for (nx, ny), H, M in (((9, 7), 12, 3), ((33, 17), 33, 2), ((18, 20), 40, 2),
                           ((40, 37), 56, 1)):
        scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, bc,
                                 precision=precision)
        samples = [hip.make_sample((nx, ny, 1), (0, 0, 0), (1, 1, 0), 1.0) for _ in range(M)]
        with hip.Ensemble(scheme, samples) as e:
            for i in range(M):
                e.upload(i, rng.random(e.volume_shape(i)) + 1.0)
            e.apply_boundary()
            vols = [e.download(i) for i in range(M)]
            cube = e.structure_cube(H, 2.0)
            basic = [e.structure_basic(H, d, 2.0) for d in (0, 1)]
        per = lambda f: sum(f(v) for v in vols) / M
        ref = per(lambda v: O.structure_cube(v, (nx, ny, 1), 2, H, 2.0, bc == O.BC_PERIODIC))
        # the parallel and the sequential sum differ by rounding that grows
        # with the number of terms: RTOL_REDUCTION is sized for about 1e5;
        # scaled by sqrt(terms / 1e5) beyond (the largest h's sum on 40 x 37
        # with H = 56 adds 6.5e5 terms: about 2.6e-12)
        terms = nx * ny * 8 * (H - 1)
        rtol = RTOL_REDUCTION * max(1.0, (terms / 1e5) ** 0.5)
        np.testing.assert_allclose(cube, ref, rtol=rtol, atol=0)
        assert np.all(cube[:, 0] == 0.0)
        for d in (0, 1):
            ref_b = per(lambda v: O.structure_basic(v, (nx, ny, 1), 2, H, 2.0, d,
                                                    bc == O.BC_PERIODIC))
            np.testing.assert_allclose(basic[d], ref_b, rtol=RTOL_REDUCTION, atol=0)


def _rounded_power_moment(vols, p):
    """Per cell, the sum in sample order of the correctly rounded u**p
    (Fraction arithmetic is exact; its conversion to float rounds once),
    divided by the number of samples."""
    flat = np.asarray(vols, dtype=np.float64).reshape(len(vols), -1)
    out = np.empty(flat.shape[1])
    for i in range(flat.shape[1]):
        acc = 0.0
        for s in range(flat.shape[0]):
            acc += float(Fraction(float(flat[s, i])) ** p)
        out[i] = acc / flat.shape[0]
    return out


@pytest.mark.parametrize("dim", [1, 2])
def test_onepoint_moment_and_bv_match_oracle(hip, dim):
    """OnePointMoment (p = 1, 2 bit for bit: u and u*u summed in sample order;
    p = 3 within 4 ulp per term of the oracle, whose libm pow is within 1 ulp
    of the correctly rounded power; p = 3, 4 bit for bit against the
    correctly rounded powers, from exact rational arithmetic, summed in sample
    order) and BoundedVariation (each sample's sum sequential in the
    reference's loop order, so bit for bit for p = 0, 1, 2; p = 3 within 1e-14)
    over 5 samples with random data, ghost layers filled by the boundary,
    against the oracle."""
    _case("onepoint/bv dim=%d" % dim)
    rng = np.random.default_rng(7 + dim)
    M = 5
    if dim == 1:
        scheme = hip.make_scheme(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2,
                                 O.BC_NEUMANN)
        samples = [hip.make_sample((301, 1, 1), (0, 0, 0), (1, 0, 0), 1.0) for _ in range(M)]
        ghosts = (2, 0, 0)
    else:
        scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3,
                                 O.BC_PERIODIC)
        samples = [hip.make_sample((37, 29, 1), (0, 0, 0), (1, 1, 0), 1.0) for _ in range(M)]
        ghosts = (2, 2, 0)
    with hip.Ensemble(scheme, samples) as e:
        for i in range(M):
            e.upload(i, rng.standard_normal(e.volume_shape(i)) + 2.0)
        e.apply_boundary()
        vols = np.stack([e.download(i) for i in range(M)])
        m = {p: e.onepoint_moment(p) for p in (1, 2, 3, 4)}
        bv = {p: e.bounded_variation(p) for p in (0, 1, 2, 3)}
        order = [3, 0, 4]
        m_sub = e.onepoint_moment(2, samples=order)
        bv_sub = e.bounded_variation(1, samples=order)
    for p in (1, 2):
        ref = O.onepoint_moment(vols, p)
        assert np.array_equal(m[p].view(np.uint64), ref.view(np.uint64)), "m%d" % p
    ref3 = O.onepoint_moment(vols, 3)
    scale = O.onepoint_moment(np.abs(vols), 3)
    assert np.all(np.abs(m[3] - ref3) <= 4 * np.spacing(scale))
    for p in (3, 4):
        exact = _rounded_power_moment(vols, p).reshape(np.shape(m[p]))
        assert np.array_equal(m[p].view(np.uint64), exact.view(np.uint64)), "m%d" % p
    assert np.array_equal(m_sub, O.onepoint_moment(vols[order], 2))
    for p in (0, 1, 2):
        ref = O.bounded_variation(vols, ghosts, p)
        assert np.array_equal(bv[p].view(np.uint64), ref.view(np.uint64)), "bv%d" % p
    # p = 0: every term is 1, so the total counts the cells of the range and
    # nothing past it
    wx = vols.shape[-1] - 2 * ghosts[0] - 1
    wy = vols.shape[-2] - 2 * ghosts[1] - 1 if dim == 2 else 1
    assert np.all(bv[0] == wx * wy)
    np.testing.assert_allclose(bv[3], O.bounded_variation(vols, ghosts, 3), rtol=1e-14)
    assert np.array_equal(bv_sub, O.bounded_variation(vols[order], ghosts, 1))


def test_bv_rejects_3d(hip):
    """The reference's getTotalVariation throws for 3D volumes."""
    _case("bv 3d rejected")
    scheme = hip.make_scheme(O.EQ_EULER3, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    with hip.Ensemble(scheme, [hip.make_sample((8, 8, 8), (0, 0, 0), (1, 1, 1), 1.0)]) as e:
        with pytest.raises(hip.HipError, match="3d"):
            e.bounded_variation(1)


def test_structure_cube_linear_closed_form(hip):
    """StructureTestBoundary.cpp:255-290 on the GPU: u_i = i in 1D."""
    _case("structure linear 1d")
    N, H = 16, 8
    scheme = hip.make_scheme(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2, O.BC_PERIODIC)
    with hip.Ensemble(scheme, [hip.make_sample((N, 1, 1), (0, 0, 0), (1, 0, 0), 1.0)]) as e:
        u = np.zeros(e.volume_shape(0))
        u[0, 0, 0, 2:-2] = np.arange(N)
        e.upload(0, u)
        out = e.structure_cube(H, 2.0)
    for h in range(1, H):
        tot = 0.0
        for i in range(N):
            for j in (-h, h):
                k = i + j
                tot += float((i - (k % N)) ** 2)
        assert out[0, h] == pytest.approx(tot / N, rel=1e-15)


def test_full_size_kh_conserves_totals(hip):
    """Full configuration size (256^2 x 64 samples, periodic): the FP64
    conserved totals of every sample are unchanged after 20 steps up to
    roundoff (relative 1e-12), and no sample leaves the constraint set."""
    _case("full size 256^2 x 64 totals")
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    M, n = 64, 256
    X = sampling.stl_mersenne_uniform(M, 40)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    with hip.Ensemble(scheme, [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), 2.0,
                                               max_steps=20) for _ in range(M)]) as e:
        for i in range(M):
            e.upload(i, ic.kelvin_helmholtz_2d(n, X[i]))
        e.apply_boundary()
        t0 = e.totals()
        e.run(10)
        t1 = e.totals()
        assert all(e.info(i)["steps"] == 20 and e.info(i)["status"] == 0 for i in (0, M - 1))
    scale = np.abs(t0).max(axis=0) + 1.0
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale * n * n / 1000)


@pytest.mark.parametrize("case", ["kh2d_256x1024_fp64", "kh2d_1024_single_fp64",
                                  "kh3d_256cubed_x64_fp32"])
def test_largest_configurations_conserve_totals(hip, case):
    """The benchmark configurations at full size: 1024 samples of 256^2
    (FP64, each from its own draws of the reference sampler), one 1024^2
    solve (FP64) and 64 samples of 256^3 (FP32 state, FP64 totals: four
    sampler draws, sample i the draw i % 4 with every conserved variable
    scaled by 1 + i/64, i.e. the same velocities and a proportionally
    denser, higher-pressure state), periodic: a few steps later every
    sample's conserved totals are unchanged to roundoff, every sample took
    the steps and none failed."""
    _case("largest " + case)
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    if case == "kh2d_256x1024_fp64":
        M, shape, prec, steps, tol = 1024, (256, 256, 1), hip.FP64, 6, 1e-12
        X = sampling.stl_mersenne_uniform(M, 40)
        make = lambda i: ic.kelvin_helmholtz_2d(256, X[i])
        eq = O.EQ_EULER2
    elif case == "kh2d_1024_single_fp64":
        M, shape, prec, steps, tol = 1, (1024, 1024, 1), hip.FP64, 12, 1e-12
        X = sampling.stl_mersenne_uniform(1, 40)
        make = lambda i: ic.kelvin_helmholtz_2d(1024, X[i])
        eq = O.EQ_EULER2
    else:
        M, shape, prec, steps, tol = 64, (256, 256, 256), hip.FP32, 3, 1e-6
        X = sampling.stl_mersenne_uniform(4, 400)
        base = [ic.kelvin_helmholtz_3d(256, X[k]) for k in range(4)]
        make = lambda i: base[i % 4] * (1.0 + i / 64.0)
        eq = O.EQ_EULER3
    top = (1, 1, 1 if shape[2] > 1 else 0)
    scheme = hip.make_scheme(eq, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                             precision=prec)
    samples = [hip.make_sample(shape, (0, 0, 0), top, 2.0, max_steps=steps) for _ in range(M)]
    distinct = set()
    l1 = np.zeros((M, 1))  # sum over cells of the largest |u_v| of the cell
    with hip.Ensemble(scheme, samples) as e:
        for i in range(M):
            u = make(i)
            distinct.add(hashlib.sha1(np.ascontiguousarray(u).tobytes()).hexdigest())
            l1[i] = np.abs(u).reshape(u.shape[0], -1).max(axis=0).sum()
            e.upload(i, u)
        e.apply_boundary()
        t0 = e.totals()
        e.run(steps)
        t1 = e.totals()
        _, nsteps, active, status = e.states()
    assert np.all(nsteps == steps) and np.all(status == 0) and not np.any(active)
    assert len(distinct) == M  # every sample starts from its own initial data
    # FP64: sums of per-cell roundings (the same bound as the 256^2 x 64 test
    # above); FP32 state: every cell's update rounds at a relative 6e-8 of
    # the fluxes entering it, which are of the size of the cell's largest
    # conserved variable (the y and z momenta start at 0 and are driven by
    # pressure differences of O(p)), so a total moves by at most a small
    # multiple (16 roundings: 1e-6) of the sum over cells of that magnitude
    cells = shape[0] * shape[1] * shape[2]
    if prec == hip.FP64:
        bound = tol * (np.abs(t0).max(axis=0) + 1.0) * cells / 1000
    else:
        bound = tol * l1
    assert np.all(np.abs(t1 - t0) <= bound)


def test_mixed_sizes_at_benchmark_size(hip):
    """The mixed-size benchmark ensemble, 171 x 128^2 + 171 x 256^2 + 170 x
    512^2 (FP64, reference sampler draws per sample): every sample conserves
    its totals over 4 steps, and the first sample of each size equals its own
    oracle run bit for bit."""
    _case("mixed 171x128^2 + 171x256^2 + 170x512^2")
    sizes = [128] * 171 + [256] * 171 + [512] * 170
    scheme, samples, u0 = _kh_ensemble(hip, sizes, T=2.0)
    for s in samples:
        s.max_steps = 4
    with hip.Ensemble(scheme, samples) as e:
        for i, u in enumerate(u0):
            e.upload(i, u)
        e.apply_boundary()
        t0 = e.totals()
        e.run(4)
        t1 = e.totals()
        _, nsteps, active, status = e.states()
        firsts = {n: sizes.index(n) for n in (128, 256, 512)}
        got = {n: e.download(i) for n, i in firsts.items()}
    assert np.all(nsteps == 4) and np.all(status == 0)
    cells = np.array([n * n for n in sizes], dtype=float)[:, None]
    scale = np.abs(t0).max(axis=0) + 1.0
    assert np.all(np.abs(t1 - t0) <= 1e-12 * scale * cells / 1000)
    for n, i in firsts.items():
        s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                         (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=2.0, max_steps=4)
        rc, ref, t, st, _ = O.run(s, u0[i])
        assert rc == 0 and st == 4
        assert np.array_equal(ref.view(np.uint64), got[n].view(np.uint64)), n


def test_fp32_state_at_benchmark_size(hip):
    """FP32 state on the headline batch (1024 samples of 256^2, reference
    sampler draws) against the FP64 batch after 10 SSP-RK3 steps: the
    relative L1 difference of rho is below 2e-5 for every sample (FP32
    roundoff of O(1e-7) per operation, amplified by 10 steps of the scheme),
    and of the ensemble mean of rho below 5e-6."""
    _case("fp32 vs fp64 256^2 x 1024")
    M, n = 1024, 256
    sizes = [n] * M
    res = {}
    for prec in (hip.FP64, hip.FP32):
        scheme, samples, u0 = _kh_ensemble(hip, sizes, T=2.0)
        scheme.precision = prec
        for s in samples:
            s.max_steps = 10
        with hip.Ensemble(scheme, samples) as e:
            for i, u in enumerate(u0):
                e.upload(i, u)
            e.apply_boundary()
            e.run(10)
            _, nsteps, _, status = e.states()
            assert np.all(nsteps == 10) and np.all(status == 0)
            res[prec] = ([e.download(i)[0, 0, 2:-2, 2:-2] for i in range(M)], e.meanvar()[0][0])
    errs = np.array([np.abs(a - b).sum() / np.abs(a).sum()
                     for a, b in zip(res[hip.FP64][0], res[hip.FP32][0])])
    print("rho rel L1: max %.3e mean %.3e" % (errs.max(), errs.mean()), flush=True)
    assert errs.max() < 2e-5
    m64, m32 = res[hip.FP64][1], res[hip.FP32][1]
    assert np.abs(m64 - m32).sum() / np.abs(m64).sum() < 5e-6


def test_full_time_uq_run_bit_exact(hip):
    """The whole UQ run to T = 2 (reference sampler, KH WENO2+HLL3 SSP-RK3,
    periodic) on 48^2 x 4 samples: every final state, the step counts, meanvar
    (bit for bit) and structure_cube(H=9, p=2) (1e-12) against the oracle's
    sequential sample loop."""
    _case("full T=2 uq 48^2 x 4")
    from alsvinn_amd import initial_data as ic
    from alsvinn_amd import sampling
    n, M, T = 48, 4, 2.0
    X = sampling.stl_mersenne_uniform(M, 40)
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    u0 = [ic.kelvin_helmholtz_2d(n, X[i]) for i in range(M)]
    with hip.Ensemble(scheme, [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), T)
                               for _ in range(M)]) as e:
        for i in range(M):
            e.upload(i, u0[i])
        e.apply_boundary()
        e.autotune()
        e.run(32)
        finals = [e.download(i) for i in range(M)]
        steps = [e.info(i)["steps"] for i in range(M)]
        mean, var = e.meanvar()
        sf = e.structure_cube(9, 2.0)
    refs = []
    for i in range(M):
        out, st = _oracle_final(n, u0[i], T)
        assert st == steps[i]
        assert np.array_equal(out.view(np.uint64), finals[i].view(np.uint64)), i
        refs.append(out)
    mr, vr = O.meanvar(np.stack(refs))
    assert np.array_equal(mean.view(np.uint64), mr.view(np.uint64))
    assert np.array_equal(var.view(np.uint64), vr.view(np.uint64))
    sref = sum(O.structure_cube(r, (n, n, 1), 2, 9, 2.0, True) for r in refs) / M
    np.testing.assert_allclose(sf, sref, rtol=RTOL_REDUCTION, atol=0)
