"""Whole-step launches of small 1D solves (alsvinn_hip_ensemble_set_small_1d):
one workgroup per sample runs whole timesteps in one launch, with workgroup
barriers where the multi-launch step has kernel boundaries.

The states (interior and ghost cells), times, step counts and dt sequences
must be the same bits as the multi-launch step's. The oracle parity tests
pin both paths: Sod 1024 and the other 1D schemes of test_gpu_parity.py run
through the whole-step launch by default, Burgers 4096 (69 strips, more
than the default four turns of 16 wavefronts) through the multi-launch step.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("CASE", name, flush=True)
    sys.stdout.flush()


def _initial(hip, eq, n, g, seed):
    from alsvinn_amd import initial_data as ic
    if eq == hip.EQ_BURGERS:
        return ic.burgers_fbm(n, 0.25 + 0.5 * (seed % 5) / 4, seed=seed, ghosts=g)
    return ic.sod(n, ghosts=g)


def _solve(hip, scheme, samples, u0s, rounds, drive):
    with hip.Ensemble(scheme, samples) as e:
        threads = e.set_small_1d(rounds)
        e.record_dt(4096)
        for s, u in enumerate(u0s):
            e.upload(s, u)
        e.apply_boundary()
        drive(e)
        out = [e.download(s) for s in range(len(samples))]
        infos = [e.info(s) for s in range(len(samples))]
        dts = [e.dt_history(s) for s in range(len(samples))]
    return threads, out, infos, dts


def _same_bits(a, b):
    assert a.shape == b.shape
    diff = np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))
    assert diff.size == 0, "%d values differ; first at %d: %r vs %r" % (
        diff.size, diff[0], a.flat[diff[0]], b.flat[diff[0]])


# (name, equation, flux, reconstruction, integrator, boundary, sizes, end time)
CASES = [
    ("sod fe hll none", "EQ_EULER1", "FLUX_HLL", "RECON_NONE", "INT_FE", "BC_NEUMANN",
     [1024], 0.2),
    ("euler1 rk3 weno2 hll3 ragged", "EQ_EULER1", "FLUX_HLL3", "RECON_WENO2", "INT_RK3",
     "BC_NEUMANN", [200, 61, 3000, 5, 60], 0.1),
    ("euler1 rk2 wenof2 hll periodic", "EQ_EULER1", "FLUX_HLL", "RECON_WENOF2", "INT_RK2",
     "BC_PERIODIC", [256, 119, 2], 0.1),
    ("burgers rk2 eno2 godunov ragged", "EQ_BURGERS", "FLUX_GODUNOV", "RECON_ENO2", "INT_RK2",
     "BC_PERIODIC", [3000, 60, 61, 7, 1024], 0.05),
    ("burgers fe none rusanov", "EQ_BURGERS", "FLUX_RUSANOV", "RECON_NONE", "INT_FE",
     "BC_PERIODIC", [999, 62, 63], 0.05),
    ("burgers rk3 mm roe", "EQ_BURGERS", "FLUX_ROE", "RECON_MM", "INT_RK3", "BC_NEUMANN",
     [500, 2], 0.05),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_whole_step_launch_matches_multi_launch(hip, case):
    name, eq, flux, recon, integ, bc, sizes, T = case
    _print_case("small 1d: " + name)
    eq, flux, recon = getattr(hip, eq), getattr(hip, flux), getattr(hip, recon)
    integ, bc = getattr(hip, integ), getattr(hip, bc)
    scheme = hip.make_scheme(eq, flux, recon, integ, bc)
    g = 1 if recon == hip.RECON_NONE else 2
    lo, hi = ((-1, 0, 0), (1, 0, 0)) if eq == hip.EQ_EULER1 else ((0, 0, 0), (1, 0, 0))
    # mixed end times and save counts: samples finish and clamp at different steps
    samples = [hip.make_sample((n, 1, 1), lo, hi, T * (0.5 + 0.25 * (i % 3)), saves=(1 + i % 3,))
               for i, n in enumerate(sizes)]
    u0s = [_initial(hip, eq, n, g, seed=3 + i) for i, n in enumerate(sizes)]
    run = lambda e: e.run(steps_per_sync=8)
    t1, out1, info1, dts1 = _solve(hip, scheme, samples, u0s, 64, run)
    t0, out0, info0, dts0 = _solve(hip, scheme, samples, u0s, 0, run)
    assert t0 == 0
    assert t1 == 64 * min(16, max(-(-n // (64 - 2 * g)) for n in sizes))
    for s in range(len(sizes)):
        assert info1[s]["steps"] == info0[s]["steps"] > 0
        assert info1[s]["t"] == info0[s]["t"]
        assert info1[s]["status"] == info0[s]["status"] == 0
        _same_bits(dts1[s], dts0[s])
        _same_bits(out1[s], out0[s])


@pytest.mark.parametrize("eq", ["EQ_EULER1", "EQ_BURGERS"])
def test_whole_step_launch_fp32_state(hip, eq):
    """FP32-state ensembles take the same whole-step launch; its results
    are the multi-launch step's, bit for bit."""
    _print_case("small 1d: fp32 " + eq)
    eq = getattr(hip, eq)
    if eq == hip.EQ_BURGERS:
        scheme = hip.make_scheme(eq, hip.FLUX_GODUNOV, hip.RECON_ENO2, hip.INT_RK2, hip.BC_PERIODIC,
                                 precision=hip.FP32)
        lo, hi, g = (0, 0, 0), (1, 0, 0), 2
    else:
        scheme = hip.make_scheme(eq, hip.FLUX_HLL3, hip.RECON_WENO2, hip.INT_RK3, hip.BC_NEUMANN,
                                 precision=hip.FP32)
        lo, hi, g = (-1, 0, 0), (1, 0, 0), 2
    sizes = [1000, 130, 7]
    samples = [hip.make_sample((n, 1, 1), lo, hi, 0.05 * (1 + i), saves=(2,))
               for i, n in enumerate(sizes)]
    u0s = [_initial(hip, eq, n, g, seed=20 + i) for i, n in enumerate(sizes)]
    run = lambda e: e.run(steps_per_sync=8)
    t1, out1, info1, dts1 = _solve(hip, scheme, samples, u0s, 64, run)
    t0, out0, info0, dts0 = _solve(hip, scheme, samples, u0s, 0, run)
    assert t1 > 0 and t0 == 0
    for s in range(len(sizes)):
        assert info1[s]["steps"] == info0[s]["steps"] > 0
        assert info1[s]["t"] == info0[s]["t"]
        _same_bits(dts1[s], dts0[s])
        _same_bits(out1[s], out0[s])  # FP32 states, widened exactly by the download


def test_whole_step_launch_pauses_at_holds_and_step_counts(hip):
    """Samples held at a save stop inside the launch; run_steps advances each
    runnable sample exactly nsteps steps; releasing the hold continues the
    same sequence."""
    _print_case("small 1d: holds and partial step counts")
    scheme = hip.make_scheme(hip.EQ_EULER1, hip.FLUX_HLL3, hip.RECON_WENO2, hip.INT_RK3,
                             hip.BC_NEUMANN)
    sizes = [512, 100, 1500]
    samples = [hip.make_sample((n, 1, 1), (-1, 0, 0), (1, 0, 0), 0.2, saves=(4,)) for n in sizes]
    u0s = [_initial(hip, hip.EQ_EULER1, n, 2, 0) for n in sizes]

    def drive(e):
        e.set_hold(0, 3)  # pause after the second save (counters start at 1)
        e.run(steps_per_sync=16)
        for s in range(len(sizes)):
            assert e.saves(s)[0] == 3
        e.set_hold(0, None)
        e.run_steps(30, 10)

    t1, out1, info1, dts1 = _solve(hip, scheme, samples, u0s, 64, drive)
    t0, out0, info0, dts0 = _solve(hip, scheme, samples, u0s, 0, drive)
    assert t1 > 0 and t0 == 0
    for s in range(len(sizes)):
        assert info1[s]["steps"] == info0[s]["steps"]
        assert info1[s]["t"] == info0[s]["t"]
        _same_bits(dts1[s], dts0[s])
        _same_bits(out1[s], out0[s])
    # the coarse sample reaches its end time; the finest is still running
    assert info1[1]["active"] == 0 and abs(info1[1]["t"] - 0.2) < 1e-12
    assert info1[2]["t"] < 0.2 and info1[2]["active"] == 1


def test_whole_step_launch_eligibility(hip):
    """The whole-step launch is used for 1D ensembles of at most 256 samples
    whose strips fit the allowed turns, and never for RK4, 2D or large
    ensembles."""
    _print_case("small 1d: eligibility")

    def threads(eq, recon, integ, n, nsamples=1, rounds=None):
        scheme = hip.make_scheme(eq, hip.FLUX_GODUNOV if eq == hip.EQ_BURGERS else hip.FLUX_HLL,
                                 recon, integ, hip.BC_PERIODIC)
        samples = [hip.make_sample(n, (0, 0, 0), (1, 1, 0), 0.1)] * nsamples
        with hip.Ensemble(scheme, samples) as e:
            return e.small_1d() if rounds is None else e.set_small_1d(rounds)

    E1, B = hip.EQ_EULER1, hip.EQ_BURGERS
    assert threads(E1, hip.RECON_NONE, hip.INT_FE, (1024, 1, 1)) == 1024  # 17 strips of 62
    assert threads(B, hip.RECON_ENO2, hip.INT_RK2, (120, 1, 1)) == 128    # 2 strips of 60
    assert threads(B, hip.RECON_ENO2, hip.INT_RK2, (3840, 1, 1)) == 1024  # 64 strips: 4 turns
    assert threads(B, hip.RECON_ENO2, hip.INT_RK2, (3841, 1, 1)) == 0     # 65 strips
    assert threads(B, hip.RECON_ENO2, hip.INT_RK2, (3841, 1, 1), rounds=5) == 1024
    assert threads(E1, hip.RECON_NONE, hip.INT_RK4, (1024, 1, 1)) == 0
    assert threads(E1, hip.RECON_ENO3, hip.INT_RK3, (1024, 1, 1)) == 0  # wide-stencil kernel
    assert threads(hip.EQ_EULER2, hip.RECON_NONE, hip.INT_FE, (64, 64, 1)) == 0
    assert threads(B, hip.RECON_ENO2, hip.INT_RK2, (256, 1, 1), nsamples=256) == 320
    assert threads(B, hip.RECON_ENO2, hip.INT_RK2, (256, 1, 1), nsamples=257) == 0
