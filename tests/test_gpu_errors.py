"""Failure handling on the GPU, forced from the host with valid kernels:
a state that leaves the constraint set stops its sample with status 2 and
the run reports the reference's constraint error (Simulator.cpp:180-189);
other samples of the ensemble are unaffected."""
import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_constraint_violation_stops_only_that_sample(hip):
    print("CASE constraint violation", flush=True)
    from alsvinn_amd import initial_data as ic
    n = 64
    scheme = hip.make_scheme(O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.INT_FE, O.BC_NEUMANN)
    samples = [hip.make_sample((n, 1, 1), (-1, 0, 0), (1, 0, 0), 0.2) for _ in range(2)]
    good = ic.sod(n)
    bad = good.copy()
    bad[2, 0, 0, 10:20] = -100.0  # negative energy -> negative pressure after the first step
    with hip.Ensemble(scheme, samples) as e:
        e.upload(0, good)
        e.upload(1, bad)
        e.apply_boundary()
        with pytest.raises(hip.HipError, match="does not obey constraints") as err:
            e.run(4)
        assert err.value.code == 3
        i0, i1 = e.info(0), e.info(1)
        got = e.download(0)
    s = O.make_setup(O.EQ_EULER1, O.FLUX_HLL, O.RECON_NONE, O.INT_FE, O.BC_NEUMANN, (n, 1, 1),
                     (-1, 0, 0), (1, 0, 0), cfl=0.9, end_time=0.2)
    rc, ref, t, steps, _ = O.run(s, good)
    assert rc == 0 and i0["status"] == 0 and i0["t"] == t and i0["steps"] == steps
    assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))
    rc_bad, _, t_bad, steps_bad, _ = O.run(s, bad)
    assert rc_bad == 2
    assert i1["status"] == 2 and i1["active"] == 0 and i1["steps"] == steps_bad


def test_invalid_arguments_fail_without_gpu_fault(hip):
    scheme = hip.make_scheme(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC)
    with pytest.raises(hip.HipError, match="ghost"):
        hip.Ensemble(scheme, [hip.make_sample((1, 8, 1), (0, 0, 0), (1, 1, 0), 1.0)])
    with pytest.raises(hip.HipError, match="same dimension"):
        hip.Ensemble(scheme, [hip.make_sample((8, 8, 1), (0, 0, 0), (1, 1, 0), 1.0),
                              hip.make_sample((8, 1, 1), (0, 0, 0), (1, 1, 0), 1.0)])
    burgers_wenof2 = hip.make_scheme(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_WENOF2, O.INT_RK2,
                                     O.BC_PERIODIC)
    with pytest.raises(hip.HipError, match="wenof2 needs the Euler equations"):
        hip.Ensemble(burgers_wenof2, [hip.make_sample((8, 1, 1), (0, 0, 0), (1, 0, 0), 1.0)])
    with hip.Ensemble(scheme, [hip.make_sample((8, 8, 1), (0, 0, 0), (1, 1, 0), 1.0)]) as e:
        with pytest.raises(hip.HipError, match="bad sample"):
            e.info(3)
        with pytest.raises(hip.HipError):
            e.run_steps(15, 10)
