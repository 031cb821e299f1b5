"""Continuous batching: results are the same alone and in the mix, and each
equals the oracle run of the same request."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def _requests():
    # a burst: all eight arrive at once and the arenas hold fewer, so the
    # later ones are admitted while earlier ones are on the GPU
    from alsvinn_amd import serving
    rs = serving.make_requests(8, rate=0.0, seed=7, t_range=(0.005, 0.02),
                               burgers_n=(2000, 6000), euler_n=(40, 90))
    return rs


def _small_config():
    from alsvinn_amd import serving
    cfg = serving.ServerConfig()
    cfg.slots = {"burgers": 3, "euler2d": 3}
    cfg.arena_elements = {"burgers": 3 * 6004, "euler2d": 2 * 94 * 94 * 4 + 10}
    cfg.steps_per_tick = 10
    return cfg


def test_serving_results_same_alone_and_in_mix(hip):
    print("CASE serving mix vs alone", flush=True)
    from alsvinn_amd import serving
    mix = _requests()
    srv = serving.Server(_small_config())
    try:
        srv.serve(mix)
    finally:
        srv.close()
    assert all(r.finished is not None and r.status == 0 for r in mix)
    # arena smaller than the offered load: some requests had to queue
    assert any(r.admitted > r.arrival + 1e-4 for r in mix)
    for r in mix:
        alone = _requests()[r.rid]
        alone.arrival = 0.0
        s1 = serving.Server(_small_config())
        try:
            s1.serve([alone])
        finally:
            s1.close()
        assert alone.steps == r.steps
        assert np.array_equal(alone.result.view(np.uint64), r.result.view(np.uint64)), r.rid
    sys.stdout.flush()


def test_serving_matches_oracle(hip):
    print("CASE serving vs oracle", flush=True)
    from alsvinn_amd import serving
    rs = _requests()[:4]
    srv = serving.Server(_small_config())
    try:
        srv.serve(rs)
    finally:
        srv.close()
    for r in rs:
        u0 = serving.initial_state(r)
        if r.kind == "burgers":
            s = O.make_setup(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2, O.BC_PERIODIC,
                             (r.n, 1, 1), (0, 0, 0), (1, 0, 0), cfl=0.475, end_time=r.end_time)
        else:
            s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_NEUMANN,
                             (r.n, r.n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=r.end_time)
        rc, ref, t, steps, _ = O.run(s, u0)
        assert rc == 0 and steps == r.steps
        assert np.array_equal(ref.view(np.uint64), r.result.view(np.uint64)), r.rid
