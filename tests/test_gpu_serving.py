"""Continuous batching at the serving configuration's sizes (1D Burgers of
4096-8192 cells, 2D Euler four-quadrant Riemann problems of 256^2-320^2,
short end times): every result is bit-identical alone and in the mix, and
equal to the oracle; requests queue and are admitted while others run; the
two pools' kernels run at the same time on the device."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def _requests():
    # a burst: all arrive at once and the arenas hold two of each class, so
    # the later ones are admitted while earlier ones are on the GPU
    from alsvinn_amd import serving
    return serving.make_requests(8, rate=0.0, seed=7, t_range=(0.004, 0.012),
                                 burgers_n=(4096, 8192), euler_n=(256, 320))


def _config():
    from alsvinn_amd import serving
    cfg = serving.ServerConfig()
    cfg.slots = {"burgers": 3, "euler2d": 3}
    cfg.max_n = {"burgers": (8192, 1, 1), "euler2d": (320, 320, 1)}
    cfg.arena_elements = {"burgers": 2 * 8196 + 10, "euler2d": 2 * 324 * 324 * 4 + 10}
    cfg.steps_per_tick = 10
    return cfg


def _serve(reqs):
    from alsvinn_amd import serving
    srv = serving.Server(_config())
    try:
        srv.serve(reqs)
    finally:
        srv.close()
    return reqs


def test_serving_results_same_alone_and_in_mix_and_oracle(hip):
    print("CASE serving mix vs alone vs oracle", flush=True)
    from alsvinn_amd import serving
    mix = _serve(_requests())
    assert all(r.finished is not None and r.status == 0 for r in mix)
    # the arenas hold fewer than the offered load: some requests had to queue
    assert any(r.admitted > r.arrival + 1e-4 for r in mix)
    for r in mix:
        alone = _requests()[r.rid]
        _serve([alone])
        assert alone.steps == r.steps
        assert np.array_equal(alone.result.view(np.uint64), r.result.view(np.uint64)), r.rid
        u0 = serving.initial_state(r)
        if r.kind == "burgers":
            s = O.make_setup(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2, O.BC_PERIODIC,
                             (r.n, 1, 1), (0, 0, 0), (1, 0, 0), cfl=0.475, end_time=r.end_time)
        else:
            s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_NEUMANN,
                             (r.n, r.n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=r.end_time)
        rc, ref, t, steps, _ = O.run(s, u0)
        assert rc == 0 and steps == r.steps, (r.rid, steps, r.steps)
        assert np.array_equal(ref.view(np.uint64), r.result.view(np.uint64)), r.rid
    sys.stdout.flush()


def test_pools_overlap_on_the_device(hip):
    """Both pools stepping: their kernels' intervals (HIP events on the two
    streams) overlap, i.e. the Burgers and Euler work runs at the same time."""
    print("CASE pools overlap", flush=True)
    from alsvinn_amd import serving
    cfg = _config()
    cfg.arena_elements = {"burgers": 4 * 8196, "euler2d": 2 * 324 * 324 * 4}
    srv = serving.Server(cfg)
    try:
        rs = serving.make_requests(6, rate=0.0, seed=3, t_range=(1.0, 1.0),
                                   burgers_n=(8192, 8192), euler_n=(320, 320),
                                   kinds=("burgers", "euler2d"))
        for r in rs:
            if srv.pools[r.kind].fits(serving.request_sample(r)):
                srv.try_admit(r, 0.0)
        pb, pe = srv.pools["burgers"], srv.pools["euler2d"]
        assert srv.live["burgers"] and srv.live["euler2d"]
        for _ in range(3):
            pe.run_steps(40, 10)  # enqueued back to back on the two streams
            pb.run_steps(400, 10)
            s_b, e_b = pb.interval_ms(pe)
            s_e, e_e = pe.interval_ms(pe)
            print("euler [%.3f, %.3f] ms, burgers [%.3f, %.3f] ms" % (s_e, e_e, s_b, e_b), flush=True)
            if s_b < e_e and s_e < e_b:
                break
        assert s_b < e_e and s_e < e_b, "the two pools' kernels did not overlap"
    finally:
        srv.close()


def test_assignment_keeps_the_captured_graph(hip):
    """Admitting and releasing solves while a pool steps leaves results
    bit-identical to a pool that never changed (no re-capture needed: the
    work list is read on the device)."""
    print("CASE admission during stepping", flush=True)
    from alsvinn_amd import serving
    from alsvinn_amd.ensemble import Pool
    sch = serving.schemes()["burgers"]
    rs = serving.make_requests(3, rate=0.0, seed=5, t_range=(0.01, 0.01), burgers_n=(4096, 6000),
                               kinds=("burgers",))
    ref = []
    for r in rs:  # each alone, to completion
        p = Pool(sch, 2, (8192, 1, 1), 8196)
        slot = p.assign(serving.request_sample(r))
        p.upload(slot, serving.initial_state(r))
        while True:
            snap = p.poll(p.tick(20, 10), wait=True)
            if not snap[2][slot]:
                break
        ref.append((int(snap[1][slot]), p.download(slot)))
        p.close()
    p = Pool(sch, 3, (8192, 1, 1), 3 * 8196)
    s0 = p.assign(serving.request_sample(rs[0]))
    p.upload(s0, serving.initial_state(rs[0]))
    p.tick(20, 10)  # graph captured with one solve
    s1 = p.assign(serving.request_sample(rs[1]))  # admitted while the first steps
    p.upload(s1, serving.initial_state(rs[1]))
    p.tick(20, 10)
    s2 = p.assign(serving.request_sample(rs[2]))
    p.upload(s2, serving.initial_state(rs[2]))
    done = {}
    while len(done) < 3:
        snap = p.poll(p.tick(20, 10), wait=True)
        for i, s in enumerate((s0, s1, s2)):
            if i not in done and not snap[2][s]:
                done[i] = (int(snap[1][s]), p.download(s))
    p.close()
    for i in range(3):
        assert done[i][0] == ref[i][0]
        assert np.array_equal(done[i][1].view(np.uint64), ref[i][1].view(np.uint64)), i
