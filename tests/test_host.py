"""Host-side logic of the hip platform and its C ABI, on the CPU (no GPU call
is made): the library loads and exports every function the header declares,
argument validation fails with the reference's error behaviour, the sampler
and the initial data are bit-identical to the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "alsvinn_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(alsvinn_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    import alsvinn_amd
    L = alsvinn_amd._lib.load()
    names = _header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(alsvinn_amd._lib.EXPORTS) == names


def test_ghost_cells_and_validation():
    """NumericalFluxCPU.cpp:226-231: ghosts = max(stencil-1, reconstruction);
    unknown names are errors (NumericalFluxFactory.cpp:194-197 THROWs)."""
    import alsvinn_amd as hip
    L = hip._lib.load()
    s = hip.make_scheme(hip.EQ_EULER2, hip.FLUX_HLL3, hip.RECON_WENO2, hip.INT_RK3, hip.BC_PERIODIC)
    assert L.alsvinn_hip_ghost_cells(ctypes.byref(s)) == 2
    s.reconstruction = hip.RECON_NONE
    assert L.alsvinn_hip_ghost_cells(ctypes.byref(s)) == 1
    s.reconstruction = hip.RECON_ENO3
    assert L.alsvinn_hip_ghost_cells(ctypes.byref(s)) == 3
    s.reconstruction = hip.RECON_ENO4
    assert L.alsvinn_hip_ghost_cells(ctypes.byref(s)) == 4
    s.reconstruction = hip.RECON_WENO3
    assert L.alsvinn_hip_ghost_cells(ctypes.byref(s)) == 3
    s.reconstruction = 9
    assert L.alsvinn_hip_ghost_cells(ctypes.byref(s)) == -1
    assert b"unknown reconstruction" in L.alsvinn_hip_last_error()
    bad = hip.make_scheme(hip.EQ_BURGERS, hip.FLUX_HLL, hip.RECON_NONE, hip.INT_FE, hip.BC_NEUMANN)
    with pytest.raises(hip.HipError, match="godunov"):
        hip.Ensemble(bad, [hip.make_sample((8, 1, 1), (0, 0, 0), (1, 0, 0), 1.0)])


def test_cfl_auto_rule():
    """SimulatorSetup.cpp:306-320: cfl auto = 0.9 without reconstruction, else 0.475."""
    import alsvinn_amd as hip
    assert hip.make_scheme(1, 0, hip.RECON_NONE, 0, 0).cfl == 0.9
    assert hip.make_scheme(2, 1, hip.RECON_WENO2, 2, 1).cfl == 0.475


def test_product_sampler_matches_libstdcxx():
    """The product's std::mt19937_64 + uniform_real_distribution restatement
    equals libstdc++ (through the oracle) bit for bit, including the
    skip-ahead of STLMersenne and the Uniform scaling."""
    from alsvinn_amd import sampling
    for (n, d, lo, hi, start) in [(4, 40, 0.0, 1.0, 0), (3, 7, -2.0, 3.0, 5), (2, 400, 0.0, 1.0, 1)]:
        a = sampling.stl_mersenne_uniform(n, d, lo, hi, sample_start=start)
        b = O.stl_mersenne_uniform(n, d, lo, hi, sample_start=start)
        assert np.array_equal(a, b)
    e = sampling.MT19937_64()
    e.discard(9999)
    assert int(e.raw(1)[0]) == 9981545732273789042


def test_seeded_sampler_is_per_sample_engine():
    from alsvinn_amd import sampling
    a = sampling.seeded_uniform(100, 3, 5)
    assert np.array_equal(a[1], sampling.seeded_uniform(101, 1, 5)[0])
    assert not np.array_equal(a[0], a[1])


def test_kh_initial_data_band_and_energy():
    """KH: rho in {1, 2}, u = +-0.5, v = 0, E = 2.5/(gamma-1) + 0.5*rho*0.25; the
    band edges are 0.25 + s0(x), 0.75 + s1(x) with |s| <= 0.01."""
    from alsvinn_amd import initial_data as ic, sampling
    X = sampling.stl_mersenne_uniform(1, 40)[0]
    u = ic.kelvin_helmholtz_2d(64, X)
    rho = u[0, 0, 2:-2, 2:-2]
    assert set(np.unique(rho)) == {1.0, 2.0}
    y = ic.midpoints(64, 0, 1)
    assert np.all(rho[y < 0.24] == 1.0) and np.all(rho[(y > 0.26) & (y < 0.74)] == 2.0)
    np.testing.assert_array_equal(u[1, 0, 2:-2, 2:-2], np.where(rho == 2.0, -1.0, 0.5))
    assert np.all(u[2, 0, 2:-2, 2:-2] == 0.0)
    np.testing.assert_array_equal(u[3, 0, 2:-2, 2:-2], 2.5 / (1.4 - 1) + (0.5 * rho) * 0.25)


def test_riemann_and_fbm_initial_data():
    from alsvinn_amd import initial_data as ic
    prm = ic.riemann_parameters(np.linspace(0, 1, 16))
    assert np.all((prm[:, 0] >= 0.1) & (prm[:, 0] <= 1.5)) and np.all(np.abs(prm[:, 1:3]) <= 1)
    u = ic.riemann_four_quadrant(32, prm)
    assert u[0, 0, 2, 2] == prm[0, 0] and u[0, 0, -3, -3] == prm[3, 0]
    b = ic.fbm_rmd(4096, 0.5, seed=3)
    assert b.shape == (4096,) and np.isfinite(b).all()
    # Hurst scaling: increments over lag k have variance ~ k^(2H); H = 0.5 ->
    # ratio 4 between lags 64 and 16 (sampling noise allowed)
    for H, lo, hi in ((0.5, 2.5, 6.5), (0.75, 4.5, 13.0)):
        b = np.mean([np.var(ic.fbm_rmd(4096, H, seed=s)[64:] - ic.fbm_rmd(4096, H, seed=s)[:-64])
                     / np.var(ic.fbm_rmd(4096, H, seed=s)[16:] - ic.fbm_rmd(4096, H, seed=s)[:-16])
                     for s in range(8)])
        assert lo < b < hi, (H, b)


def test_first_fit_arena_placement():
    """Pool placement: lowest gap that fits, None when the arena is full."""
    from alsvinn_amd.ensemble import first_fit
    assert first_fit([], 100, 30) == 0
    assert first_fit([(0, 30)], 100, 30) == 30
    assert first_fit([(0, 30), (60, 40)], 100, 30) == 30
    assert first_fit([(0, 30), (60, 40)], 100, 31) is None
    assert first_fit([(10, 10)], 100, 10) == 0
    assert first_fit([(0, 50), (50, 50)], 100, 1) is None


def test_poisson_request_stream():
    from alsvinn_amd import serving
    rs = serving.make_requests(200, rate=10.0, seed=1, t_range=(0.1, 1.0))
    arr = np.array([r.arrival for r in rs])
    assert np.all(np.diff(arr) > 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.1) < 0.03
    for r in rs:
        assert 0.1 <= r.end_time <= 1.0
        if r.kind == "burgers":
            assert 4096 <= r.n <= 65536 and 0.25 <= r.hurst <= 0.75
        else:
            assert 256 <= r.n <= 1024 and r.states.shape == (4, 4)


def test_empty_ensemble_and_pool_rejected_before_any_gpu_call():
    """An ensemble with no samples, or a pool with no slots, is refused by
    argument checks that run before any HIP call (this test runs without a
    GPU)."""
    import alsvinn_amd as hip
    s = hip.make_scheme(hip.EQ_EULER2, hip.FLUX_HLL3, hip.RECON_WENO2, hip.INT_RK3,
                        hip.BC_PERIODIC)
    with pytest.raises(hip.HipError, match="bad ensemble arguments"):
        hip.Ensemble(s, [])
    L = hip._lib.load()
    out = ctypes.c_void_p()
    mx = (ctypes.c_int * 3)(64, 64, 1)
    assert L.alsvinn_hip_pool_create(ctypes.byref(s), 0, mx, 1000, None, ctypes.byref(out)) != 0
    assert b"bad pool arguments" in L.alsvinn_hip_last_error()


def test_per_sample_bv_rejects_memory_deeper_than_one_cell_before_any_gpu_call():
    """HostMemory::getTotalVariation (HostMemory.cpp:333-335) throws for any
    memory with more than one cell in z, ghost layers included: a 2D shape
    with z ghost layers is refused like a 3D one, and a plain 2D shape
    passes the argument checks (without a GPU it then stops at the first HIP
    call, which reports the missing device)."""
    import alsvinn_amd as hip
    from alsvinn_amd import _lib, operators

    for n, g in [((16, 16, 1), (2, 2, 1)), ((8, 8, 8), (1, 1, 1))]:
        shape = operators.make_volume(4, n, g, [0, 0, 0, 0])
        with pytest.raises(hip.HipError, match="not supported for 3d"):
            operators.Statistics(_lib.STAT_BV, shape, p=1.0)
    shape = operators.make_volume(4, (16, 16, 1), (2, 2, 0), [0, 0, 0, 0])
    try:
        operators.Statistics(_lib.STAT_BV, shape, p=1.0).close()
    except hip.HipError as err:
        assert "not supported for 3d" not in str(err)


class _FakePool:
    """A pool without a GPU for the serving loop: every solve finishes after
    int(end_time * 1000) steps; snapshots are taken at tick time."""

    def __init__(self, nslots, arena, max_n, ghosts=2):
        self.nslots, self.arena, self.max_n, self.ghosts = nslots, arena, max_n, ghosts
        self.allocs, self.need, self.done, self.snaps = {}, {}, {}, []

    def footprint(self, n):
        return (n[0] + 2 * self.ghosts) * ((n[1] + 2 * self.ghosts) if n[1] > 1 else 1)

    def fits(self, s):
        n = (s.nx, s.ny, s.nz)
        return (all(a <= b for a, b in zip(n, self.max_n)) and n[0] >= self.ghosts
                and self.footprint(n) <= self.arena)

    def assign(self, s):
        from alsvinn_amd.ensemble import first_fit
        free = [k for k in range(self.nslots) if k not in self.allocs]
        size = self.footprint((s.nx, s.ny, s.nz))
        off = first_fit(self.allocs.values(), self.arena, size)
        if not free or off is None:
            return None
        self.allocs[free[0]] = (off, size)
        self.need[free[0]] = int(s.end_time * 1000)
        self.done[free[0]] = 0
        return free[0]

    def upload(self, slot, u):
        pass

    def release(self, slot):
        del self.allocs[slot], self.need[slot], self.done[slot]

    def download(self, slot):
        return np.zeros(1)

    def tick(self, k, g):
        n = self.nslots
        steps = np.zeros(n, dtype=np.int64)
        act = np.zeros(n, dtype=np.int32)
        for s in self.allocs:
            self.done[s] = min(self.need[s], self.done[s] + k)
            steps[s], act[s] = self.done[s], int(self.done[s] < self.need[s])
        self.snaps.append((np.zeros(n), steps, act, np.zeros(n, dtype=np.int32)))
        return len(self.snaps) - 1

    def poll(self, ticket, wait=False):
        return self.snaps[ticket]

    def close(self):
        pass


def test_serving_loop_terminates_and_rejects_unfittable_requests():
    """serve() with pools that need no GPU: requests larger than a pool's
    arena or grid, or smaller than the ghost width, are rejected at arrival
    (status STATUS_REJECTED, no latency) instead of waiting forever; the
    others all finish, with queueing when the arena is full."""
    from alsvinn_amd import serving
    cfg = serving.ServerConfig()
    cfg.steps_per_tick = 7
    pools = {"burgers": _FakePool(2, 2 * 104, (100, 1, 1)),
             "euler2d": _FakePool(2, 2 * 24 * 24, (20, 20, 1))}
    reqs = serving.make_requests(6, rate=0.0, seed=1, t_range=(0.02, 0.05),
                                 burgers_n=(40, 100), euler_n=(8, 20))
    reqs[0].kind, reqs[0].n = "burgers", 500      # larger than the pool's grid and arena
    reqs[1].kind, reqs[1].n = "euler2d", 1        # smaller than the ghost width
    for r in reqs[:2]:
        r.states, r.hurst, r.seed = serving.ic.riemann_parameters(np.full(16, 0.5)), 0.5, 1
    srv = serving.Server(cfg, pools=pools)
    srv.serve(reqs)
    assert [r.status for r in reqs[:2]] == [serving.STATUS_REJECTED] * 2
    assert all(r.finished is not None and r.status == 0 for r in reqs[2:])
    assert all(r.steps == int(r.end_time * 1000) for r in reqs[2:])
    assert serving.latency_stats(reqs)["count"] == 4
