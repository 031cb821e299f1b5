"""Continuous-batching request loop over the hip platform.

Solve requests (equation, grid, initial-data parameters, end time) arrive
over time; each is admitted into the pool of its scheme as soon as a slot and
arena memory are free (first-fit by footprint), all live solves of a pool
advance together in the pool's batched step (HIP graph replays on the pool's
own stream, so the pools' work overlaps on the device), and a finished solve
is read back and its slot released while the others keep stepping.

Per-request arithmetic never depends on the batch composition (no
cross-request reductions: every sample has its own CFL max, dt and clock), so
a result is bit-identical alone and in the mix.

Request classes (the serving configuration):
  burgers: 1D periodic Burgers on [0,1], fractional-Brownian initial data
           (random midpoint displacement, Hurst H), ENO2 + Godunov, SSP-RK2.
  euler2d: 2D Euler four-quadrant Riemann problem on [0,1]^2, quadrant
           rho, p in [0.1, 1.5], u, v in [-1, 1], WENO2 + HLL3, SSP-RK3,
           Neumann (outflow) boundaries.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import initial_data as ic
from .ensemble import Pool, make_sample, make_scheme


@dataclass
class Request:
    rid: int
    kind: str            # "burgers" | "euler2d"
    n: int               # cells per axis
    end_time: float
    arrival: float       # seconds after the stream starts
    hurst: float = 0.5   # burgers
    seed: int = 0        # burgers fBm seed
    states: np.ndarray = None  # euler2d quadrant states (4, 4)
    admitted: float = None
    finished: float = None
    steps: int = 0
    result: np.ndarray = None
    status: int = 0

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def cells(self):
        return self.n if self.kind == "burgers" else self.n * self.n


def schemes():
    return {
        "burgers": make_scheme(_lib.EQ_BURGERS, _lib.FLUX_GODUNOV, _lib.RECON_ENO2, _lib.INT_RK2,
                               _lib.BC_PERIODIC),
        "euler2d": make_scheme(_lib.EQ_EULER2, _lib.FLUX_HLL3, _lib.RECON_WENO2, _lib.INT_RK3,
                               _lib.BC_NEUMANN),
    }


def request_sample(r):
    if r.kind == "burgers":
        return make_sample((r.n, 1, 1), (0, 0, 0), (1, 0, 0), r.end_time)
    return make_sample((r.n, r.n, 1), (0, 0, 0), (1, 1, 0), r.end_time)


def initial_state(r):
    if r.kind == "burgers":
        return ic.burgers_fbm(r.n, r.hurst, r.seed)
    return ic.riemann_four_quadrant(r.n, r.states)


def make_requests(count, rate, seed=0, t_range=(0.1, 1.0), burgers_n=(4096, 65536),
                  euler_n=(256, 1024), kinds=("burgers", "euler2d")):
    """A Poisson stream (exponential inter-arrival times, mean 1/rate; rate 0
    means every request arrives at t = 0) mixing the request classes."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        if rate > 0:
            t += rng.exponential(1.0 / rate)
        kind = kinds[int(rng.integers(len(kinds)))]
        T = float(rng.uniform(*t_range))
        if kind == "burgers":
            r = Request(i, kind, int(rng.integers(burgers_n[0], burgers_n[1] + 1)), T, t,
                        hurst=float(rng.uniform(0.25, 0.75)), seed=int(rng.integers(1 << 31)))
        else:
            r = Request(i, kind, int(rng.integers(euler_n[0], euler_n[1] + 1)), T, t,
                        states=ic.riemann_parameters(rng.uniform(0, 1, 16)))
        out.append(r)
    return out


@dataclass
class ServerConfig:
    slots: dict = field(default_factory=lambda: {"burgers": 64, "euler2d": 32})
    # arena per pool in state elements per buffer (its share of the HBM budget)
    arena_elements: dict = field(default_factory=lambda: {"burgers": 64 * 65540,
                                                          "euler2d": 16 * 1028 * 1028 * 4})
    steps_per_tick: int = 20
    steps_per_graph: int = 10
    keep_results: bool = True


class Server:
    def __init__(self, config=None):
        self.cfg = config or ServerConfig()
        sch = schemes()
        self.pools = {k: Pool(sch[k], self.cfg.slots[k], 1 if k == "burgers" else 2,
                              self.cfg.arena_elements[k]) for k in sch}
        self.live = {k: {} for k in self.pools}  # kind -> slot -> Request

    def close(self):
        for p in self.pools.values():
            p.close()

    def footprint_bytes(self, r):
        p = self.pools[r.kind]
        esize = 8
        buffers = 3 if r.kind == "euler2d" else 2
        return p.footprint((r.n, r.n if r.kind == "euler2d" else 1, 1)) * esize * buffers

    def try_admit(self, r, now):
        p = self.pools[r.kind]
        slot = p.assign(request_sample(r))
        if slot is None:
            return False
        p.upload(slot, initial_state(r))
        self.live[r.kind][slot] = r
        r.admitted = now
        return True

    def serve(self, requests, clock_start=None):
        """Runs the stream to completion; returns the requests with their
        results, step counts and latencies (seconds after arrival)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []
        t0 = time.perf_counter() if clock_start is None else clock_start
        K, G = self.cfg.steps_per_tick, self.cfg.steps_per_graph
        cell_updates = 0
        while pending or queue or any(self.live[k] for k in self.live):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            admitted = []
            for r in queue:
                if self.try_admit(r, now):
                    admitted.append(r)
            for r in admitted:
                queue.remove(r)
            if admitted:
                for k in self.pools:
                    if any(rr.kind == k for rr in admitted):
                        self.pools[k].apply_boundary()
            busy = [k for k in self.pools if self.live[k]]
            if not busy:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.05)))
                continue
            for k in busy:  # enqueue on every pool's stream, then wait
                self.pools[k].run_steps(K, G)
            for k in busy:
                p = self.pools[k]
                t, steps, act, st = p.states()
                done_at = time.perf_counter() - t0
                for slot, r in list(self.live[k].items()):
                    if not act[slot]:
                        if self.cfg.keep_results:
                            r.result = p.download(slot)
                        r.steps = int(steps[slot])
                        r.status = int(st[slot])
                        r.finished = done_at
                        cell_updates += r.steps * r.cells
                        p.release(slot)
                        del self.live[k][slot]
        self.elapsed = time.perf_counter() - t0
        self.cell_updates = cell_updates
        return requests


def latency_stats(requests):
    lat = np.array([r.latency for r in requests if r.latency is not None])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "count": int(len(lat))}
