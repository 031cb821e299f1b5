"""Continuous-batching request loop over the hip platform.

Solve requests (equation, grid, initial-data parameters, end time) arrive
over time; each is admitted into the pool of its scheme as soon as a slot and
arena memory are free (first-fit by footprint; the arenas are the pools'
shares of the device memory budget), all live solves of a pool advance
together in the pool's batched step (HIP graph replays on the pool's own
stream, so the pools' work overlaps on the device), and a finished solve is
read back and its slot released while the others keep stepping. Admission,
retirement and progress checks never wait for a pool's stream.

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
    first_ticket: int = 0
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


STATUS_REJECTED = -1  # the request can never fit its pool (grid or footprint)


@dataclass
class ServerConfig:
    slots: dict = field(default_factory=lambda: {"burgers": 256, "euler2d": 128})
    # the largest grid of each request class (the pools' work lists are sized for it)
    max_n: dict = field(default_factory=lambda: {"burgers": (65536, 1, 1),
                                                 "euler2d": (1024, 1024, 1)})
    # arena per pool in state elements per buffer; None: each pool's share of
    # the device memory budget
    arena_elements: dict = None
    # bytes of device memory the pools may take; None: free memory at start
    # minus budget_margin_bytes
    hbm_budget_bytes: int = None
    budget_margin_bytes: int = 8 << 30
    budget_share: dict = field(default_factory=lambda: {"burgers": 0.05, "euler2d": 0.95})
    steps_per_tick: int = 20
    steps_per_graph: int = 10
    ticks_in_flight: int = 2
    keep_results: bool = True


# state buffers per solve: U and A for SSP-RK2, U, A, B for SSP-RK3
BUFFERS = {"burgers": 2, "euler2d": 3}


def arena_sizes(cfg, free_bytes):
    """State elements per buffer of each pool: its share of the budget (the
    device's free memory minus the margin, unless given) over its buffers."""
    budget = cfg.hbm_budget_bytes if cfg.hbm_budget_bytes is not None else \
        max(0, free_bytes - cfg.budget_margin_bytes)
    return {k: int(budget * cfg.budget_share[k] // (8 * BUFFERS[k])) for k in cfg.budget_share}


class Server:
    def __init__(self, config=None, pools=None):
        """pools: ready-made pools by request class (tests); by default one
        Pool per class over its arena."""
        self.cfg = config or ServerConfig()
        if pools is None:
            from .ensemble import device_memory
            sch = schemes()
            arenas = self.cfg.arena_elements
            if arenas is None:
                arenas = arena_sizes(self.cfg, device_memory()[0])
            self.arena_elements = dict(arenas)
            pools = {k: Pool(sch[k], self.cfg.slots[k], self.cfg.max_n[k], arenas[k]) for k in sch}
        self.pools = pools
        self.live = {k: {} for k in self.pools}  # kind -> slot -> Request
        self.inflight = {k: [] for k in self.pools}  # kind -> tickets in order
        self.issued = {k: 0 for k in self.pools}  # snapshots enqueued per pool

    def close(self):
        for p in self.pools.values():
            p.close()

    def footprint_bytes(self, r):
        p = self.pools[r.kind]
        return p.footprint((r.n, r.n if r.kind == "euler2d" else 1, 1)) * 8 * BUFFERS[r.kind]

    def try_admit(self, r, now):
        p = self.pools[r.kind]
        slot = p.assign(request_sample(r))
        if slot is None:
            return False
        p.upload(slot, initial_state(r))
        r.first_ticket = self.issued[r.kind]  # snapshots taken before it do not describe it
        self.live[r.kind][slot] = r
        r.admitted = now
        return True

    def _retire(self, k, snap, ticket, now):
        """Reads back and releases the solves a snapshot shows finished."""
        p = self.pools[k]
        t, steps, act, st = snap
        done = 0
        for slot, r in list(self.live[k].items()):
            if r.first_ticket > ticket or act[slot]:
                continue
            if self.cfg.keep_results:
                r.result = p.download(slot)
            r.steps = int(steps[slot])
            r.status = int(st[slot])
            r.finished = now
            self.cell_updates += r.steps * r.cells
            p.release(slot)
            del self.live[k][slot]
            done += 1
        return done

    def serve(self, requests, clock_start=None):
        """Runs the stream to completion; returns the requests with their
        results, step counts and latencies (seconds after arrival). Requests
        that can never fit their pool get status STATUS_REJECTED at arrival.

        Every busy pool keeps ticks_in_flight ticks (steps plus a state
        snapshot) enqueued; the host polls the oldest snapshot of each pool
        without blocking, reads back and releases the solves it shows
        finished, and admits queued requests while the device keeps
        stepping."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []
        t0 = time.perf_counter() if clock_start is None else clock_start
        K, G = self.cfg.steps_per_tick, self.cfg.steps_per_graph
        self.cell_updates = 0
        self.rejected = 0
        while pending or queue or any(self.live[k] or self.inflight[k] for k in self.pools):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.pop(0)
                if not self.pools[r.kind].fits(request_sample(r)):
                    r.status = STATUS_REJECTED
                    self.rejected += 1
                    continue
                queue.append(r)
            if queue:
                queue = [r for r in queue if not self.try_admit(r, now)]
            progressed = False
            for k, p in self.pools.items():
                while self.live[k] and len(self.inflight[k]) < self.cfg.ticks_in_flight:
                    self.inflight[k].append(p.tick(K, G))
                    self.issued[k] += 1
                if self.inflight[k]:
                    snap = p.poll(self.inflight[k][0])
                    if snap is not None:
                        ticket = self.inflight[k].pop(0)
                        self._retire(k, snap, ticket, time.perf_counter() - t0)
                        progressed = True
            if not progressed:
                wait = pending[0].arrival - now if pending else 0.0
                time.sleep(min(max(wait, 0.0), 2e-4) if (queue or any(self.live.values()))
                           else min(max(wait, 0.0), 0.05))
        self.elapsed = time.perf_counter() - t0
        return requests


def latency_stats(requests):
    lat = np.array([r.latency for r in requests if r.latency is not None and r.status >= 0])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "count": int(len(lat))}
