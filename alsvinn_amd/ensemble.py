"""Python host interface of the hip platform (thin layer over the C ABI).

`Ensemble` is a batch of solves of one scheme advanced together on one GPU:
the on-device analogue of the reference's per-sample Simulator loop
(alsuq/src/run/Runner.cpp:42-78, alsfvm/src/simulator/Simulator.cpp:77-85).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_scheme(equation, flux, reconstruction, integrator, boundary, gamma=1.4, cfl=None,
                precision=_lib.FP64):
    s = _lib.Scheme()
    s.equation, s.flux, s.reconstruction = equation, flux, reconstruction
    s.integrator, s.boundary, s.precision = integrator, boundary, precision
    s.gamma = gamma
    if cfl is None:
        # SimulatorSetup.cpp:306-320: cfl "auto" is 0.9 without reconstruction, else 0.475
        cfl = 0.9 if reconstruction == _lib.RECON_NONE else 0.475
    s.cfl = cfl
    return s


def make_sample(n, origin, top, end_time, saves=(1,), max_steps=-1):
    """One solve: grid n=(nx,ny,nz) on [origin, top], FixedInterval adjusters
    with `saves` saves each (interval end_time/saves, first save at 1*interval)."""
    p = _lib.Sample()
    p.nx, p.ny, p.nz = n
    for k in range(3):
        p.origin[k] = origin[k]
        p.top[k] = top[k]
    p.end_time = end_time
    p.num_adjusters = len(saves)
    for k, ns in enumerate(saves):
        p.interval[k] = end_time / ns
        p.saved0[k] = 1
    p.max_steps = max_steps
    return p


def state_elements(scheme, samples):
    """Reals of the concatenated sample volumes (the state buffer an
    Ensemble(..., state=ptr) advances in place)."""
    samples = list(samples)
    arr = (_lib.Sample * len(samples))(*samples)
    n = ctypes.c_longlong()
    check(_lib.load().alsvinn_hip_state_elements(ctypes.byref(scheme), len(samples), arr,
                                                 ctypes.byref(n)))
    return n.value


class Ensemble:
    """Samples of one scheme advanced together. state: optional device pointer
    (int) of a caller-owned buffer of state_elements(scheme, samples) reals of
    the scheme's precision, advanced in place (the library never frees it)."""

    def __init__(self, scheme, samples, stream=None, state=None):
        L = _lib.load()
        self.scheme = scheme
        self.samples = list(samples)
        arr = (_lib.Sample * len(self.samples))(*self.samples)
        h = ctypes.c_void_p()
        check(L.alsvinn_hip_ensemble_create_bound(ctypes.byref(scheme), len(self.samples), arr,
                                                  stream, state, ctypes.byref(h)))
        self.h = h
        self.nvar = 1 if scheme.equation == _lib.EQ_BURGERS else scheme.equation + 2

    def close(self):
        if self.h:
            _lib.load().alsvinn_hip_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def layout(self, sample):
        dims = (ctypes.c_int * 6)()
        off = ctypes.c_longlong()
        n = ctypes.c_longlong()
        check(_lib.load().alsvinn_hip_ensemble_layout(self.h, sample, dims, ctypes.byref(off),
                                                      ctypes.byref(n)))
        return tuple(dims), off.value, n.value

    def volume_shape(self, sample):
        (tx, ty, tz, gx, gy, gz), _, _ = self.layout(sample)
        return (self.nvar, tz, ty, tx)

    def upload(self, sample, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == self.volume_shape(sample), (u.shape, self.volume_shape(sample))
        check(_lib.load().alsvinn_hip_ensemble_upload(self.h, sample, _dp(u)))

    def download(self, sample):
        u = np.empty(self.volume_shape(sample))
        check(_lib.load().alsvinn_hip_ensemble_download(self.h, sample, _dp(u)))
        return u

    def apply_boundary(self):
        check(_lib.load().alsvinn_hip_ensemble_apply_boundary(self.h))

    def reset_time(self):
        check(_lib.load().alsvinn_hip_ensemble_reset_time(self.h))

    def set_schedule(self, rows, variant):
        check(_lib.load().alsvinn_hip_ensemble_set_schedule(self.h, rows, variant))

    def schedule(self):
        r, v, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.load().alsvinn_hip_ensemble_get_schedule(self.h, ctypes.byref(r), ctypes.byref(v),
                                                            ctypes.byref(n)))
        waves = (ctypes.c_int * 3)()
        check(_lib.load().alsvinn_hip_ensemble_get_stage_waves(self.h, waves))
        return {"rows": r.value, "variant": v.value, "waves_per_simd": waves[0],
                "stage_waves_per_simd": list(waves), "tasks": n.value}

    def set_small_1d(self, max_rounds):
        """Whole-step launches for small 1D solves: eligible when each sample's
        strips take at most max_rounds turns of 16 wavefronts (0: never).
        Returns the workgroup size in use (0: the multi-launch step)."""
        check(_lib.load().alsvinn_hip_ensemble_set_small_1d(self.h, max_rounds))
        return self.small_1d()

    def small_1d(self):
        t = ctypes.c_int()
        check(_lib.load().alsvinn_hip_ensemble_get_small_1d(self.h, ctypes.byref(t)))
        return t.value

    def autotune(self):
        r, v = ctypes.c_int(), ctypes.c_int()
        check(_lib.load().alsvinn_hip_ensemble_autotune(self.h, ctypes.byref(r), ctypes.byref(v)))
        return self.schedule()

    def step(self, n=1):
        check(_lib.load().alsvinn_hip_ensemble_step(self.h, n))

    def run(self, steps_per_sync=16):
        check(_lib.load().alsvinn_hip_ensemble_run(self.h, steps_per_sync))

    def run_steps(self, nsteps, steps_per_graph=10):
        check(_lib.load().alsvinn_hip_ensemble_run_steps(self.h, nsteps, steps_per_graph))

    def step_timed(self, nsteps):
        ms = np.zeros(3)
        n = (ctypes.c_longlong * 3)()
        check(_lib.load().alsvinn_hip_ensemble_step_timed(self.h, nsteps, _dp(ms), n))
        return ms, list(n)

    def synchronize(self):
        check(_lib.load().alsvinn_hip_ensemble_synchronize(self.h))

    def info(self, sample):
        t = ctypes.c_double()
        steps = ctypes.c_longlong()
        dt = ctypes.c_double()
        st = ctypes.c_int()
        act = ctypes.c_int()
        check(_lib.load().alsvinn_hip_ensemble_sample_info(self.h, sample, ctypes.byref(t),
                                                           ctypes.byref(steps), ctypes.byref(dt),
                                                           ctypes.byref(st), ctypes.byref(act)))
        return {"t": t.value, "steps": steps.value, "last_dt": dt.value, "status": st.value,
                "active": act.value}

    def set_hold(self, adjuster, saved_limit):
        """Samples pause once adjuster's save counter reaches saved_limit
        (None removes the hold); run() returns when every sample is paused
        or finished."""
        check(_lib.load().alsvinn_hip_ensemble_set_hold(self.h, adjuster,
                                                        -1 if saved_limit is None else saved_limit))

    def saves(self, sample):
        out = (ctypes.c_longlong * 2)()
        check(_lib.load().alsvinn_hip_ensemble_saves(self.h, sample, out))
        return list(out)

    def record_dt(self, capacity):
        check(_lib.load().alsvinn_hip_ensemble_record_dt(self.h, capacity))
        self._dt_cap = capacity

    def dt_history(self, sample):
        out = np.zeros(self._dt_cap)
        check(_lib.load().alsvinn_hip_ensemble_dt_history(self.h, sample, _dp(out), self._dt_cap))
        steps = self.info(sample)["steps"]
        return out[: min(steps, self._dt_cap)]

    def rhs(self, sample):
        out = np.zeros(self.volume_shape(sample))
        ws = np.zeros(3)
        check(_lib.load().alsvinn_hip_ensemble_rhs(self.h, sample, _dp(out), _dp(ws)))
        return out, ws

    def interval_ms(self, ref):
        """(start, end) of this ensemble's last step/run call on the device,
        in ms after the start of ref's last call."""
        a = ctypes.c_double()
        b = ctypes.c_double()
        check(_lib.load().alsvinn_hip_ensemble_interval_ms(self.h, ref.h, ctypes.byref(a),
                                                           ctypes.byref(b)))
        return a.value, b.value

    def last_elapsed_ms(self):
        ms = ctypes.c_double()
        check(_lib.load().alsvinn_hip_ensemble_last_elapsed_ms(self.h, ctypes.byref(ms)))
        return ms.value

    # ---- statistics (alsuq::stats on the device)

    def _group(self, samples):
        if samples is None:
            samples = range(len(self.samples))
        arr = np.ascontiguousarray(list(samples), dtype=np.int32)
        return arr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(arr)

    def meanvar(self, samples=None, welford=False):
        arr, ptr, m = self._group(samples)
        shape = self.volume_shape(int(arr[0]))
        mean = np.empty(shape)
        var = np.empty(shape)
        check(_lib.load().alsvinn_hip_stats_meanvar(self.h, ptr, m, int(welford), _dp(mean),
                                                    _dp(var)))
        return mean, var

    def onepoint_moment(self, p, samples=None):
        """OnePointMoment: per cell mean over the samples of u**p."""
        arr, ptr, m = self._group(samples)
        out = np.empty(self.volume_shape(int(arr[0])))
        check(_lib.load().alsvinn_hip_stats_onepoint_moment(self.h, ptr, m, float(p), _dp(out)))
        return out

    def bounded_variation(self, p=1, samples=None):
        """BoundedVariation: per variable mean over the samples of the total
        variation (p-th powers of the gradient magnitudes), 1D and 2D."""
        arr, ptr, m = self._group(samples)
        out = np.zeros(self.nvar)
        check(_lib.load().alsvinn_hip_stats_bounded_variation(self.h, ptr, m, float(p), _dp(out)))
        return out

    def structure_cube(self, number_of_h, p=2.0, samples=None):
        arr, ptr, m = self._group(samples)
        out = np.zeros((self.nvar, number_of_h))
        check(_lib.load().alsvinn_hip_stats_structure_cube(self.h, ptr, m, number_of_h, p, _dp(out)))
        return out

    def structure_basic(self, number_of_h, direction=0, p=2.0, samples=None):
        arr, ptr, m = self._group(samples)
        out = np.zeros((self.nvar, number_of_h))
        check(_lib.load().alsvinn_hip_stats_structure_basic(self.h, ptr, m, number_of_h, p,
                                                            direction, _dp(out)))
        return out

    def states(self):
        """(t, steps, active, status) of every sample (synchronous)."""
        n = len(self.samples)
        t = np.zeros(n)
        steps = np.zeros(n, dtype=np.int64)
        act = np.zeros(n, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        check(_lib.load().alsvinn_hip_ensemble_states(
            self.h, _dp(t), steps.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
            act.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return t, steps, act, st

    def totals(self):
        out = np.zeros((len(self.samples), self.nvar))
        check(_lib.load().alsvinn_hip_ensemble_totals(self.h, _dp(out)))
        return out


def first_fit(allocations, arena, size):
    """Lowest offset where `size` elements fit between the (offset, size)
    allocations inside [0, arena), or None when no gap is large enough."""
    pos = 0
    for off, sz in sorted(allocations):
        if off - pos >= size:
            return pos
        pos = max(pos, off + sz)
    return pos if arena - pos >= size else None


class Pool(Ensemble):
    """Slots over an HBM arena for continuous batching (C ABI pool).

    Solves are placed first-fit by footprint (elements per state buffer) and
    released when their result has been read; the other slots keep stepping.
    Assigning, uploading into, releasing and downloading a slot do not wait
    for the pool's stream; `tick` enqueues steps plus a state snapshot and
    `poll` reads a snapshot without blocking, so the host admits and retires
    solves while the device steps.
    """

    def __init__(self, scheme, nslots, max_n, arena_elements, stream=None):
        L = _lib.load()
        self.scheme = scheme
        self.samples = [None] * nslots
        h = ctypes.c_void_p()
        mx = (ctypes.c_int * 3)(*max_n)
        check(L.alsvinn_hip_pool_create(ctypes.byref(scheme), nslots, mx, arena_elements, stream,
                                        ctypes.byref(h)))
        self.h = h
        self.max_n = tuple(max_n)
        self.nvar = 1 if scheme.equation == _lib.EQ_BURGERS else scheme.equation + 2
        self.ghosts = L.alsvinn_hip_ghost_cells(ctypes.byref(scheme))
        self.arena = arena_elements
        self.allocs = {}  # slot -> (offset, size)

    def footprint(self, n):
        g = self.ghosts
        nx, ny, nz = n
        return self.nvar * (nx + 2 * g) * (ny + (2 * g if ny > 1 else 0)) * (nz + (2 * g if nz > 1 else 0))

    def fits(self, sample):
        """Whether a solve could ever be admitted: its grid within max_n and
        of the pool's dimension, at least the ghost width, its footprint
        within the arena."""
        n = (sample.nx, sample.ny, sample.nz)
        dim = 1 + (n[1] > 1) + (n[2] > 1)
        pool_dim = 1 + (self.max_n[1] > 1) + (self.max_n[2] > 1)
        if dim != pool_dim or any(a > b for a, b in zip(n, self.max_n)):
            return False
        if n[0] < self.ghosts or (n[1] > 1 and n[1] < self.ghosts) or (n[2] > 1 and n[2] < self.ghosts):
            return False
        return self.footprint(n) <= self.arena

    def place(self, size):
        return first_fit(self.allocs.values(), self.arena, size)

    def free_slot(self):
        for s in range(len(self.samples)):
            if s not in self.allocs:
                return s
        return None

    def assign(self, sample):
        """Places a solve; returns its slot or None (no slot / no memory)."""
        slot = self.free_slot()
        if slot is None:
            return None
        size = self.footprint((sample.nx, sample.ny, sample.nz))
        off = self.place(size)
        if off is None:
            return None
        check(_lib.load().alsvinn_hip_pool_assign(self.h, slot, ctypes.byref(sample), off))
        self.allocs[slot] = (off, size)
        self.samples[slot] = sample
        return slot

    def release(self, slot):
        check(_lib.load().alsvinn_hip_pool_release(self.h, slot))
        self.allocs.pop(slot, None)
        self.samples[slot] = None

    def upload(self, slot, u):
        """Initial state of a just-assigned slot, then its ghost cells."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == self.volume_shape(slot), (u.shape, self.volume_shape(slot))
        check(_lib.load().alsvinn_hip_pool_upload(self.h, slot, _dp(u)))
        check(_lib.load().alsvinn_hip_pool_apply_boundary(self.h, slot))

    def download(self, slot):
        """Result of a finished slot (inactive in a snapshot)."""
        out = np.empty(self.volume_shape(slot))
        check(_lib.load().alsvinn_hip_pool_download(self.h, slot, _dp(out)))
        return out

    def tick(self, nsteps, steps_per_graph):
        """Enqueues nsteps steps and a state snapshot; returns its ticket."""
        self.run_steps(nsteps, steps_per_graph)
        t = ctypes.c_longlong()
        check(_lib.load().alsvinn_hip_pool_snapshot(self.h, ctypes.byref(t)))
        return t.value

    def poll(self, ticket, wait=False):
        """(t, steps, active, status) of a snapshot, or None while its copy
        has not run (wait=False)."""
        n = len(self.samples)
        t = np.zeros(n)
        steps = np.zeros(n, dtype=np.int64)
        act = np.zeros(n, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        ready = ctypes.c_int()
        check(_lib.load().alsvinn_hip_pool_poll(
            self.h, ticket, int(wait), _dp(t), steps.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
            act.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(ready)))
        return (t, steps, act, st) if ready.value else None



def device_memory():
    """(free, total) bytes of device memory (hipMemGetInfo)."""
    f = ctypes.c_longlong()
    t = ctypes.c_longlong()
    check(_lib.load().alsvinn_hip_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value
