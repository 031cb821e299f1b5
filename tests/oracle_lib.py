"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

EQ_BURGERS, EQ_EULER1, EQ_EULER2, EQ_EULER3 = 0, 1, 2, 3
FLUX_HLL, FLUX_HLL3, FLUX_GODUNOV, FLUX_CENTRAL, FLUX_RUSANOV, FLUX_ROE = 0, 1, 2, 3, 4, 5
RECON_NONE, RECON_WENO2, RECON_ENO2, RECON_MC, RECON_MM, RECON_WENOF2 = 0, 1, 2, 3, 4, 5
RECON_ENO3, RECON_ENO4, RECON_WENO3 = 6, 7, 8
INT_FE, INT_RK2, INT_RK3, INT_RK4 = 0, 1, 2, 3
BC_NEUMANN, BC_PERIODIC = 0, 1


class OracleSetup(ctypes.Structure):
    _fields_ = [
        ("equation", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("recon", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("boundary", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("origin", ctypes.c_double * 3),
        ("top", ctypes.c_double * 3),
        ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("end_time", ctypes.c_double),
        ("num_adjusters", ctypes.c_int),
        ("adj_interval", ctypes.c_double * 2),
        ("adj_saved0", ctypes.c_long * 2),
        ("max_steps", ctypes.c_long),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_long)
        sp = ctypes.POINTER(OracleSetup)
        L.oracle_ghost_cells.argtypes = [sp]
        L.oracle_nvar.argtypes = [ctypes.c_int]
        L.oracle_run.argtypes = [sp, dp, dp, dp, lp, dp, ctypes.c_long]
        L.oracle_system_rhs.argtypes = [sp, dp, dp, dp]
        L.oracle_apply_boundary.argtypes = [sp, dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_compute_flux.argtypes = [sp, dp, ip, ip, dp, dp]
        L.oracle_reconstruct.argtypes = [sp, dp, ctypes.c_int, ip, ip, dp, dp]
        L.oracle_obeys_constraints.argtypes = [sp, dp]
        L.oracle_face_flux.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, dp, dp, dp]
        L.oracle_face_flux.restype = ctypes.c_double
        L.oracle_point_flux.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp]
        L.oracle_reconstruct_1d.argtypes = [ctypes.c_int] * 3 + [dp, dp, dp, ctypes.c_double]
        L.oracle_compute_timestep.argtypes = [dp, dp, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_int, dp, lp]
        L.oracle_compute_timestep.restype = ctypes.c_double
        L.oracle_stl_mersenne_uniform.argtypes = [ctypes.c_long] * 3 + [ctypes.c_double] * 2 + [dp]
        L.oracle_mt19937_64_nth.argtypes = [ctypes.c_long]
        L.oracle_mt19937_64_nth.restype = ctypes.c_ulonglong
        L.oracle_meanvar.argtypes = [ctypes.c_long, ctypes.c_long, dp, dp, dp]
        L.oracle_onepoint_moment.argtypes = [ctypes.c_long, ctypes.c_long, dp, ctypes.c_double, dp]
        L.oracle_total_variation.argtypes = [ctypes.c_int] * 6 + [dp, ctypes.c_double]
        L.oracle_total_variation.restype = ctypes.c_double
        L.oracle_structure_cube.argtypes = [ctypes.c_int] * 5 + [dp, ctypes.c_int, ctypes.c_double,
                                                                ctypes.c_int, dp]
        L.oracle_structure_basic.argtypes = [ctypes.c_int] * 5 + [dp, ctypes.c_int, ctypes.c_double,
                                                                 ctypes.c_int, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_setup(equation, flux, recon, integrator, boundary, n, origin, top, gamma=1.4,
               cfl=0.9, end_time=1.0, saves=(1,), max_steps=-1, intervals=None):
    """intervals: explicit adjuster intervals (default end_time / saves[k]),
    e.g. to stop a run at an intermediate save of a longer schedule."""
    s = OracleSetup()
    s.equation, s.flux, s.recon, s.integrator, s.boundary = equation, flux, recon, integrator, boundary
    s.nx, s.ny, s.nz = n
    for k in range(3):
        s.origin[k] = origin[k]
        s.top[k] = top[k]
    s.gamma, s.cfl, s.end_time = gamma, cfl, end_time
    s.num_adjusters = len(saves)
    for k, ns in enumerate(saves):
        s.adj_interval[k] = end_time / ns if intervals is None else intervals[k]
        s.adj_saved0[k] = 1
    s.max_steps = max_steps
    return s


def nvar(setup):
    return lib().oracle_nvar(setup.equation)


def ghosts(setup):
    return lib().oracle_ghost_cells(ctypes.byref(setup))


def volume_shape(setup):
    """(nvar, tz, ty, tx) of a full volume incl. ghost layers."""
    g = ghosts(setup)
    tx = setup.nx + 2 * g
    ty = setup.ny + (2 * g if setup.ny > 1 else 0)
    tz = setup.nz + (2 * g if setup.nz > 1 else 0)
    return (nvar(setup), tz, ty, tx)


def run(setup, u0, dt_cap=100000):
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    assert u0.shape == volume_shape(setup)
    out = np.empty_like(u0)
    t = ctypes.c_double()
    steps = ctypes.c_long()
    dts = np.zeros(dt_cap)
    rc = lib().oracle_run(ctypes.byref(setup), _dp(u0), _dp(out), ctypes.byref(t),
                          ctypes.byref(steps), _dp(dts), dt_cap)
    return rc, out, t.value, steps.value, dts[: min(steps.value, dt_cap)]


def system_rhs(setup, u, out_init=None):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u) if out_init is None else np.array(out_init, dtype=np.float64)
    ws = np.zeros(3)
    rc = lib().oracle_system_rhs(ctypes.byref(setup), _dp(u), _dp(out), _dp(ws))
    assert rc == 0
    return out, ws


def _ivec(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def compute_flux(setup, u, start=(0, 0, 0), end=(0, 0, 0), out=None):
    """NumericalFlux::computeFlux over the region (start, end); `out` holds
    the output volume's prior contents (zeros by default). Returns (out, ws)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u) if out is None else np.array(out, dtype=np.float64, copy=True)
    ws = np.zeros(3)
    rc = lib().oracle_compute_flux(ctypes.byref(setup), _dp(u), _ivec(start), _ivec(end), _dp(out),
                                   _dp(ws))
    assert rc == 0
    return out, ws


def ghosts3(setup):
    """Ghost layers per axis (none on inactive axes)."""
    g = ghosts(setup)
    return (g, g if setup.ny > 1 else 0, g if setup.nz > 1 else 0)


def conserved_system_regions(setup):
    """The computeFlux regions (start, end) of ConservedSystem.cpp:31-103:
    the inner region (ghost layers cut off each active axis), then for each
    active direction d the low side (end[d] = -size[d] + 3g) and the high
    side (start[d] = size[d] - 3g), size the padded extent."""
    dim = 1 + (setup.ny > 1) + (setup.nz > 1)
    gv = ghosts3(setup)
    size = (setup.nx + 2 * gv[0], setup.ny + 2 * gv[1], setup.nz + 2 * gv[2])
    regs = [(gv, tuple(-x for x in gv))]
    for d in range(dim):
        eb = [0, 0, 0]
        eb[d] = -size[d] + 3 * gv[d]
        st = [0, 0, 0]
        st[d] = size[d] - 3 * gv[d]
        regs.append(((0, 0, 0), tuple(eb)))
        regs.append((tuple(st), (0, 0, 0)))
    return regs


def reconstruct(setup, u, direction, start=(0, 0, 0), end=(0, 0, 0), left=None, right=None):
    """Reconstruction::performReconstruction of the setup's reconstruction;
    left/right hold the prior contents (zeros by default)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    left = np.zeros_like(u) if left is None else np.array(left, dtype=np.float64, copy=True)
    right = np.zeros_like(u) if right is None else np.array(right, dtype=np.float64, copy=True)
    rc = lib().oracle_reconstruct(ctypes.byref(setup), _dp(u), direction, _ivec(start), _ivec(end),
                                  _dp(left), _dp(right))
    assert rc == 0
    return left, right


def apply_boundary(setup, u):
    u = np.array(u, dtype=np.float64)
    assert lib().oracle_apply_boundary(ctypes.byref(setup), _dp(u)) == 0
    return u


def obeys_constraints(setup, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().oracle_obeys_constraints(ctypes.byref(setup), _dp(u))


def face_flux(equation, flux, direction, gamma, qL, qR):
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    F = np.zeros(len(qL))
    ws = lib().oracle_face_flux(equation, flux, direction, gamma, _dp(qL), _dp(qR), _dp(F))
    return F, ws


def point_flux(nsd, direction, gamma, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    F = np.zeros(len(q))
    lib().oracle_point_flux(nsd, direction, gamma, _dp(q), _dp(F))
    return F


def reconstruct_1d(recon, q, gamma=1.4):
    q = np.ascontiguousarray(q, dtype=np.float64)
    nv, n = q.shape
    left = np.full_like(q, np.nan)
    right = np.full_like(q, np.nan)
    assert lib().oracle_reconstruct_1d(recon, nv, n, _dp(q), _dp(left), _dp(right), gamma) == 0
    return left, right


def compute_timestep(ws3, dx3, cfl, t, intervals=(), saved=()):
    ws3 = np.ascontiguousarray(ws3, dtype=np.float64)
    dx3 = np.ascontiguousarray(dx3, dtype=np.float64)
    iv = np.ascontiguousarray(list(intervals) + [0.0], dtype=np.float64)
    sv = np.ascontiguousarray(list(saved) + [0], dtype=np.int64)
    return lib().oracle_compute_timestep(_dp(ws3), _dp(dx3), cfl, t, len(intervals), _dp(iv),
                                         sv.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))


def stl_mersenne_uniform(nsamples, dim, lower=0.0, upper=1.0, sample_start=0):
    out = np.zeros((nsamples, dim))
    lib().oracle_stl_mersenne_uniform(sample_start, nsamples, dim, lower, upper, _dp(out))
    return out


def mt19937_64_nth(n):
    return lib().oracle_mt19937_64_nth(n)


def meanvar(samples):
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    M = samples.shape[0]
    flat = samples.reshape(M, -1)
    mean = np.zeros(flat.shape[1])
    var = np.zeros(flat.shape[1])
    lib().oracle_meanvar(M, flat.shape[1], _dp(flat), _dp(mean), _dp(var))
    return mean.reshape(samples.shape[1:]), var.reshape(samples.shape[1:])


def onepoint_moment(samples, p):
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    M = samples.shape[0]
    flat = samples.reshape(M, -1)
    out = np.zeros(flat.shape[1])
    lib().oracle_onepoint_moment(M, flat.shape[1], _dp(flat), float(p), _dp(out))
    return out.reshape(samples.shape[1:])


def total_variation(u, ghosts, p):
    """u: one variable (tz, ty, tx) incl. ghost layers ghosts = (gx, gy, gz)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    tz, ty, tx = u.shape
    return lib().oracle_total_variation(tx, ty, tz, ghosts[0], ghosts[1], ghosts[2], _dp(u),
                                        float(p))


def bounded_variation(samples, ghosts, p):
    """BoundedVariation over samples (M, nvar, tz, ty, tx): per variable, the
    sum over samples in order of the total variation, divided by M."""
    M, nvar = samples.shape[:2]
    out = np.zeros(nvar)
    for var in range(nvar):
        acc = 0.0
        for k in range(M):
            acc += total_variation(samples[k, var], ghosts, p)
        out[var] = acc / M
    return out


def structure_cube(u, n, ng, H, p, periodic, out=None):
    u = np.ascontiguousarray(u, dtype=np.float64)
    nv = u.shape[0]
    if out is None:
        out = np.zeros((nv, H))
    lib().oracle_structure_cube(nv, n[0], n[1], n[2], ng, _dp(u), H, p, int(periodic), _dp(out))
    return out


def structure_basic(u, n, ng, H, p, direction, periodic, out=None):
    u = np.ascontiguousarray(u, dtype=np.float64)
    nv = u.shape[0]
    if out is None:
        out = np.zeros((nv, H))
    lib().oracle_structure_basic(nv, n[0], n[1], n[2], ng, _dp(u), H, p, direction,
                                 int(periodic), _dp(out))
    return out
