"""Initial data of the benchmark configurations, generated on the host.

The reference evaluates initial data on the host too (embedded CPython,
alsfvm/src/init/PythonInitialData.cpp:84-281) at the cell midpoints
x_i = origin + dx*i + dx/2 (Grid.cpp:148-171) and converts primitive to
conserved variables with m = rho*u, E = p/(gamma-1) + (0.5*rho)*(u.u)
(Euler.hpp:236-245). Each generator returns the full padded volume
(nvar, tz, ty, tx) with ghost layers left at zero; the solver fills them
(Simulator::setInitialValue applies the boundary, Simulator.cpp:160).
"""
import numpy as np


def midpoints(n, lo, hi):
    dx = (hi - lo) / n
    return lo + dx * np.arange(n) + dx / 2


def _volume(nvar, n, g):
    nx, ny, nz = n
    gy = g if ny > 1 else 0
    gz = g if nz > 1 else 0
    return np.zeros((nvar, nz + 2 * gz, ny + 2 * gy, nx + 2 * g)), (g, gy, gz)


def _interior(u, gs, n):
    g, gy, gz = gs
    nx, ny, nz = n
    return u[:, gz:gz + nz, gy:gy + ny, g:g + nx]


def conserved_euler(rho, vel, p, gamma):
    """(rho, m..., E) from primitives; vel is a list of velocity arrays."""
    dot = vel[0] * vel[0]
    for v in vel[1:]:
        dot = dot + v * v
    E = p / (gamma - 1) + (0.5 * rho) * dot
    return [rho] + [rho * v for v in vel] + [E]


def sod(n=1024, ghosts=1, gamma=1.4, lo=-1.0, hi=1.0):
    """Sod shock tube: (1, 0, 1) left of x=0, (0.125, 0, 0.1) right."""
    x = midpoints(n, lo, hi)
    rho = np.where(x < 0, 1.0, 0.125)
    p = np.where(x < 0, 1.0, 0.1)
    u, gs = _volume(3, (n, 1, 1), ghosts)
    cons = conserved_euler(rho, [np.zeros(n)], p, gamma)
    inner = _interior(u, gs, (n, 1, 1))
    for k in range(3):
        inner[k, 0, 0, :] = cons[k]
    return u


def kh_perturbation(x, a, b):
    """s(x) = 0.01 * sum_k a_k cos(b_k + 2*pi*k*x) / sum(a), k = 1..len(a)."""
    norm = 0.0
    for ak in a:
        norm += ak
    if abs(norm) < 1e-10:
        norm = 1.0
    S = np.zeros_like(x)
    for k in range(len(a)):
        S = S + a[k] * np.cos(b[k] + (2 * np.pi * (k + 1)) * x)
    return (0.01 * S) / norm


def kh_parameters(X):
    """Maps a 40-vector of U[0,1] draws to (a0, a1, b0, b1): a^j in [0,1],
    b^j in [-pi, pi] via the Uniform scaling x*(pi-(-pi)) + (-pi)."""
    X = np.asarray(X, dtype=np.float64)
    a0, a1 = X[0:10], X[10:20]
    b0 = X[20:30] * (np.pi - (-np.pi)) + (-np.pi)
    b1 = X[30:40] * (np.pi - (-np.pi)) + (-np.pi)
    return a0, a1, b0, b1


def kelvin_helmholtz_2d(n, X, ghosts=2, gamma=1.4):
    """2D Kelvin-Helmholtz on [0,1]^2: rho=2, u=-0.5 inside the band
    0.25+s0(x) < y < 0.75+s1(x), else rho=1, u=0.5; v=0, p=2.5."""
    a0, a1, b0, b1 = kh_parameters(X)
    x = midpoints(n, 0.0, 1.0)
    y = midpoints(n, 0.0, 1.0)
    lo = 0.25 + kh_perturbation(x, a0, b0)
    hi = 0.75 + kh_perturbation(x, a1, b1)
    Y = y[:, None]
    middle = (Y > lo[None, :]) & (Y < hi[None, :])
    rho = np.where(middle, 2.0, 1.0)
    ux = np.where(middle, -0.5, 0.5)
    uy = np.zeros_like(rho)
    p = np.full_like(rho, 2.5)
    cons = conserved_euler(rho, [ux, uy], p, gamma)
    u, gs = _volume(4, (n, n, 1), ghosts)
    inner = _interior(u, gs, (n, n, 1))
    for k in range(4):
        inner[k, 0] = cons[k]
    return u


def kh_parameters_3d(X):
    """400-vector of U[0,1] draws -> (a0, a1, b0, b1), each 10x10 modes."""
    X = np.asarray(X, dtype=np.float64)
    a0, a1 = X[0:100].reshape(10, 10), X[100:200].reshape(10, 10)
    b0 = (X[200:300] * (np.pi - (-np.pi)) + (-np.pi)).reshape(10, 10)
    b1 = (X[300:400] * (np.pi - (-np.pi)) + (-np.pi)).reshape(10, 10)
    return a0, a1, b0, b1


def kh_perturbation_3d(x, z, a, b):
    norm = 0.0
    for ak in a.ravel():
        norm += ak
    if abs(norm) < 1e-10:
        norm = 1.0
    X, Z = np.meshgrid(x, z)  # (nz, nx)
    S = np.zeros_like(X)
    for k in range(a.shape[0]):
        for l in range(a.shape[1]):
            S = S + a[k, l] * np.cos(b[k, l] + (2 * np.pi) * ((k + 1) * X + (l + 1) * Z))
    return (0.01 * S) / norm


def kelvin_helmholtz_3d(n, X, ghosts=2, gamma=1.4):
    """3D Kelvin-Helmholtz on [0,1]^3: the 2D band in y with perturbations
    summed over 10x10 modes in x and z; u = (+-0.5, 0, 0), p = 2.5."""
    a0, a1, b0, b1 = kh_parameters_3d(X)
    x = midpoints(n, 0.0, 1.0)
    lo = 0.25 + kh_perturbation_3d(x, x, a0, b0)  # (nz, nx)
    hi = 0.75 + kh_perturbation_3d(x, x, a1, b1)
    y = midpoints(n, 0.0, 1.0)
    Y = y[None, :, None]
    middle = (Y > lo[:, None, :]) & (Y < hi[:, None, :])  # (nz, ny, nx)
    rho = np.where(middle, 2.0, 1.0)
    ux = np.where(middle, -0.5, 0.5)
    z0 = np.zeros_like(rho)
    p = np.full_like(rho, 2.5)
    cons = conserved_euler(rho, [ux, z0, z0], p, gamma)
    u, gs = _volume(5, (n, n, n), ghosts)
    inner = _interior(u, gs, (n, n, n))
    for k in range(5):
        inner[k] = cons[k]
    return u


def riemann_four_quadrant(n, params, ghosts=2, gamma=1.4):
    """2D four-quadrant Riemann problem on [0,1]^2 split at (0.5, 0.5).
    params: (4, 4) array of (rho, u, v, p) for quadrants
    (x<.5,y<.5), (x>=.5,y<.5), (x<.5,y>=.5), (x>=.5,y>=.5)."""
    params = np.asarray(params, dtype=np.float64).reshape(4, 4)
    x = midpoints(n, 0.0, 1.0)
    X, Y = np.meshgrid(x, x)  # (ny, nx)
    q = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
    rho = params[q, 0]
    vx = params[q, 1]
    vy = params[q, 2]
    p = params[q, 3]
    cons = conserved_euler(rho, [vx, vy], p, gamma)
    u, gs = _volume(4, (n, n, 1), ghosts)
    inner = _interior(u, gs, (n, n, 1))
    for k in range(4):
        inner[k, 0] = cons[k]
    return u


def riemann_parameters(X16):
    """16 U[0,1] draws -> quadrant states: rho, p in [0.1, 1.5], u, v in [-1, 1]."""
    X = np.asarray(X16, dtype=np.float64).reshape(4, 4)
    out = np.empty((4, 4))
    out[:, 0] = X[:, 0] * (1.5 - 0.1) + 0.1
    out[:, 1] = X[:, 1] * (1.0 - (-1.0)) + (-1.0)
    out[:, 2] = X[:, 2] * (1.0 - (-1.0)) + (-1.0)
    out[:, 3] = X[:, 3] * (1.5 - 0.1) + 0.1
    return out


def fbm_rmd(n, hurst, seed):
    """Fractional Brownian motion on [0,1] by random midpoint displacement
    (Fournier-Fussell-Carpenter), sampled at the n cell midpoints by linear
    interpolation of the 2^k+1 grid points (2^k >= n). B(0) = 0."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(np.log2(max(n, 2))))
    m = 1 << k
    B = np.zeros(m + 1)
    B[m] = rng.standard_normal()
    step = m
    level = 1
    while step > 1:
        half = step // 2
        sigma = np.sqrt((1.0 - 2.0 ** (2 * hurst - 2)) / 2.0 ** (2 * hurst * level))
        idx = np.arange(half, m, step)
        B[idx] = 0.5 * (B[idx - half] + B[idx + half]) + sigma * rng.standard_normal(len(idx))
        step = half
        level += 1
    t = np.linspace(0.0, 1.0, m + 1)
    return np.interp(midpoints(n, 0.0, 1.0), t, B)


def burgers_fbm(n, hurst, seed, ghosts=2):
    u, gs = _volume(1, (n, 1, 1), ghosts)
    _interior(u, gs, (n, 1, 1))[0, 0, 0, :] = fbm_rmd(n, hurst, seed)
    return u
