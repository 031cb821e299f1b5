# Kelvin-Helmholtz initial data (init_global form): a band 0.25 + s0(x) < y <
# 0.75 + s1(x) of rho = 2, ux = -0.5 in rho = 1, ux = 0.5, uy = 0, p = 2.5,
# s_j(x) = 0.01 * sum_k a_k cos(b_k + 2 pi k x) / sum(a), from the 40-vector a.


def _perturbation(x, amp, phase):
    norm = 0.0
    for v in amp:
        norm += v
    if abs(norm) < 1e-10:
        norm = 1.0
    s = zeros_like(x)
    for k in range(len(amp)):
        s = s + amp[k] * cos(phase[k] + (2 * pi * (k + 1)) * x)
    return (0.01 * s) / norm


def init_global(rho, ux, uy, p, nx, ny, nz, ax, ay, az, bx, by, bz):
    dx = (bx - ax) / nx
    dy = (by - ay) / ny
    x = ax + dx * arange(nx) + dx / 2
    y = ay + dy * arange(ny) + dy / 2
    a0, a1 = a[0:10], a[10:20]
    b0 = a[20:30] * (pi - (-pi)) + (-pi)
    b1 = a[30:40] * (pi - (-pi)) + (-pi)
    lo = 0.25 + _perturbation(x, a0, b0)
    hi = 0.75 + _perturbation(x, a1, b1)
    middle = (y[None, :] > lo[:, None]) & (y[None, :] < hi[:, None])
    rho[:, :, 0] = where(middle, 2.0, 1.0)
    ux[:, :, 0] = where(middle, -0.5, 0.5)
    uy[:, :, 0] = 0.0
    p[:, :, 0] = 2.5
