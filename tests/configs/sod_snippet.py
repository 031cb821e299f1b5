# Sod shock tube, per-cell form: (rho, ux, p) = (1, 0, 1) left of x = 0,
# (0.125, 0, 0.1) right.
if x < 0:
    rho = 1.0
    p = 1.0
else:
    rho = 0.125
    p = 0.1
ux = 0.0
