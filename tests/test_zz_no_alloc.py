"""The allocation-free computeFlux (alsvinn_hip_numflux_compute_noalloc)
against the operator entry point alsvinn_hip_numflux_compute, through the C
ABI on hipMalloc'd per-variable volumes: output cells and wave speeds bit for
bit, on two calls in a row (the second on a smaller grid, so the shared
device scratch is rewritten with a shorter work list), plus a 3D FP32 call
and the size limit (rejected on the host, before any GPU work)."""
import sys

import numpy as np
import pytest

from tests import oracle_lib as O
from tests.test_gpu_operators import DevVolume, _bits_equal, _ghosts3, _regions, _valid_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(hip):
    from tests import hip_runtime
    hip_runtime.rt()
    return hip_runtime


def _one(hip, dev, eq, flux, recon, bc, n, region, precision, seed):
    from alsvinn_amd import operators as ops
    setup = O.make_setup(eq, flux, recon, O.INT_RK3, bc, n, (0, 0, 0), (1, 1, 1))
    g = O.ghosts(setup)
    gv = _ghosts3(setup, g)
    rng = np.random.default_rng(seed)
    u = _valid_state(setup, rng)
    prior = rng.standard_normal(u.shape)
    scheme = hip.make_scheme(eq, flux, recon, hip.INT_RK3, bc, gamma=1.4, precision=precision)
    U = DevVolume(dev, u, n, gv, precision)
    A = DevVolume(dev, prior, n, gv, precision)
    B = DevVolume(dev, prior, n, gv, precision)
    ws_dev = dev.DeviceArray(3)
    ws_dev.upload(np.full(3, -1.0))
    start, end = region
    ops.compute_flux_noalloc(scheme, U.vol, A.vol, start, end, device_wave_speeds=ws_dev.address)
    dev.synchronize()
    with ops.NumericalFlux(scheme, n) as nf:
        ws_ref = nf.compute_flux(U.vol, B.vol, start, end)
    _bits_equal(B.host(), A.host(), "noalloc output vs computeFlux")
    dim = 1 + (n[1] > 1) + (n[2] > 1)
    got_ws = ws_dev.download()
    _bits_equal(ws_ref[:dim].copy(), got_ws[:dim].copy(), "noalloc wave speeds")
    assert np.all(got_ws[dim:] == -1.0)  # inactive directions untouched
    return u


def test_noalloc_compute_flux_matches_operator(hip, dev):
    print("CASE noalloc: kh 2d 150x90 whole interior, then 40x30 side region", flush=True)
    _one(hip, dev, O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.BC_PERIODIC, (150, 90, 1),
         ((0, 0, 0), (0, 0, 0)), hip.FP64, 1)
    _one(hip, dev, O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.BC_PERIODIC, (40, 30, 1),
         ((0, 0, 0), (0, -34 + 6, 0)), hip.FP64, 2)
    print("CASE noalloc: 3d fp32 eno2 and 1d wide eno3", flush=True)
    _one(hip, dev, O.EQ_EULER3, O.FLUX_HLL, O.RECON_ENO2, O.BC_NEUMANN, (33, 12, 10),
         ((2, 2, 2), (-2, -2, -2)), hip.FP32, 3)
    _one(hip, dev, O.EQ_EULER1, O.FLUX_HLL3, O.RECON_ENO3, O.BC_PERIODIC, (300, 1, 1),
         ((0, 0, 0), (0, 0, 0)), hip.FP64, 4)
    sys.stdout.flush()


def test_noalloc_rejects_regions_beyond_its_scratch(hip, dev):
    """A 1D grid of 700000 cells needs more work items than the scratch
    array holds: rejected with a message, on the host."""
    from alsvinn_amd import operators as ops
    scheme = hip.make_scheme(O.EQ_BURGERS, O.FLUX_GODUNOV, O.RECON_ENO2, O.INT_RK2, O.BC_PERIODIC)
    n = 700000
    vol = ops.make_volume(1, (n, 1, 1), (2, 0, 0), [0x1000])
    out = ops.make_volume(1, (n, 1, 1), (2, 0, 0), [0x2000])
    with pytest.raises(hip.HipError, match="allocation-free"):
        ops.compute_flux_noalloc(scheme, vol, out)
