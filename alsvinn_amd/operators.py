"""The hip platform's operators on a reference Volume, from Python.

A reference `volume::Volume` keeps every conserved variable in its own
memory area of the padded extents (alsfvm/src/volume/Volume.cpp:44-60). The
classes here take such volumes as device pointers, one per variable (any
allocator: hipMalloc, a framework's tensors), and mirror the reference's
interfaces with the same names and argument meaning:

  NumericalFlux.compute_flux            numflux::NumericalFlux::computeFlux
                                        (NumericalFlux.hpp:44-52)
  NumericalFlux.perform_reconstruction  reconstruction::Reconstruction::performReconstruction
                                        (Reconstruction.hpp:46-58)
  apply_boundary_conditions             boundary::Boundary::applyBoundaryConditions
                                        (Boundary.hpp:30-31)
  Statistics.compute_statistics /       alsuq::stats::Statistics (Statistics.hpp:25-62):
  combine_statistics /                  computeStatistics per sample, then combineStatistics
  finalize_statistics                   and finalizeStatistics

Errors raise HipError with the library's message, as the reference THROWs.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def make_volume(nvar, n, ghosts, pointers, precision=_lib.FP64):
    """alsvinn_hip_volume of nvar variables on n = (nx, ny, nz) interior
    cells with `ghosts` = (gx, gy, gz) layers; `pointers` are the variables'
    device addresses (ints)."""
    v = _lib.Volume()
    v.nvar = nvar
    v.precision = precision
    v.nx, v.ny, v.nz = n
    v.gx, v.gy, v.gz = ghosts
    for k, p in enumerate(pointers):
        v.var[k] = int(p)
    return v


def _ivec(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _stream(stream):
    return ctypes.c_void_p(int(stream)) if stream else None


class NumericalFlux:
    """A numerical flux with its reconstruction for one grid
    (NumericalFluxFactory::createNumericalFlux(grid), NumericalFluxFactory.cpp:159-200).
    `stream`: a HIP stream handle (int) or None."""

    def __init__(self, scheme, n, stream=None):
        self.scheme = scheme
        self.n = tuple(n)
        self._h = ctypes.c_void_p()
        check(_lib.load().alsvinn_hip_operator_create(ctypes.byref(scheme), self.n[0], self.n[1],
                                                      self.n[2], _stream(stream),
                                                      ctypes.byref(self._h)))

    def get_number_of_ghost_cells(self):
        return _lib.load().alsvinn_hip_ghost_cells(ctypes.byref(self.scheme))

    def compute_flux(self, conserved, output, start=(0, 0, 0), end=(0, 0, 0),
                     compute_wave_speed=True):
        """Writes the divergence of the face fluxes into the region's cells of
        `output`; returns the max wave speed per direction (3 floats, the
        inactive directions 0)."""
        ws = np.zeros(3)
        check(_lib.load().alsvinn_hip_numflux_compute(
            self._h, ctypes.byref(conserved),
            ws.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(compute_wave_speed),
            ctypes.byref(output), _ivec(start), _ivec(end)))
        return ws

    def perform_reconstruction(self, inp, direction, left, right, start=(0, 0, 0), end=(0, 0, 0),
                               indicator_variable=0):
        check(_lib.load().alsvinn_hip_reconstruction_compute(
            self._h, ctypes.byref(inp), direction, indicator_variable, ctypes.byref(left),
            ctypes.byref(right), _ivec(start), _ivec(end)))

    def close(self):
        if self._h:
            _lib.load().alsvinn_hip_operator_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compute_flux_noalloc(scheme, conserved, output, start=(0, 0, 0), end=(0, 0, 0),
                         device_wave_speeds=None, stream=None):
    """computeFlux without an operator object and without any allocation
    (alsvinn_hip_numflux_compute_noalloc): asynchronous on `stream`; the wave
    speeds go to `device_wave_speeds` (device address of 3 doubles) if given."""
    check(_lib.load().alsvinn_hip_numflux_compute_noalloc(
        ctypes.byref(scheme), ctypes.byref(conserved), ctypes.byref(output), _ivec(start),
        _ivec(end), ctypes.c_void_p(device_wave_speeds) if device_wave_speeds else None,
        _stream(stream)))


def apply_boundary_conditions(boundary, volume, stream=None):
    """Fills every ghost layer of the volume (BC_NEUMANN or BC_PERIODIC)."""
    check(_lib.load().alsvinn_hip_boundary_apply(boundary, ctypes.byref(volume), _stream(stream)))


class Statistics:
    """One statistic of one snapshot, fed a sample at a time: kind is one of
    _lib.STAT_*; `shape` is a Volume with the samples' layout (its pointers
    are not used)."""

    def __init__(self, kind, shape, p=2.0, number_of_h=1, direction=0, boundary=_lib.BC_PERIODIC,
                 stream=None):
        self._h = ctypes.c_void_p()
        check(_lib.load().alsvinn_hip_statistics_create(kind, float(p), int(number_of_h),
                                                        int(direction), int(boundary),
                                                        ctypes.byref(shape), _stream(stream),
                                                        ctypes.byref(self._h)))
        count = ctypes.c_int()
        elements = ctypes.c_longlong()
        check(_lib.load().alsvinn_hip_statistics_outputs(self._h, ctypes.byref(count),
                                                         ctypes.byref(elements)))
        self.outputs, self.elements = count.value, elements.value

    def compute_statistics(self, sample):
        check(_lib.load().alsvinn_hip_statistics_compute(self._h, ctypes.byref(sample)))

    def combine_statistics(self, samples):
        check(_lib.load().alsvinn_hip_statistics_combine(self._h, int(samples)))

    def finalize_statistics(self):
        check(_lib.load().alsvinn_hip_statistics_finalize(self._h))

    def reset(self):
        check(_lib.load().alsvinn_hip_statistics_reset(self._h))

    def result_pointer(self, index=0):
        ptr = ctypes.POINTER(ctypes.c_double)()
        check(_lib.load().alsvinn_hip_statistics_result(self._h, index, ctypes.byref(ptr)))
        return ctypes.cast(ptr, ctypes.c_void_p).value

    def download(self, index=0):
        out = np.empty(self.elements)
        check(_lib.load().alsvinn_hip_statistics_download(
            self._h, index, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def close(self):
        if self._h:
            _lib.load().alsvinn_hip_statistics_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
