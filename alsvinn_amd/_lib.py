"""ctypes binding of libalsvinn_hip.so (include/alsvinn_hip.h).

The library must be built in-tree (`python -c "import __graft_entry__ as g; g.build()"`
or `python alsvinn_amd/build.py`). There is no fallback: loading fails loudly
when the library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; ALSVINN_HIP_LIBRARY names another build of the same sources
LIB_PATH = os.environ.get("ALSVINN_HIP_LIBRARY") or os.path.join(HERE, "libalsvinn_hip.so")

EQ_BURGERS, EQ_EULER1, EQ_EULER2, EQ_EULER3 = 0, 1, 2, 3
FLUX_HLL, FLUX_HLL3, FLUX_GODUNOV, FLUX_CENTRAL, FLUX_RUSANOV, FLUX_ROE = 0, 1, 2, 3, 4, 5
RECON_NONE, RECON_WENO2, RECON_ENO2, RECON_MC, RECON_MM, RECON_WENOF2 = 0, 1, 2, 3, 4, 5
RECON_ENO3, RECON_ENO4, RECON_WENO3 = 6, 7, 8
INT_FE, INT_RK2, INT_RK3, INT_RK4 = 0, 1, 2, 3
BC_NEUMANN, BC_PERIODIC = 0, 1
FP64, FP32 = 0, 1

E_OK, E_INVALID, E_HIP, E_CONSTRAINT = 0, 1, 2, 3

# Every function include/alsvinn_hip.h declares.
EXPORTS = [
    "alsvinn_hip_last_error",
    "alsvinn_hip_ghost_cells",
    "alsvinn_hip_ensemble_create",
    "alsvinn_hip_ensemble_destroy",
    "alsvinn_hip_state_elements",
    "alsvinn_hip_ensemble_create_bound",
    "alsvinn_hip_pool_create",
    "alsvinn_hip_pool_assign",
    "alsvinn_hip_pool_release",
    "alsvinn_hip_pool_upload",
    "alsvinn_hip_pool_apply_boundary",
    "alsvinn_hip_pool_download",
    "alsvinn_hip_pool_snapshot",
    "alsvinn_hip_pool_poll",
    "alsvinn_hip_device_memory",
    "alsvinn_hip_ensemble_states",
    "alsvinn_hip_ensemble_layout",
    "alsvinn_hip_ensemble_upload",
    "alsvinn_hip_ensemble_download",
    "alsvinn_hip_ensemble_state_pointer",
    "alsvinn_hip_ensemble_apply_boundary",
    "alsvinn_hip_ensemble_reset_time",
    "alsvinn_hip_ensemble_set_schedule",
    "alsvinn_hip_ensemble_get_schedule",
    "alsvinn_hip_ensemble_autotune",
    "alsvinn_hip_ensemble_get_stage_variants",
    "alsvinn_hip_ensemble_get_stage_waves",
    "alsvinn_hip_ensemble_set_small_1d",
    "alsvinn_hip_ensemble_get_small_1d",
    "alsvinn_hip_ensemble_step",
    "alsvinn_hip_ensemble_run",
    "alsvinn_hip_ensemble_set_hold",
    "alsvinn_hip_ensemble_saves",
    "alsvinn_hip_ensemble_run_steps",
    "alsvinn_hip_ensemble_step_timed",
    "alsvinn_hip_ensemble_synchronize",
    "alsvinn_hip_ensemble_sample_info",
    "alsvinn_hip_ensemble_record_dt",
    "alsvinn_hip_ensemble_dt_history",
    "alsvinn_hip_ensemble_rhs",
    "alsvinn_hip_ensemble_last_elapsed_ms",
    "alsvinn_hip_ensemble_interval_ms",
    "alsvinn_hip_stats_meanvar",
    "alsvinn_hip_stats_onepoint_moment",
    "alsvinn_hip_stats_bounded_variation",
    "alsvinn_hip_stats_structure_cube",
    "alsvinn_hip_stats_structure_basic",
    "alsvinn_hip_ensemble_totals",
    "alsvinn_hip_operator_create",
    "alsvinn_hip_operator_destroy",
    "alsvinn_hip_numflux_compute",
    "alsvinn_hip_reconstruction_compute",
    "alsvinn_hip_boundary_apply",
    "alsvinn_hip_statistics_create",
    "alsvinn_hip_statistics_destroy",
    "alsvinn_hip_statistics_compute",
    "alsvinn_hip_statistics_combine",
    "alsvinn_hip_statistics_finalize",
    "alsvinn_hip_statistics_reset",
    "alsvinn_hip_statistics_outputs",
    "alsvinn_hip_statistics_result",
    "alsvinn_hip_statistics_download",
    "alsvinn_hip_numflux_compute_noalloc",
]

STAT_MEANVAR, STAT_ONEPOINTMOMENT, STAT_STRUCTURE_CUBE, STAT_STRUCTURE_BASIC, STAT_BV = 0, 1, 2, 3, 4


class Scheme(ctypes.Structure):
    _fields_ = [
        ("equation", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("reconstruction", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("boundary", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("cfl", ctypes.c_double),
    ]


class Sample(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("num_adjusters", ctypes.c_int),
        ("origin", ctypes.c_double * 3),
        ("top", ctypes.c_double * 3),
        ("end_time", ctypes.c_double),
        ("interval", ctypes.c_double * 2),
        ("saved0", ctypes.c_longlong * 2),
        ("max_steps", ctypes.c_longlong),
    ]


class Volume(ctypes.Structure):
    """alsvinn_hip_volume: a reference Volume as per-variable device pointers."""
    _fields_ = [
        ("nvar", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("gx", ctypes.c_int),
        ("gy", ctypes.c_int),
        ("gz", ctypes.c_int),
        ("var", ctypes.c_void_p * 5),
    ]


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("alsvinn_hip error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load():
    """Loads the library (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libalsvinn_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    llp = ctypes.POINTER(ctypes.c_longlong)
    ip = ctypes.POINTER(ctypes.c_int)
    L.alsvinn_hip_last_error.restype = ctypes.c_char_p
    L.alsvinn_hip_last_error.argtypes = []
    L.alsvinn_hip_ghost_cells.argtypes = [ctypes.POINTER(Scheme)]
    L.alsvinn_hip_ensemble_create.argtypes = [ctypes.POINTER(Scheme), ctypes.c_int,
                                              ctypes.POINTER(Sample), vp, ctypes.POINTER(vp)]
    L.alsvinn_hip_ensemble_destroy.argtypes = [vp]
    L.alsvinn_hip_ensemble_set_hold.argtypes = [vp, ctypes.c_int, ctypes.c_longlong]
    L.alsvinn_hip_ensemble_saves.argtypes = [vp, ctypes.c_int, llp]
    L.alsvinn_hip_state_elements.argtypes = [ctypes.POINTER(Scheme), ctypes.c_int,
                                             ctypes.POINTER(Sample), llp]
    L.alsvinn_hip_ensemble_create_bound.argtypes = [ctypes.POINTER(Scheme), ctypes.c_int,
                                                    ctypes.POINTER(Sample), vp, vp, ctypes.POINTER(vp)]
    L.alsvinn_hip_pool_create.argtypes = [ctypes.POINTER(Scheme), ctypes.c_int, ip,
                                          ctypes.c_longlong, vp, ctypes.POINTER(vp)]
    L.alsvinn_hip_pool_assign.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Sample), ctypes.c_longlong]
    L.alsvinn_hip_pool_release.argtypes = [vp, ctypes.c_int]
    L.alsvinn_hip_pool_upload.argtypes = [vp, ctypes.c_int, dp]
    L.alsvinn_hip_pool_apply_boundary.argtypes = [vp, ctypes.c_int]
    L.alsvinn_hip_pool_download.argtypes = [vp, ctypes.c_int, dp]
    L.alsvinn_hip_pool_snapshot.argtypes = [vp, llp]
    L.alsvinn_hip_pool_poll.argtypes = [vp, ctypes.c_longlong, ctypes.c_int, dp, llp, ip, ip, ip]
    L.alsvinn_hip_device_memory.argtypes = [llp, llp]
    L.alsvinn_hip_ensemble_states.argtypes = [vp, dp, llp, ip, ip]
    L.alsvinn_hip_ensemble_layout.argtypes = [vp, ctypes.c_int, ip, llp, llp]
    L.alsvinn_hip_ensemble_upload.argtypes = [vp, ctypes.c_int, dp]
    L.alsvinn_hip_ensemble_download.argtypes = [vp, ctypes.c_int, dp]
    L.alsvinn_hip_ensemble_state_pointer.argtypes = [vp, ctypes.POINTER(vp)]
    L.alsvinn_hip_ensemble_apply_boundary.argtypes = [vp]
    L.alsvinn_hip_ensemble_reset_time.argtypes = [vp]
    L.alsvinn_hip_ensemble_set_schedule.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.alsvinn_hip_ensemble_get_schedule.argtypes = [vp, ip, ip, ip]
    L.alsvinn_hip_ensemble_get_stage_variants.argtypes = [vp, ip]
    L.alsvinn_hip_ensemble_get_stage_waves.argtypes = [vp, ip]
    L.alsvinn_hip_ensemble_set_small_1d.argtypes = [vp, ctypes.c_int]
    L.alsvinn_hip_ensemble_get_small_1d.argtypes = [vp, ip]
    L.alsvinn_hip_ensemble_autotune.argtypes = [vp, ip, ip]
    L.alsvinn_hip_ensemble_step.argtypes = [vp, ctypes.c_int]
    L.alsvinn_hip_ensemble_run.argtypes = [vp, ctypes.c_int]
    L.alsvinn_hip_ensemble_run_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.alsvinn_hip_ensemble_step_timed.argtypes = [vp, ctypes.c_int, dp, llp]
    L.alsvinn_hip_ensemble_synchronize.argtypes = [vp]
    L.alsvinn_hip_ensemble_sample_info.argtypes = [vp, ctypes.c_int, dp, llp, dp, ip, ip]
    L.alsvinn_hip_ensemble_record_dt.argtypes = [vp, ctypes.c_longlong]
    L.alsvinn_hip_ensemble_dt_history.argtypes = [vp, ctypes.c_int, dp, ctypes.c_longlong]
    L.alsvinn_hip_ensemble_rhs.argtypes = [vp, ctypes.c_int, dp, dp]
    L.alsvinn_hip_ensemble_last_elapsed_ms.argtypes = [vp, dp]
    L.alsvinn_hip_ensemble_interval_ms.argtypes = [vp, vp, dp, dp]
    L.alsvinn_hip_stats_meanvar.argtypes = [vp, ip, ctypes.c_int, ctypes.c_int, dp, dp]
    L.alsvinn_hip_stats_onepoint_moment.argtypes = [vp, ip, ctypes.c_int, ctypes.c_double, dp]
    L.alsvinn_hip_stats_bounded_variation.argtypes = [vp, ip, ctypes.c_int, ctypes.c_double, dp]
    L.alsvinn_hip_stats_structure_cube.argtypes = [vp, ip, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, dp]
    L.alsvinn_hip_stats_structure_basic.argtypes = [vp, ip, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_double, ctypes.c_int, dp]
    L.alsvinn_hip_ensemble_totals.argtypes = [vp, dp]
    volp = ctypes.POINTER(Volume)
    L.alsvinn_hip_operator_create.argtypes = [ctypes.POINTER(Scheme), ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, vp, ctypes.POINTER(vp)]
    L.alsvinn_hip_operator_destroy.argtypes = [vp]
    L.alsvinn_hip_numflux_compute.argtypes = [vp, volp, dp, ctypes.c_int, volp, ip, ip]
    L.alsvinn_hip_reconstruction_compute.argtypes = [vp, volp, ctypes.c_int, ctypes.c_int, volp,
                                                     volp, ip, ip]
    L.alsvinn_hip_boundary_apply.argtypes = [ctypes.c_int, volp, vp]
    L.alsvinn_hip_statistics_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, volp, vp,
                                                ctypes.POINTER(vp)]
    L.alsvinn_hip_statistics_destroy.argtypes = [vp]
    L.alsvinn_hip_statistics_compute.argtypes = [vp, volp]
    L.alsvinn_hip_statistics_combine.argtypes = [vp, ctypes.c_longlong]
    L.alsvinn_hip_statistics_finalize.argtypes = [vp]
    L.alsvinn_hip_statistics_reset.argtypes = [vp]
    L.alsvinn_hip_statistics_outputs.argtypes = [vp, ip, llp]
    L.alsvinn_hip_statistics_result.argtypes = [vp, ctypes.c_int, ctypes.POINTER(dp)]
    L.alsvinn_hip_statistics_download.argtypes = [vp, ctypes.c_int, dp]
    L.alsvinn_hip_numflux_compute_noalloc.argtypes = [ctypes.POINTER(Scheme), volp, volp, ip, ip,
                                                      vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise HipError(rc, load().alsvinn_hip_last_error().decode())
    return rc
