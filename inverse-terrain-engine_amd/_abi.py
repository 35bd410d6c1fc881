"""ctypes mirror of include/itr_hip.h and the loader of the HIP library.

The library is loaded from this package's ``_lib/`` directory (built in-tree by
``build.py``).  There is no fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# ITR_HIP_LIB selects another build of the same library (for example a tuning variant).
LIB_PATH = os.environ.get("ITR_HIP_LIB") or os.path.join(PKG, "_lib", "libitr_hip.so")

ITR_OK, ITR_DECLINED, ITR_EINVAL, ITR_EDEVICE, ITR_ENOMEM = range(5)
ITR_PHYSICS_HLL, ITR_PHYSICS_LOCAL_INERTIAL = 0, 1
STATUS_NAMES = {0: "ok", 1: "declined", 2: "invalid argument", 3: "device error", 4: "out of memory"}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class SolverParamsC(C.Structure):
    _fields_ = [("duration_s", C.c_double), ("sync_interval_s", C.c_double), ("cfl", C.c_double),
                ("dt_max_s", C.c_double), ("h_dry", C.c_double), ("h_eps", C.c_double),
                ("infil_capacity", C.c_double), ("initial_depth", C.c_double), ("initial_stage", C.c_double),
                ("cfl_hard", C.c_double), ("inflow_scale", C.c_double), ("rain_t", _dp), ("rain_r", _dp),
                ("locked_dt", _dp), ("n_locked", C.c_int64), ("max_steps", C.c_int64), ("n_rain", C.c_int32),
                ("has_initial_stage", C.c_int32), ("skip_dry", C.c_int32), ("physics", C.c_int32)]


class DesignC(C.Structure):
    _fields_ = [("prim_start", _ip), ("prim", _dp), ("bbox", _ip), ("z_fine", _dp),
                ("wall", C.POINTER(C.c_uint8)), ("editable", C.POINTER(C.c_uint8)),
                ("origin_x", C.c_double), ("origin_y", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("max_abs_dz", C.c_double), ("z_wall", C.c_double), ("volumes", _dp),
                ("level", C.c_int32), ("fine_nx", C.c_int32), ("fine_ny", C.c_int32), ("reserved", C.c_int32)]


class ReplayC(C.Structure):
    _fields_ = [("mode", C.c_int32), ("c0", C.c_int32), ("r0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32),
                ("reserved", C.c_int32), ("ring", _fp), ("band", _fp), ("ring_steps", C.c_int64),
                ("band_syncs", C.c_int64), ("tol_h", C.c_double), ("tol_q", C.c_double),
                ("n_steps_out", C.POINTER(C.c_int64)), ("n_syncs_out", C.POINTER(C.c_int64))]


class GridC(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dy", C.c_double), ("g", C.c_double),
                ("z_ref", C.c_double), ("crest", C.c_int32), ("wet_count", C.c_int32), ("wet", _fp), ("gn2", _fp),
                ("ghost", _ip), ("n_cand", C.c_int32), ("n_mon", C.c_int32), ("dzx", _fp), ("dzy", _fp),
                ("dzx_r", _fp), ("dzy_r", _fp), ("z_level", _dp), ("mon_cells", _up), ("baseline_max", _dp),
                ("guard_pos", _ip), ("n_guard", C.c_int32), ("reserved", C.c_int32), ("guard_tol", C.c_double),
                ("ghost_value", _dp), ("ghost_hydro", _ip), ("hydro_start", _ip), ("hydro_t", _dp), ("hydro_q", _dp),
                ("n_hydro", C.c_int32), ("reserved2", C.c_int32), ("params", C.POINTER(SolverParamsC)),
                ("design", C.POINTER(DesignC)), ("replay", C.POINTER(ReplayC))]


class RunOutC(C.Structure):
    _fields_ = [("t_end", C.c_double), ("volume_m3", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("dt_mean", C.c_double), ("lts", C.c_double), ("v0", C.c_double), ("rel_residual", C.c_double),
                ("abort_value", C.c_double), ("steps", C.c_uint64), ("cell_updates", C.c_uint64),
                ("rows_skipped", C.c_uint64), ("bad", C.c_int32), ("aborted", C.c_int32), ("abort_cell", C.c_int32),
                ("abort_sync", C.c_int32), ("sync_count", C.c_int32), ("n_ledger", C.c_int32),
                ("monitor_max", _dp), ("monitor_tmax", _dp), ("depth", _fp), ("qx", _fp), ("qy", _fp),
                ("ledger", _dp), ("max_ledger", C.c_int32), ("reserved", C.c_int32), ("dt_fallback_t", C.c_double)]


class StatsC(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("device_ms", C.c_double), ("cell_steps", C.c_double),
                ("launches", C.c_uint64), ("syncs", C.c_uint64), ("band_rows", C.c_int32),
                ("waves_per_task", C.c_int32), ("steps_per_chunk", C.c_int32), ("use_graph", C.c_int32),
                ("sim_s", C.c_double), ("launch_lists", C.c_int32), ("reserved", C.c_int32)]


class ObjectiveC(C.Structure):
    _fields_ = [("n_assets", C.c_int32), ("reserved", C.c_int32), ("asset_start", _ip), ("asset_pos", _ip),
                ("threshold_m", _dp), ("weight", _dp), ("tau_m", C.c_double)]


class EarthworkC(C.Structure):
    _fields_ = [("cut_m3", C.c_double), ("fill_m3", C.c_double), ("cut_cost", C.c_double), ("fill_cost", C.c_double),
                ("earth_weight", C.c_double)]


class AdjointOutC(C.Structure):
    _fields_ = [("j_risk", C.c_double), ("peak_m", _dp), ("dJ_dz_face_x", _dp), ("dJ_dz_face_y", _dp),
                ("dJ_dz_cell", _dp), ("steps", C.c_uint64), ("segment_steps", C.c_int32), ("checkpoints", C.c_int32),
                ("replay_mismatches", C.c_uint64), ("plain_forward_ms", C.c_double), ("forward_ms", C.c_double),
                ("replay_ms", C.c_double), ("reverse_ms", C.c_double), ("dJ_drise_face_x", _dp),
                ("dJ_drise_face_y", _dp)]


class TuneResultC(C.Structure):
    _fields_ = [("band_rows", C.c_int32), ("reserved", C.c_int32), ("us_per_step", C.c_double)]


SYNC_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_uint32, _dp,
                      C.c_int32, _dp)

EXPORTS = ["itr_hip_create", "itr_hip_destroy", "itr_hip_device_info", "itr_hip_set_option", "itr_hip_run_batch",
           "itr_hip_run_batch_noalloc",
           "itr_hip_batch_create", "itr_hip_batch_run", "itr_hip_batch_results", "itr_hip_batch_stats",
           "itr_hip_batch_destroy", "itr_hip_batch_advance", "itr_hip_adjoint",
           "itr_hip_dt_schedule", "itr_hip_autotune", "itr_hip_math_selfcheck", "itr_hip_objective_j"]

_lib = None


def load():
    """Load the in-tree library (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with inverse-terrain-engine_amd/build.py "
                      "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.itr_hip_create.restype = C.c_int32
    L.itr_hip_create.argtypes = [C.c_int32, C.POINTER(vp), C.c_char_p, C.c_size_t]
    L.itr_hip_destroy.restype = None
    L.itr_hip_destroy.argtypes = [vp]
    L.itr_hip_device_info.restype = C.c_char_p
    L.itr_hip_device_info.argtypes = [vp]
    L.itr_hip_set_option.restype = C.c_int32
    L.itr_hip_set_option.argtypes = [vp, C.c_char_p, C.c_int64]
    L.itr_hip_run_batch.restype = C.c_int32
    L.itr_hip_run_batch.argtypes = [vp, C.POINTER(SolverParamsC), C.POINTER(GridC), C.c_int32, SYNC_FN, vp,
                                    C.POINTER(RunOutC), C.c_char_p, C.c_size_t]
    L.itr_hip_run_batch_noalloc.restype = C.c_int32
    L.itr_hip_run_batch_noalloc.argtypes = L.itr_hip_run_batch.argtypes
    L.itr_hip_batch_create.restype = C.c_int32
    L.itr_hip_batch_create.argtypes = [vp, C.POINTER(SolverParamsC), C.POINTER(GridC), C.c_int32, C.POINTER(vp),
                                       C.c_char_p, C.c_size_t]
    L.itr_hip_batch_run.restype = C.c_int32
    L.itr_hip_batch_run.argtypes = [vp, SYNC_FN, vp, C.c_char_p, C.c_size_t]
    L.itr_hip_batch_results.restype = C.c_int32
    L.itr_hip_batch_results.argtypes = [vp, C.POINTER(RunOutC), C.c_char_p, C.c_size_t]
    L.itr_hip_batch_stats.restype = C.c_int32
    L.itr_hip_batch_stats.argtypes = [vp, C.POINTER(StatsC)]
    L.itr_hip_batch_advance.restype = C.c_int32
    L.itr_hip_batch_advance.argtypes = [vp, C.c_int64, C.c_char_p, C.c_size_t]
    L.itr_hip_adjoint.restype = C.c_int32
    L.itr_hip_adjoint.argtypes = [vp, C.POINTER(SolverParamsC), C.POINTER(GridC), C.POINTER(ObjectiveC),
                                  C.POINTER(AdjointOutC), C.c_char_p, C.c_size_t]
    L.itr_hip_dt_schedule.restype = C.c_int32
    L.itr_hip_dt_schedule.argtypes = [vp, C.POINTER(SolverParamsC), C.POINTER(GridC), C.c_int32, _dp, C.c_int64,
                                      C.POINTER(C.c_int64), C.c_char_p, C.c_size_t]
    L.itr_hip_autotune.restype = C.c_int32
    L.itr_hip_autotune.argtypes = [vp, C.POINTER(SolverParamsC), C.POINTER(GridC), C.c_int32, C.c_int32,
                                   C.POINTER(TuneResultC), C.c_char_p, C.c_size_t]
    L.itr_hip_math_selfcheck.restype = C.c_int32
    L.itr_hip_math_selfcheck.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]
    L.itr_hip_objective_j.restype = C.c_int32
    L.itr_hip_objective_j.argtypes = [C.POINTER(ObjectiveC), C.POINTER(EarthworkC), _dp, C.c_int32, _dp, _dp, _dp,
                                      C.c_char_p, C.c_size_t]
    L.itr_hip_batch_destroy.restype = None
    L.itr_hip_batch_destroy.argtypes = [vp]
    _lib = L
    return L
