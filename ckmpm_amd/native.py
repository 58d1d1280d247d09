"""ctypes binding of libckmpm.so (the C-ABI declared in include/ckmpm.h).

The library is loaded from this package directory only; a missing library raises
immediately (there is no fallback implementation).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libckmpm.so")

OK, ERR_INVALID, ERR_HIP, ERR_CAPACITY, ERR_NONFINITE, ERR_OUT_OF_WINDOW, ERR_OUT_OF_DOMAIN, ERR_NO_DEVICE = range(8)
ERROR_NAMES = {
    OK: "OK", ERR_INVALID: "INVALID", ERR_HIP: "HIP", ERR_CAPACITY: "CAPACITY", ERR_NONFINITE: "NONFINITE",
    ERR_OUT_OF_WINDOW: "OUT_OF_WINDOW", ERR_OUT_OF_DOMAIN: "OUT_OF_DOMAIN", ERR_NO_DEVICE: "NO_DEVICE",
}
FLAG_OUT_OF_WINDOW, FLAG_OUT_OF_DOMAIN, FLAG_NONFINITE, FLAG_CAPACITY, FLAG_MISSING_BLOCK = 1, 2, 4, 8, 16
MOVE_SHIFT = 27
SLOT_MASK = (1 << MOVE_SHIFT) - 1

# every entry point declared in include/ckmpm.h
EXPORTS = [
    "ckmpm_create", "ckmpm_add_model", "ckmpm_initial_setup", "ckmpm_substep", "ckmpm_run_frames",
    "ckmpm_synchronize", "ckmpm_get_counters", "ckmpm_clear_flags", "ckmpm_get_positions", "ckmpm_model_info",
    "ckmpm_export_layout", "ckmpm_export_ordering", "ckmpm_export_state", "ckmpm_import_state",
    "ckmpm_export_grid_m", "ckmpm_export_grid_v", "ckmpm_set_profiling", "ckmpm_get_phase_times",
    "ckmpm_set_conservation", "ckmpm_get_conservation", "ckmpm_set_boundary", "ckmpm_set_rigid",
    "ckmpm_destroy", "ckmpm_last_error", "ckmpm_autotune", "ckmpm_set_tuning", "ckmpm_get_tuning",
    "ckmpm_hash_create", "ckmpm_hash_build", "ckmpm_hash_build_device", "ckmpm_hash_find",
    "ckmpm_hash_find_device", "ckmpm_hash_synchronize", "ckmpm_hash_info", "ckmpm_hash_export",
    "ckmpm_hash_destroy", "ckmpm_hash_last_error", "ckmpm_hash_build_noalloc",
    "ckmpm_batch_create", "ckmpm_batch_run", "ckmpm_batch_synchronize", "ckmpm_batch_captures",
    "ckmpm_batch_autotune", "ckmpm_batch_get_serial", "ckmpm_batch_set_serial",
    "ckmpm_batch_destroy", "ckmpm_batch_last_error",
    "ckmpm_adjoint_enable", "ckmpm_checkpoint_save", "ckmpm_checkpoint_restore", "ckmpm_checkpoint_free",
    "ckmpm_adjoint_record", "ckmpm_adjoint_reset", "ckmpm_adjoint_seed", "ckmpm_adjoint_substep",
    "ckmpm_adjoint_gradients", "ckmpm_get_positions_input_order", "ckmpm_adjoint_rewind",
]


class DomainDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32 * 3), ("offset", ctypes.c_int32 * 3)]


class SimDesc(ctypes.Structure):
    _fields_ = [("dx", ctypes.c_float), ("gravity", ctypes.c_float), ("mass_clamp", ctypes.c_float),
                ("cfl", ctypes.c_float), ("dt", ctypes.c_float), ("fps", ctypes.c_uint32),
                ("block_capacity", ctypes.c_uint32)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("lam", ctypes.c_float), ("mu", ctypes.c_float), ("bulk", ctypes.c_float),
                ("gamma", ctypes.c_float), ("viscosity", ctypes.c_float), ("J0", ctypes.c_float),
                ("friction_angle_deg", ctypes.c_float), ("yield_stress", ctypes.c_float), ("alpha0", ctypes.c_float),
                ("beta", ctypes.c_float), ("hardening_factor", ctypes.c_float), ("m", ctypes.c_float),
                ("hardening_on", ctypes.c_int32)]


class Region(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_int32 * 3), ("hi", ctypes.c_int32 * 3), ("kind", ctypes.c_int32),
                ("value", ctypes.c_float * 3), ("center", ctypes.c_float * 3)]


MAX_REGIONS = 16


class BoundaryDesc(ctypes.Structure):
    _fields_ = [("walls", ctypes.c_int32), ("nregions", ctypes.c_uint32), ("regions", Region * MAX_REGIONS)]


class HashInfo(ctypes.Structure):
    _fields_ = [("buckets", ctypes.c_uint32), ("slots_per_bucket", ctypes.c_uint32),
                ("stash_capacity", ctypes.c_uint32), ("keys", ctypes.c_uint32), ("occupied_slots", ctypes.c_uint32),
                ("stash", ctypes.c_uint32), ("seeds", ctypes.c_uint32 * 3), ("fill_histogram", ctypes.c_uint32 * 17)]


class Counters(ctypes.Structure):
    _fields_ = [("active_blocks", ctypes.c_uint32), ("neighbor_blocks", ctypes.c_uint32),
                ("exterior_blocks", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("max_speed", ctypes.c_float),
                ("p2g_global_deposits", ctypes.c_uint32), ("substeps", ctypes.c_uint64)]


_lib = None
_vp = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)


def load():
    """Load libckmpm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libckmpm.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    # One HIP runtime per process: torch bundles its own libamdhip64.so.7 (same soname
    # as /opt/rocm's).  Whichever is loaded first serves both; torch only initialises
    # with its own, so let it load first when it is installed.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    L.ckmpm_create.argtypes = [ctypes.POINTER(DomainDesc), ctypes.POINTER(SimDesc), ctypes.POINTER(_vp)]
    L.ckmpm_add_model.argtypes = [_vp, ctypes.POINTER(MaterialDesc), ctypes.c_float, ctypes.c_float, _f32p,
                                  ctypes.c_uint64, _f32p, ctypes.c_uint64, _f32p, _f32p, ctypes.POINTER(ctypes.c_int32)]
    L.ckmpm_initial_setup.argtypes = [_vp]
    L.ckmpm_substep.argtypes = [_vp, ctypes.c_double, ctypes.c_double]
    L.ckmpm_run_frames.argtypes = [_vp, ctypes.c_uint32]
    L.ckmpm_synchronize.argtypes = [_vp]
    L.ckmpm_get_counters.argtypes = [_vp, ctypes.POINTER(Counters)]
    L.ckmpm_clear_flags.argtypes = [_vp]
    L.ckmpm_get_positions.argtypes = [_vp, ctypes.c_int32, _f32p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.ckmpm_model_info.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32)]
    L.ckmpm_export_layout.argtypes = [_vp, ctypes.c_int32, _u32p, ctypes.c_uint32, _u32p]
    L.ckmpm_export_ordering.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _u32p, _u32p, ctypes.c_uint32]
    L.ckmpm_export_state.argtypes = [_vp, ctypes.c_int32, _f32p]
    L.ckmpm_import_state.argtypes = [_vp, ctypes.c_int32, _f32p]
    L.ckmpm_export_grid_m.argtypes = [_vp, _f32p, ctypes.c_uint64]
    L.ckmpm_export_grid_v.argtypes = [_vp, _f32p, ctypes.c_uint64]
    L.ckmpm_set_profiling.argtypes = [_vp, ctypes.c_int32]
    L.ckmpm_set_conservation.argtypes = [_vp, ctypes.c_uint32]
    L.ckmpm_set_boundary.argtypes = [_vp, ctypes.POINTER(BoundaryDesc)]
    L.ckmpm_get_conservation.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_float), ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]
    L.ckmpm_get_phase_times.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]
    L.ckmpm_destroy.argtypes = [_vp]
    L.ckmpm_set_rigid.argtypes = [_vp, ctypes.c_int32, ctypes.c_uint64, _f32p]
    L.ckmpm_autotune.argtypes = [_vp, ctypes.c_uint32]
    L.ckmpm_set_tuning.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32]
    L.ckmpm_get_tuning.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.ckmpm_last_error.argtypes = [_vp]
    L.ckmpm_last_error.restype = ctypes.c_char_p
    _u64p = ctypes.POINTER(ctypes.c_uint64)
    L.ckmpm_hash_create.argtypes = [ctypes.POINTER(DomainDesc), ctypes.c_uint32, ctypes.POINTER(_vp)]
    L.ckmpm_hash_build.argtypes = [_vp, _u32p, _u32p, ctypes.c_uint32]
    L.ckmpm_hash_build_device.argtypes = [_vp, _vp, _vp, ctypes.c_uint32]
    L.ckmpm_hash_find.argtypes = [_vp, _u32p, ctypes.c_uint32, _u32p]
    L.ckmpm_hash_find_device.argtypes = [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_int32]
    L.ckmpm_hash_synchronize.argtypes = [_vp]
    L.ckmpm_hash_info.argtypes = [_vp, ctypes.POINTER(HashInfo)]
    L.ckmpm_hash_export.argtypes = [_vp, _u64p, ctypes.c_uint64, _u64p, ctypes.c_uint32, _u32p]
    L.ckmpm_hash_destroy.argtypes = [_vp]
    L.ckmpm_hash_build_noalloc.argtypes = [ctypes.POINTER(DomainDesc), ctypes.c_uint32, _vp, _vp, ctypes.c_uint32,
                                           _vp, _vp, _vp, _vp]
    L.ckmpm_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(_vp)]
    L.ckmpm_batch_run.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int32]
    L.ckmpm_batch_synchronize.argtypes = [_vp]
    L.ckmpm_batch_captures.argtypes = [_vp, ctypes.POINTER(ctypes.c_uint64)]
    L.ckmpm_batch_autotune.argtypes = [_vp, ctypes.c_uint32]
    L.ckmpm_batch_get_serial.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32)]
    L.ckmpm_batch_set_serial.argtypes = [_vp, ctypes.c_int32]
    L.ckmpm_batch_destroy.argtypes = [_vp]
    L.ckmpm_batch_last_error.argtypes = [_vp]
    L.ckmpm_batch_last_error.restype = ctypes.c_char_p
    L.ckmpm_adjoint_enable.argtypes = [_vp, ctypes.c_uint32]
    L.ckmpm_checkpoint_save.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.ckmpm_checkpoint_restore.argtypes = [_vp, _vp]
    L.ckmpm_checkpoint_free.argtypes = [_vp]
    L.ckmpm_adjoint_record.argtypes = [_vp, ctypes.c_int32]
    L.ckmpm_adjoint_reset.argtypes = [_vp]
    L.ckmpm_adjoint_seed.argtypes = [_vp, ctypes.c_int32, _f32p]
    L.ckmpm_adjoint_substep.argtypes = [_vp]
    L.ckmpm_adjoint_gradients.argtypes = [_vp, ctypes.c_int32, _f32p, _f32p, _f32p]
    L.ckmpm_get_positions_input_order.argtypes = [_vp, ctypes.c_int32, _f32p]
    L.ckmpm_adjoint_rewind.argtypes = [_vp, _f32p, ctypes.c_uint64, _f32p, _f32p]
    L.ckmpm_hash_last_error.argtypes = [_vp]
    L.ckmpm_hash_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if not name.endswith("last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


class CkmpmError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)
