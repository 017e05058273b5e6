"""ctypes binding of libfullmag_fdm.so.0 (include/fullmag_fdm.h, fullmag_fdm_ext.h).

The struct classes mirror the C layouts field for field, exactly as the Rust
binding crates/fullmag-fdm-sys/src/lib.rs:67-184 does, so the same plan
descriptor reaches the library from Python as from fullmag's runner.

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback behind this module.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfullmag_fdm.so.0")

FULLMAG_FDM_OK = 0
FULLMAG_FDM_ERR_INVALID = -1
FULLMAG_FDM_ERR_CUDA = -2
FULLMAG_FDM_ERR_INTERNAL = -3
MAX_EXCHANGE_REGIONS = 256

PRECISION_SINGLE = 1
PRECISION_DOUBLE = 2
PRECISION_MIXED = 3

INTEGRATOR_HEUN = 1
INTEGRATOR_DP45 = 2
INTEGRATOR_ABM3 = 3
INTEGRATOR_RK4 = 4
INTEGRATOR_RK23 = 5

BOUNDARY_NONE = 0
BOUNDARY_VOLUME = 1
BOUNDARY_FULL = 2

OBSERVABLE_M = 1
OBSERVABLE_H_EX = 2
OBSERVABLE_H_DEMAG = 3
OBSERVABLE_H_EXT = 4
OBSERVABLE_H_EFF = 5

SNAPSHOT_F32 = 1
SNAPSHOT_F64 = 2

_d = ctypes.c_double
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)
_pf = ctypes.POINTER(ctypes.c_float)


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", _u32), ("ny", _u32), ("nz", _u32), ("dx", _d), ("dy", _d), ("dz", _d)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("saturation_magnetisation", _d), ("exchange_stiffness", _d), ("damping", _d),
                ("gyromagnetic_ratio", _d)]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("grid", GridDesc),
        ("material", MaterialDesc),
        ("precision", _i32),
        ("integrator", _i32),
        ("disable_precession", _i32),
        ("enable_exchange", _i32),
        ("enable_demag", _i32),
        ("has_external_field", _i32),
        ("external_field_am", _d * 3),
        ("demag_kernel_xx_spectrum", _pd),
        ("demag_kernel_yy_spectrum", _pd),
        ("demag_kernel_zz_spectrum", _pd),
        ("demag_kernel_xy_spectrum", _pd),
        ("demag_kernel_xz_spectrum", _pd),
        ("demag_kernel_yz_spectrum", _pd),
        ("demag_kernel_spectrum_len", _u64),
        ("active_mask", ctypes.POINTER(ctypes.c_uint8)),
        ("active_mask_len", _u64),
        ("region_mask", ctypes.POINTER(ctypes.c_uint32)),
        ("region_mask_len", _u64),
        ("exchange_lut", _pd),
        ("exchange_lut_len", _u64),
        ("boundary_correction", _i32),
        ("boundary_phi_floor", _d),
        ("boundary_delta_min", _d),
        ("volume_fraction", _pd),
        ("volume_fraction_len", _u64),
        ("face_link_xp", _pd), ("face_link_xm", _pd),
        ("face_link_yp", _pd), ("face_link_ym", _pd),
        ("face_link_zp", _pd), ("face_link_zm", _pd),
        ("delta_xp", _pd), ("delta_xm", _pd),
        ("delta_yp", _pd), ("delta_ym", _pd),
        ("delta_zp", _pd), ("delta_zm", _pd),
        ("has_demag_boundary_corr", _i32),
        ("demag_corr_target_idx", ctypes.POINTER(ctypes.c_int32)),
        ("demag_corr_source_idx", ctypes.POINTER(ctypes.c_int32)),
        ("demag_corr_tensor", _pd),
        ("demag_corr_target_count", _u32),
        ("demag_corr_stencil_size", _u32),
        ("initial_magnetization_xyz", _pd),
        ("initial_magnetization_len", _u64),
        ("adaptive_max_error", _d),
        ("adaptive_dt_min", _d),
        ("adaptive_dt_max", _d),
        ("adaptive_headroom", _d),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("step", _u64), ("time_seconds", _d), ("dt_seconds", _d),
        ("exchange_energy_joules", _d), ("demag_energy_joules", _d), ("external_energy_joules", _d),
        ("total_energy_joules", _d), ("max_effective_field_amplitude", _d),
        ("max_demag_field_amplitude", _d), ("max_rhs_amplitude", _d), ("suggested_next_dt", _d),
        ("wall_time_ns", _u64),
    ]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("compute_capability_major", _i32),
                ("compute_capability_minor", _i32), ("driver_version", _i32), ("runtime_version", _i32)]


class SnapshotDesc(ctypes.Structure):
    _fields_ = [("cell_count", _u64), ("component_count", _u32), ("scalar_bytes", _u32), ("scalar_type", _i32)]


class ExtOptions(ctypes.Structure):
    _fields_ = [("struct_size", _u32), ("precision_override", ctypes.c_int32),
                ("newell_on_device", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("ms_cell", _pd), ("ms_cell_len", _u64)]


class StackDesc(ctypes.Structure):
    """fullmag_fdm_ext_stack_desc."""
    _fields_ = [("struct_size", _u32), ("layer_count", _u32),
                ("layers", ctypes.POINTER(ctypes.POINTER(PlanDesc))), ("layer_origins", _pd),
                ("common_cells", _u32 * 3), ("reserved", _u32), ("convolution_cell_size", _d * 3),
                ("enable_demag", ctypes.c_int32), ("precision_override", ctypes.c_int32)]


# Every exported entry point and its signature: (restype, argtypes).
SIGNATURES = {
    # fullmag_fdm.h
    "fullmag_fdm_is_available": (_i32, []),
    "fullmag_fdm_backend_create": (ctypes.c_void_p, [ctypes.POINTER(PlanDesc)]),
    "fullmag_fdm_backend_step": (_i32, [ctypes.c_void_p, _d, ctypes.POINTER(StepStats)]),
    "fullmag_fdm_backend_copy_field_f64": (_i32, [ctypes.c_void_p, _i32, _pd, _u64]),
    "fullmag_fdm_backend_copy_field_f32": (_i32, [ctypes.c_void_p, _i32, _pf, _u64]),
    "fullmag_fdm_backend_copy_field_preview_f64": (_i32, [ctypes.c_void_p, _i32, _u32, _u32, _u32, _u32, _u32,
                                                          _pd, _u64]),
    "fullmag_fdm_backend_copy_field_preview_f32": (_i32, [ctypes.c_void_p, _i32, _u32, _u32, _u32, _u32, _u32,
                                                          _pf, _u64]),
    "fullmag_fdm_backend_begin_field_snapshot": (ctypes.c_void_p, [ctypes.c_void_p, _i32]),
    "fullmag_fdm_field_snapshot_wait": (_i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                               ctypes.POINTER(_u64), ctypes.POINTER(SnapshotDesc)]),
    "fullmag_fdm_field_snapshot_destroy": (None, [ctypes.c_void_p]),
    "fullmag_fdm_backend_upload_magnetization_f64": (_i32, [ctypes.c_void_p, _pd, _u64]),
    "fullmag_fdm_backend_upload_magnetization_f32": (_i32, [ctypes.c_void_p, _pf, _u64]),
    "fullmag_fdm_backend_refresh_observables": (_i32, [ctypes.c_void_p]),
    "fullmag_fdm_backend_get_device_info": (_i32, [ctypes.c_void_p, ctypes.POINTER(DeviceInfo)]),
    "fullmag_fdm_backend_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "fullmag_fdm_backend_destroy": (None, [ctypes.c_void_p]),
    # fullmag_fdm_ext.h
    "fullmag_fdm_ext_backend_create": (ctypes.c_void_p, [ctypes.POINTER(PlanDesc), ctypes.POINTER(ExtOptions)]),
    "fullmag_fdm_ext_precision": (_i32, [ctypes.c_void_p]),
    "fullmag_fdm_ext_newell_spectra": (_i32, [_u32, _u32, _u32, _d, _d, _d, _i32, _pd, _u64]),
    "fullmag_fdm_ext_set_external_field": (_i32, [ctypes.c_void_p, _pd]),
    "fullmag_fdm_ext_step_many": (_i32, [ctypes.c_void_p, _d, _u64, ctypes.POINTER(StepStats)]),
    "fullmag_fdm_ext_set_timing": (_i32, [ctypes.c_void_p, _i32]),
    "fullmag_fdm_ext_get_timing": (_i32, [ctypes.c_void_p, _i32, ctypes.POINTER(ctypes.c_char_p), _pd,
                                          ctypes.POINTER(_u64)]),
    "fullmag_fdm_ext_random_unit_vectors": (None, [_u64, _u64, _pd]),
    "fullmag_fdm_ext_device_bytes": (_u64, [ctypes.c_void_p]),
    "fullmag_fdm_ext_heun_adjoint": (_i32, [ctypes.c_void_p, _d, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, _pd]),
    "fullmag_fdm_ext_copy_state_device": (_i32, [ctypes.c_void_p, ctypes.c_void_p, _u64]),
    "fullmag_fdm_ext_batch_create": (ctypes.c_void_p, [ctypes.POINTER(ctypes.c_void_p), _u32, _pd, _u32]),
    "fullmag_fdm_ext_batch_run": (_i32, [ctypes.c_void_p, _u64, ctypes.POINTER(StepStats)]),
    "fullmag_fdm_ext_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "fullmag_fdm_ext_batch_destroy": (None, [ctypes.c_void_p]),
    "fullmag_fdm_ext_autotune": (_i32, [ctypes.c_void_p, _i32, ctypes.c_char_p, _u64]),
    "fullmag_fdm_ext_set_schedule": (_i32, [ctypes.c_void_p, ctypes.c_char_p]),
    "fullmag_fdm_ext_set_state_device": (_i32, [ctypes.c_void_p, ctypes.c_void_p, _u64]),
    "fullmag_fdm_ext_copy_field_preview_noalloc": (_i32, [ctypes.c_void_p, _i32, _u32, _u32, _u32, _u32, _u32,
                                                          _i32, ctypes.c_void_p, _u64]),
    "fullmag_fdm_ext_stack_create": (ctypes.c_void_p, [ctypes.POINTER(StackDesc)]),
    "fullmag_fdm_ext_stack_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "fullmag_fdm_ext_stack_step": (_i32, [ctypes.c_void_p, _d, _u64, ctypes.POINTER(StepStats)]),
    "fullmag_fdm_ext_stack_refresh": (_i32, [ctypes.c_void_p]),
    "fullmag_fdm_ext_stack_layer": (ctypes.c_void_p, [ctypes.c_void_p, _u32]),
    "fullmag_fdm_ext_stack_destroy": (None, [ctypes.c_void_p]),
}

# fullmag_fdm_ext.h: FULLMAG_FDM_EXT_NOALLOC_SCRATCH_BYTES
NOALLOC_SCRATCH_BYTES = 4194304

_lib = None


class LibraryNotBuilt(RuntimeError):
    pass


def lib():
    """Load the HIP library (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing: build the HIP extension first (python -c "
                f"'import __graft_entry__ as g; g.build()')")
        # torch wheels bundle their own HIP runtime under the same SONAME
        # (libamdhip64.so.7); whichever loads first serves the whole process.
        # Load torch's first when torch is installed, so that tensors and this
        # library share one runtime (device pointers pass between them).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
