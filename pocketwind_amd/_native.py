"""ctypes binding of libpocketwind.so (the C-ABI declared in include/pw.h).

The library must exist: nothing here falls back to a CPU path. Build it with
``python -m pocketwind_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

from . import build as _build

LIB_PATH = _build.LIB

PW_OK, PW_EINVAL, PW_ENOMEM, PW_EBUSY, PW_EDEVICE, PW_ESTATE = range(6)


class Layout(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32)]


class Params(C.Structure):
    _fields_ = [("tau0", C.c_double), ("u_in", C.c_double), ("les", C.c_uint32),
                ("op", C.c_uint32), ("walls", C.c_uint32), ("storage", C.c_uint32),
                ("ramp_steps", C.c_uint32), ("wall_model", C.c_uint32), ("sides", C.c_uint32),
                ("outlet", C.c_uint32), ("slip_floor", C.c_float)]


class ForceSample(C.Structure):
    _fields_ = [("step", C.c_uint64), ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double)]


class DeviceState(C.Structure):
    _fields_ = [("populations", C.c_void_p), ("macro", C.c_void_p), ("geometry", C.c_void_p),
                ("stream", C.c_void_p), ("run_elements", C.c_size_t)]


_u32p = C.POINTER(C.c_uint32)


class VoxelGridC(C.Structure):
    _fields_ = [("geometry", C.POINTER(C.c_uint32)), ("geometry_words", C.c_size_t),
                ("boundary_cells", C.POINTER(C.c_uint32)), ("n_boundary", C.c_size_t),
                ("solid_cells", C.c_uint64), ("frontal_cells", C.c_uint64),
                ("blockage", C.c_double), ("watertight", C.c_int32), ("missed_links", C.c_uint64)]
_lib = None

# Every entry point of include/pw.h with its ctypes signature (restype, argtypes).
SIGNATURES = {
    "pw_last_error_message": (C.c_char_p, []),
    "pw_abi_version": (C.c_int, []),
    "pw_device_count": (C.c_int, []),
    "pw_solver_create": (C.c_int, [C.POINTER(Layout), _u32p, C.c_size_t, _u32p, C.c_size_t,
                                   C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "pw_pack_params": (C.c_int, [C.POINTER(Layout), C.POINTER(Params), C.c_int, C.c_uint64, C.c_void_p,
                                 C.c_size_t]),
    "pw_solver_step": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int]),
    "pw_solver_encode_force": (C.c_int, [C.c_void_p]),
    "pw_solver_read_force": (C.c_int, [C.c_void_p, C.POINTER(ForceSample)]),
    "pw_solver_reset": (C.c_int, [C.c_void_p]),
    "pw_solver_step_count": (C.c_uint64, [C.c_void_p]),
    "pw_solver_destroy": (None, [C.c_void_p]),
    "pw_solver_read_macro": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_read_populations": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_write_populations": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_population_bytes": (C.c_size_t, [C.c_void_p]),
    "pw_debug_inject_nan": (C.c_int, [C.c_void_p]),
    "pw_solver_device_state": (C.c_int, [C.c_void_p, C.POINTER(DeviceState)]),
    "pw_solver_set_kernel": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pw_solver_advance_guarded": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64)]),
    "pw_solver_field_vtk_blocks": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_size_t, C.c_void_p,
                                             C.c_size_t]),
    "pw_solver_block_size": (C.c_uint32, [C.c_void_p]),
    "pw_debug_cell_path": (C.c_int, [C.c_void_p, C.c_int]),
    "pw_solver_set_rounding": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pw_debug_set_wall_stress": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_set_closure_placement": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pw_batch_set_closure_placement": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pw_solver_set_step_variant": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "pw_solver_xcd_chunks": (C.c_uint32, [C.c_void_p]),
    "pw_batch_persistent_grid": (C.c_uint32, [C.c_void_p]),
    "pw_batch_block_size": (C.c_uint32, [C.c_void_p]),
    "pw_solver_kernel": (C.c_uint32, [C.c_void_p]),
    "pw_solver_synchronize": (C.c_int, [C.c_void_p]),
    "pw_solver_step_sampled": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t]),
    "pw_solver_step_sampled_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                                 C.c_size_t]),
    "pw_solver_time_steps": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]),
    "pw_copy_bandwidth": (C.c_int, [C.c_size_t, C.c_uint32, C.POINTER(C.c_double)]),
    "pw_voxelize": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.POINTER(Layout),
                              C.c_int, C.c_int, C.POINTER(VoxelGridC)]),
    "pw_voxel_grid_free": (None, [C.POINTER(VoxelGridC)]),
    "pw_particles_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_void_p)]),
    "pw_particles_update": (C.c_int, [C.c_void_p, C.c_float]),
    "pw_particles_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_particles_write": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_particles_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "pw_particles_destroy": (None, [C.c_void_p]),
    "pw_particles_create_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_void_p)]),
    "pw_particles_create_field": (C.c_int, [C.POINTER(Layout), _u32p, C.c_size_t, C.c_void_p, C.c_uint32,
                                            C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_void_p)]),
    "pw_particles_synchronize": (C.c_int, [C.c_void_p]),
    "pw_batch_create": (C.c_int, [C.c_uint32, C.POINTER(Layout), C.POINTER(C.c_void_p),
                                  C.POINTER(C.c_size_t), C.POINTER(C.c_void_p), C.POINTER(C.c_size_t),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(Params),
                                  C.POINTER(C.c_void_p)]),
    "pw_batch_step": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int]),
    "pw_batch_step_variant": (C.c_uint32, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "pw_batch_step_sampled": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t]),
    "pw_batch_read_macro": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t]),
    "pw_batch_read_populations": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t]),
    "pw_batch_step_count": (C.c_uint64, [C.c_void_p]),
    "pw_batch_cells": (C.c_uint64, [C.c_void_p]),
    "pw_batch_time_steps": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]),
    "pw_batch_destroy": (None, [C.c_void_p]),
    "pw_batch_set_closure": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_batch_set_closure_interval": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pw_batch_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_wall_cell_count": (C.c_size_t, [C.c_void_p]),
    "pw_solver_wall_cells": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_closure_features": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_size_t]),
    "pw_solver_wall_forces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_set_closure": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_closure_output": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_solver_set_closure_interval": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pw_mlp_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "pw_mlp_destroy": (None, [C.c_void_p]),
    "pw_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "pw_mlp_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                   C.POINTER(C.c_float)]),
    "pw_mlp_set_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "pw_mlp_train": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_void_p]),
}


class _Library:
    """libpocketwind.so with every entry point typed on first use."""

    def __init__(self, handle: C.CDLL):
        self._handle = handle

    def __getattr__(self, name):
        fn = getattr(self._handle, name)
        if name in SIGNATURES:
            fn.restype, fn.argtypes = SIGNATURES[name]
        setattr(self, name, fn)
        return fn


def lib() -> "_Library":
    """Load libpocketwind.so or raise: the product has no fallback path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing; build it with `python -m pocketwind_amd.build`.")
        _lib = _Library(C.CDLL(LIB_PATH))
    return _lib


def last_error() -> str:
    msg = lib().pw_last_error_message()
    return msg.decode() if msg else ""


class DeviceError(RuntimeError):
    """A HIP runtime failure inside libpocketwind."""


def check(status: int) -> int:
    """Map a pw_status onto the reference's error kinds (RangeError -> ValueError)."""
    if status in (PW_OK, PW_EBUSY):
        return status
    message = last_error()
    if status == PW_EINVAL:
        raise ValueError(message)
    if status == PW_ENOMEM:
        raise MemoryError(message)
    if status == PW_ESTATE:
        raise RuntimeError(message)
    raise DeviceError(message)


def device_count() -> int:
    return int(lib().pw_device_count())
