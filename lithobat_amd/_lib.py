"""ctypes binding of the C ABI in include/lithobat.h.

The library is the in-tree ``liblithobat_amd.so`` built by
``lithobat_amd.build``.  torch is imported first so the library binds to the
HIP runtime torch already loaded (same soname), which makes torch's device
pointers and streams valid arguments.  There is no fallback: a missing or
unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (must load the HIP runtime before the library)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblithobat_amd.so")

LB_OK, LB_ERR_ARG, LB_ERR_UNSUPPORTED, LB_ERR_HIP, LB_ERR_OOM, LB_ERR_WORKSPACE = range(6)
POLARISATION_CODES = {"unpolarised": 0, "unpolarized": 0, "x": 1, "y": 2, "te": 3, "tm": 4}


class ImagingDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("pixel_size", C.c_double), ("wavelength", C.c_double),
        ("na", C.c_double), ("vector", C.c_int32), ("polarisation", C.c_int32),
        ("obliquity", C.c_int32), ("exact_defocus", C.c_int32), ("n_planes", C.c_int32),
        ("defocus", C.POINTER(C.c_double)), ("n_image", C.POINTER(C.c_double)),
        ("n_zernike", C.c_int32), ("zernike_j", C.POINTER(C.c_int32)),
        ("zernike_c", C.POINTER(C.c_double)), ("n_source", C.c_int32),
        ("src_weight", C.POINTER(C.c_double)), ("src_fx", C.POINTER(C.c_double)),
        ("src_fy", C.POINTER(C.c_double)), ("src_phi", C.POINTER(C.c_double)),
        ("precision", C.c_int32), ("coarse_m", C.c_int32), ("chunk", C.c_int32),
        ("n_refl", C.c_int32), ("refl_theta", C.POINTER(C.c_double)),
        ("refl_mag", C.POINTER(C.c_double)), ("refl_phase", C.POINTER(C.c_double)),
        ("refl_doff", C.c_double), ("src_sin_x", C.POINTER(C.c_double)),
        ("src_sin_y", C.POINTER(C.c_double)),
    ]


class PlanInfo(C.Structure):
    _fields_ = [(k, C.c_int32) for k in (
        "n", "m", "box", "band", "n_planes", "n_systems", "precision", "n_kernels",
        "block_threads", "column_splits", "chunk", "grid_blocks", "adj_block_threads",
        "adj_column_splits", "adj_chunk", "socs_sweeps")]


# LB_STATIC_SCRATCH_BYTES of the header: lb_forward_static's fixed device workspace
STATIC_SCRATCH_BYTES = 64 << 20

EXPORTS = [
    "lb_version", "lb_last_error", "lb_abbe_plan_create", "lb_socs_plan_create",
    "lb_plan_destroy", "lb_plan_info_get", "lb_plan_clear", "lb_socs_weights",
    "lb_socs_kernels", "lb_workspace_bytes", "lb_forward", "lb_forward_static", "lb_backward",
    "lb_fields", "lb_fields_backward", "lb_fields_backward_workspace_bytes",
    "lb_backward_spectrum_complex", "lb_pupil_dense",
    "lb_fields_workspace_bytes", "lb_profile_enable", "lb_profile_read",
    "lb_profile_read_adjoint", "lb_plan_shapes", "lb_plan_work_items",
    "lb_plan_set_shape", "lb_batch_create", "lb_batch_destroy", "lb_batch_workspace_bytes",
    "lb_batch_forward", "lb_batch_backward", "lb_dense_workspace_bytes", "lb_forward_dense",
    "lb_backward_dense", "lb_gaussian_blur_workspace_bytes", "lb_gaussian_blur",
    "lb_resist_objective_workspace_bytes", "lb_resist_objective", "lb_mask_sigmoid",
    "lb_mask_sigmoid_grad", "lb_mask_penalties", "lb_adam_step", "lb_adam_step_dev",
    "lb_ilt_schedule", "lb_mask_sigmoid_dev", "lb_mask_sigmoid_grad_dev", "lb_batch_set_shape",
    "lb_volume_interp", "lb_mask_spectrum_bytes", "lb_mask_spectrum", "lb_forward_spectrum",
    "lb_backward_spectrum", "lb_batch_backward_reuse", "lb_mask_spectrum_sigmoid",
    "lb_mask_sigmoid_grad_theta", "lb_forward_objective_workspace_bytes",
    "lb_sensitivity_band_bytes", "lb_forward_objective", "lb_backward_sensitivity_band",
    "lb_backward_objective_theta",
]

_lib = None


def load() -> C.CDLL:
    """Load the library once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m lithobat_amd.build` "
            "(the imaging path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    vp, i32, sz = C.c_void_p, C.c_int32, C.c_size_t
    plan_p = C.POINTER(vp)
    sig = {
        "lb_version": ([], i32),
        "lb_last_error": ([], C.c_char_p),
        "lb_abbe_plan_create": ([C.POINTER(ImagingDesc), plan_p], C.c_int),
        "lb_socs_plan_create": ([C.POINTER(ImagingDesc), C.c_double, i32, plan_p], C.c_int),
        "lb_plan_destroy": ([vp], C.c_int),
        "lb_pupil_dense": ([C.POINTER(ImagingDesc), C.c_double, C.c_double, i32, vp, vp],
                           C.c_int),
        "lb_plan_info_get": ([vp, C.POINTER(PlanInfo)], C.c_int),
        "lb_plan_clear": ([vp, C.POINTER(C.c_double)], C.c_int),
        "lb_socs_weights": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                             C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "lb_socs_kernels": ([vp, vp, vp], C.c_int),
        "lb_workspace_bytes": ([vp, i32, i32], sz),
        "lb_forward": ([vp, vp, i32, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_forward_static": ([vp, vp, i32, i32, vp, vp, vp], C.c_int),
        "lb_backward": ([vp, vp, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_fields": ([vp, vp, i32, i32, vp, vp, sz, vp], C.c_int),
        "lb_fields_workspace_bytes": ([vp, i32], sz),
        "lb_fields_backward_workspace_bytes": ([vp], sz),
        "lb_fields_backward": ([vp, vp, i32, vp, vp, i32, vp, sz, vp], C.c_int),
        "lb_backward_spectrum_complex": ([vp, vp, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_profile_enable": ([vp, i32], C.c_int),
        "lb_profile_read": ([vp, C.POINTER(i32), C.POINTER(C.c_double)], C.c_int),
        "lb_profile_read_adjoint": ([vp, C.POINTER(i32), C.POINTER(C.c_double)], C.c_int),
        "lb_plan_shapes": ([vp, i32, C.POINTER(i32), i32, C.POINTER(i32)], C.c_int),
        "lb_plan_set_shape": ([vp, i32, i32, i32, i32], C.c_int),
        "lb_plan_work_items": ([vp, i32, i32, C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "lb_dense_workspace_bytes": ([vp, i32], sz),
        "lb_forward_dense": ([vp, vp, i32, i32, vp, vp, sz, vp], C.c_int),
        "lb_backward_dense": ([vp, vp, i32, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_gaussian_blur_workspace_bytes": ([i32, i32, C.c_double, i32], sz),
        "lb_gaussian_blur": ([vp, vp, i32, i32, C.c_double, i32, vp, sz, vp], C.c_int),
        "lb_resist_objective_workspace_bytes": ([i32, C.c_int64], sz),
        "lb_resist_objective": ([vp, vp, i32, C.c_int64, C.c_double, C.c_double, C.c_double,
                                 C.c_double, vp, vp, vp, vp, vp, i32, vp, sz, vp], C.c_int),
        "lb_mask_sigmoid": ([vp, C.c_int64, C.c_double, vp, i32, vp], C.c_int),
        "lb_mask_sigmoid_grad": ([vp, vp, C.c_int64, C.c_double, vp, i32, vp], C.c_int),
        "lb_mask_penalties": ([vp, i32, i32, i32, C.c_double, C.c_double, vp, vp, vp, i32, vp, sz,
                               vp], C.c_int),
        "lb_adam_step": ([vp, vp, vp, vp, C.c_int64, i32, C.c_double, C.c_double, C.c_double,
                          C.c_double, i32, vp], C.c_int),
        "lb_adam_step_dev": ([vp, vp, vp, vp, C.c_int64, vp, C.c_double, C.c_double, C.c_double,
                              C.c_double, i32, vp], C.c_int),
        "lb_ilt_schedule": ([vp, C.c_double, C.c_double, C.c_double, vp], C.c_int),
        "lb_mask_sigmoid_dev": ([vp, C.c_int64, vp, vp, i32, vp], C.c_int),
        "lb_mask_sigmoid_grad_dev": ([vp, vp, C.c_int64, vp, vp, i32, vp], C.c_int),
        "lb_batch_create": ([C.POINTER(vp), i32, plan_p], C.c_int),
        "lb_batch_destroy": ([vp], C.c_int),
        "lb_batch_set_shape": ([vp, i32, i32, i32], C.c_int),
        "lb_volume_interp": ([vp, i32, i32, C.c_int64, i32, vp, vp, vp, vp, i32, vp], C.c_int),
        "lb_mask_spectrum_bytes": ([vp, i32], sz),
        "lb_mask_spectrum": ([vp, vp, i32, i32, vp, vp, sz, vp], C.c_int),
        "lb_forward_spectrum": ([vp, vp, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_backward_spectrum": ([vp, vp, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_mask_spectrum_sigmoid": ([vp, vp, i32, C.c_double, vp, vp, sz, vp], C.c_int),
        "lb_mask_sigmoid_grad_theta": ([vp, vp, C.c_int64, C.c_double, vp, vp, i32, vp],
                                       C.c_int),
        "lb_forward_objective_workspace_bytes": ([vp, i32], sz),
        "lb_sensitivity_band_bytes": ([vp, i32], sz),
        "lb_forward_objective": ([vp, vp, i32, vp, C.c_double, C.c_double, C.c_double,
                                  C.c_double, vp, vp, vp, vp, vp, vp, vp, sz, vp], C.c_int),
        "lb_backward_sensitivity_band": ([vp, vp, i32, vp, vp, vp, sz, vp], C.c_int),
        "lb_backward_objective_theta": ([vp, vp, i32, vp, vp, C.c_double, vp, vp, vp, sz, vp],
                                        C.c_int),
        "lb_batch_backward_reuse": ([vp, C.POINTER(vp), C.POINTER(vp), vp, sz, vp, vp], C.c_int),
        "lb_batch_workspace_bytes": ([vp, i32], sz),
        "lb_batch_forward": ([vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), vp, sz, vp],
                             C.c_int),
        "lb_batch_backward": ([vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), vp, sz, vp],
                              C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class LithobatError(RuntimeError):
    """A C-ABI call failed (status code and the library's message)."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def check(status: int) -> None:
    if status != LB_OK:
        msg = load().lb_last_error().decode(errors="replace")
        if status == LB_ERR_ARG:
            raise ValueError(msg)
        if status == LB_ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        raise LithobatError(status, msg)
