"""ctypes mirror of include/gargantua.h (ABI version 1).

The structs here are byte-for-byte the C structs of the header; the loader
finds libgargantua.so next to this file.  Nothing here computes anything.
"""
from __future__ import annotations

import ctypes as C
import os

ABI_VERSION = 1

GG_OK = 0
GG_ERR_INVALID_ARG = -1
GG_ERR_UNSUPPORTED = -5
GG_METRIC_SCHWARZSCHILD = 0
GG_METRIC_KERR = 1
GG_METRIC_USER = 2
GG_DISK_NONE = 0
GG_DISK_VOLUMETRIC = 1
GG_DISK_THIN = 2
GG_ERR_REFERENCE = 0
GG_ERR_RTOL_ATOL = 1
GG_PREC_FP64 = 0
GG_PREC_MIXED = 1
GG_PREC_FP64_FAST = 2
GG_CLASS_HORIZON = 0
GG_CLASS_ESCAPED = 1
GG_CLASS_DISK = 2
GG_CLASS_MAX_STEPS = 3
GG_CLASS_SCENE = 4
GG_SHAPE_SPHERE = 0
GG_SHAPE_QUAD = 1
GG_MATERIAL_LAMBERTIAN = 0
GG_MATERIAL_METAL = 1
GG_MATERIAL_DIELECTRIC = 2
GG_MATERIAL_DIFFUSE_LIGHT = 3


class gg_metric(C.Structure):
    _fields_ = [("kind", C.c_int32), ("_pad", C.c_int32), ("mass", C.c_double), ("spin", C.c_double)]


class gg_camera(C.Structure):
    _fields_ = [("position", C.c_double * 4), ("look_at", C.c_double * 3), ("up", C.c_double * 3),
                ("fov_y_radians", C.c_double), ("aspect", C.c_double)]


class gg_disk(C.Structure):
    _fields_ = [("kind", C.c_int32), ("_pad", C.c_int32), ("r_inner", C.c_double), ("r_outer", C.c_double),
                ("scale_height", C.c_double), ("density_0", C.c_double), ("temperature_0", C.c_double),
                ("absorption", C.c_double), ("emissivity_q", C.c_double), ("edge_width", C.c_double)]


class gg_integrator(C.Structure):
    _fields_ = [("initial_step", C.c_double), ("min_step", C.c_double), ("max_step", C.c_double),
                ("tolerance", C.c_double), ("max_radius", C.c_double), ("horizon_buffer", C.c_double),
                ("horizon_rel", C.c_double), ("rtol", C.c_double), ("atol", C.c_double),
                ("err_mode", C.c_int32), ("max_steps", C.c_uint32)]


class gg_image(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples_per_axis", C.c_uint32),
                ("jitter", C.c_uint32), ("seed", C.c_uint64)]


class gg_scene(C.Structure):
    _fields_ = [("metric", gg_metric), ("camera", gg_camera), ("disk", gg_disk),
                ("integrator", gg_integrator), ("image", gg_image)]


class gg_outputs(C.Structure):
    _fields_ = [("cls", C.c_void_p), ("r", C.c_void_p), ("theta", C.c_void_p), ("phi", C.c_void_p),
                ("intensity", C.c_void_p), ("attempts", C.c_void_p), ("accepted", C.c_void_p),
                ("state", C.c_void_p), ("rgb", C.c_void_p), ("color", C.c_void_p)]


class gg_hittable(C.Structure):
    _fields_ = [("shape", C.c_int32), ("material", C.c_int32), ("q", C.c_double * 3), ("u", C.c_double * 3),
                ("v", C.c_double * 3), ("radius", C.c_double), ("color", C.c_double * 3), ("fuzz", C.c_double),
                ("ior", C.c_double)]


class gg_stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("rays", C.c_uint64),
                ("attempts", C.c_uint64), ("rhs_evals", C.c_uint64), ("rhs_evals_fp32", C.c_uint64),
                ("rhs_evals_fast", C.c_uint64),
                ("attempts_full", C.c_uint64), ("reflections", C.c_uint64), ("grid_blocks", C.c_uint32),
                ("block_threads", C.c_uint32)]


class gg_launch(C.Structure):
    _fields_ = [("block_threads", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("refill_interval", C.c_uint32), ("graph", C.c_uint32), ("order", C.c_uint32)]


class gg_grad_scene(C.Structure):
    _fields_ = [("mass", C.c_double), ("spin", C.c_double), ("inclination", C.c_double),
                ("emissivity_q", C.c_double), ("r_inner", C.c_double), ("r_outer", C.c_double),
                ("edge_width", C.c_double), ("r_camera", C.c_double), ("fov_y_radians", C.c_double),
                ("integrator", gg_integrator), ("width", C.c_uint32), ("height", C.c_uint32),
                ("precision", C.c_int32), ("frozen_steps", C.c_uint32)]


class gg_grad_stats(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("adjoint_ms", C.c_double), ("total_ms", C.c_double),
                ("rays", C.c_uint64), ("disk_rays", C.c_uint64), ("steps", C.c_uint64), ("attempts", C.c_uint64),
                ("checkpoints", C.c_uint64), ("checkpoint_bytes", C.c_uint64), ("segment_bytes", C.c_uint64), ("record_bytes", C.c_uint64),
                ("grid_blocks", C.c_uint32), ("block_threads", C.c_uint32),
                ("checkpoint_every", C.c_uint32), ("mode", C.c_uint32)]


# Entry points declared in include/gargantua.h: (name, restype, argtypes).
_P = C.c_void_p
EXPORTS = [
    ("gg_strerror", C.c_char_p, [C.c_int]),
    ("gg_last_error", C.c_char_p, []),
    ("gg_abi_version", C.c_int, []),
    ("gg_context_create", C.c_int, [C.c_int, _P, C.POINTER(_P)]),
    ("gg_context_destroy", C.c_int, [_P]),
    ("gg_set_launch", C.c_int, [_P, C.POINTER(gg_launch)]),
    ("gg_set_grad_memory", C.c_int, [_P, C.c_uint64]),
    ("gg_set_sky_hdri", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_float]),
    ("gg_set_hittables", C.c_int, [_P, C.c_uint32, _P]),
    ("gg_set_user_metric", C.c_int, [_P, C.c_char_p, C.c_double]),
    ("gg_check_user_metric", C.c_int, [C.c_char_p, C.c_int32, C.c_char_p, C.c_size_t]),
    ("gg_get_launch", C.c_int, [_P, C.POINTER(gg_launch)]),
    ("gg_scene_rays", C.c_int, [C.POINTER(gg_scene), C.POINTER(C.c_uint64)]),
    ("gg_render", C.c_int, [_P, C.POINTER(gg_scene), C.c_int, C.POINTER(gg_outputs), C.POINTER(gg_stats)]),
    ("gg_render_batch", C.c_int, [_P, C.POINTER(gg_scene), C.c_uint32, C.c_int, C.POINTER(gg_outputs),
                                  C.POINTER(gg_stats)]),
    ("gg_render_batch_noalloc", C.c_int, [_P, C.POINTER(gg_scene), C.c_uint32, C.c_int, C.POINTER(gg_outputs),
                                          C.POINTER(gg_stats)]),
    ("gg_trace_states", C.c_int, [_P, C.POINTER(gg_scene), C.c_int, C.c_uint64, _P, C.POINTER(gg_outputs),
                                  C.POINTER(gg_stats)]),
    ("gg_geodesic_acceleration", C.c_int, [_P, C.POINTER(gg_metric), C.c_int, C.c_uint64, _P, _P, _P]),
    ("gg_autotune", C.c_int, [_P, C.POINTER(gg_scene), C.c_uint32, C.c_int, C.POINTER(gg_launch), C.c_uint32,
                              C.POINTER(gg_launch), C.POINTER(C.c_double)]),
    ("gg_camera_rays", C.c_int, [_P, C.POINTER(gg_scene), _P]),
    ("gg_constants_of_motion", C.c_int, [_P, C.POINTER(gg_metric), C.c_uint64, _P, _P]),
    ("gg_grad_scene_to_scene", C.c_int, [C.POINTER(gg_grad_scene), C.POINTER(gg_scene)]),
    ("gg_grad", C.c_int, [_P, C.POINTER(gg_grad_scene), _P, C.c_uint32, C.POINTER(C.c_double), _P, _P, _P,
                          C.POINTER(gg_grad_stats)]),
    ("gg_geodesic_acceleration_vjp", C.c_int, [_P, C.POINTER(gg_metric), C.c_uint64, _P, _P, _P, _P, _P, _P]),
]

LIB_NAME = "libgargantua.so"
LIB_PATH = os.environ.get("GARGANTUA_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

_lib = None


def load() -> C.CDLL:
    """Load libgargantua.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP library is the only implementation; there is no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.gg_abi_version() != ABI_VERSION:
        raise RuntimeError(f"libgargantua ABI {lib.gg_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib
