"""ctypes binding of the C ABI declared in include/path_amd.h.

The library is loaded from path_amd/_lib/libpath_amd.so (built in-tree by
`python -m path_amd.build`). There is no fallback: if the library is missing,
importing the renderer raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libpath_amd.so"

PT_OK = 0
PT_ERR_INVALID = 1
PT_ERR_IO = 2
PT_ERR_HIP = 3
PT_ERR_NO_DEVICE = 4
PT_ERR_UNSUPPORTED = 5

# enums (values match include/path_amd.h)
BSDF = {"Diffuse": 0, "Glossy": 1, "DiffuseGlossy": 2, "DiffuseMirror": 3,
        "Mirror": 4, "DielectricRefractive": 5}
KERNEL_APPROACH = {"MegaKernel": 0, "Streaming": 1}
ACCEL = {"LoopAll": 0, "NaivePartitionBVH": 1, "SBVH": 2, "DirectionGrid": 3}
LIGHT_SAMPLER = {"NoLightSampling": 0, "RandomTriangle": 1}
DIR_SAMPLER = {"Uniform": 0, "BSDF": 1}
TERM_PROB = {"Constant": 0, "NIters": 1, "MultiplierFunc": 2, "Normalize": 3}
RNG = {"Uniform": 0, "Sobel": 1}
SAMPLE_SPEC = {"SquareImage": 0, "InitialRays": 1, "InitialIdxAndDir": 2}
OUTPUT = {"BGRA": 0, "FloatRGB": 1, "OutputPerStep": 2, "DeviceFloatRGB": 3}


class PtMaterial(C.Structure):
    _fields_ = [("type", C.c_uint32), ("diffuse", C.c_float * 3),
                ("specular", C.c_float * 3), ("shininess", C.c_float),
                ("ior", C.c_float), ("diffuse_weight", C.c_float),
                ("glossy_weight", C.c_float), ("emission", C.c_float * 3)]


class PtWorkDivisionSettings(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("target_x_block_size", C.c_uint32),
                ("force_target_samples", C.c_uint32),
                ("forced_target_samples_per_thread", C.c_uint32),
                ("base_num_threads", C.c_uint32),
                ("samples_per_thread_scaling_power", C.c_float),
                ("max_samples_per_thread", C.c_uint32)]


class PtSettings(C.Structure):
    _fields_ = [
        ("kernel_approach", C.c_uint32),
        ("max_blocks_per_launch_gpu", C.c_uint32),
        ("max_blocks_per_launch_cpu", C.c_uint32),
        ("render_work_division", PtWorkDivisionSettings),
        ("reduce_work_division", PtWorkDivisionSettings),
        ("init_samples_division", PtWorkDivisionSettings),
        ("max_num_samples_per_launch", C.c_uint32),
        ("min_num_init_blocks", C.c_uint32),
        ("to_bulk_max_size", C.c_uint32),
        ("accel", C.c_uint32),
        ("overlap_threshold", C.c_float),
        ("use_spatial_splits", C.c_uint32),
        ("traversal_per_intersect_cost", C.c_float),
        ("target_objects", C.c_uint32),
        ("print_stats", C.c_uint32),
        ("light_sampler", C.c_uint32),
        ("binary_search_threshold", C.c_uint32),
        ("dir_sampler", C.c_uint32),
        ("term_prob", C.c_uint32),
        ("constant_prob", C.c_float),
        ("n_iters", C.c_uint32),
        ("multiplier_func_exp", C.c_float),
        ("multiplier_func_min_prob", C.c_float),
        ("normalize_min_prob", C.c_float),
        ("rng", C.c_uint32),
        ("max_sample_size", C.c_uint32),
        ("back_cull_emission", C.c_uint32),
    ]


class PtSampleSpec(C.Structure):
    _fields_ = [("type", C.c_uint32), ("x_dim", C.c_uint32), ("y_dim", C.c_uint32),
                ("film_to_world", C.c_float * 12), ("rays", C.POINTER(C.c_float)),
                ("initial_idx", C.POINTER(C.c_uint32)), ("n_rays", C.c_uint32)]


class PtRayBatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("rays", C.POINTER(C.c_float)),
                ("initial_idx", C.POINTER(C.c_uint32)), ("n_rays", C.c_uint32),
                ("out", C.POINTER(C.c_float))]


class PtOutput(C.Structure):
    _fields_ = [("type", C.c_uint32), ("data", C.c_void_p),
                ("per_step", C.POINTER(C.POINTER(C.c_float))), ("n_steps", C.c_uint32)]


class PtRenderStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("trace_ms", C.c_double),
                ("rays", C.c_uint64), ("samples", C.c_uint64),
                ("trace_launches", C.c_uint64), ("busy_trace_launches", C.c_uint64),
                ("ext_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("mega_ms", C.c_double),
                ("mega_launches", C.c_uint64), ("profiled", C.c_uint32),
                ("pipelines", C.c_uint32), ("replays", C.c_uint32)]


class PtBatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("x_dim", C.c_uint32), ("y_dim", C.c_uint32),
                ("film_to_world", C.c_float * 12), ("out_rgb", C.POINTER(C.c_float)),
                ("out_gbuffer", C.POINTER(C.c_float))]


# Every exported symbol with its (restype, argtypes).
_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
SIGNATURES = {
    "pt_last_error": (C.c_char_p, []),
    "pt_scene_load": (C.c_int, [C.c_char_p, C.c_float, C.c_int, C.POINTER(_P), _F]),
    "pt_scene_create": (C.c_int, [C.POINTER(_P)]),
    "pt_scene_add_mesh": (C.c_int, [_P, _F, _F, _U, C.c_uint32, C.POINTER(PtMaterial), C.c_uint32, _F]),
    "pt_scene_add_obj": (C.c_int, [_P, C.c_char_p, _F]),
    "pt_scene_destroy": (None, [_P]),
    "pt_scene_counts": (C.c_int, [_P, _U, _U, _U]),
    "pt_scene_flat": (C.c_int, [_P, _F, _F, _U, C.POINTER(PtMaterial), _F, _F]),
    "pt_camera_from_look": (C.c_int, [_F, _F, _F, C.c_float, C.c_float, _F]),
    "pt_settings_default": (None, [C.POINTER(PtSettings)]),
    "pt_settings_load_yaml": (C.c_int, [C.c_char_p, C.POINTER(PtSettings)]),
    "pt_settings_to_yaml": (C.c_int, [C.POINTER(PtSettings), C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pt_renderer_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "pt_renderer_destroy": (None, [_P]),
    "pt_render_rays_batch": (C.c_int, [_P, C.POINTER(PtRayBatchItem), C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.POINTER(PtSettings),
                                       C.POINTER(C.c_double)]),
    "pt_render_noalloc": (C.c_int, [_P, C.POINTER(PtSampleSpec), _P, _P, C.c_uint32,
                                    C.POINTER(PtSettings), C.POINTER(C.c_double)]),
    "pt_render": (C.c_int, [_P, C.POINTER(PtSampleSpec), C.POINTER(PtOutput), _P, C.c_uint32,
                            C.POINTER(PtSettings), C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "pt_intersect": (C.c_int, [_P, _P, _F, C.c_uint32, _F, _U, _U]),
    "pt_sobol_table": (C.c_int, [_P, C.c_uint32, C.c_uint32, _F]),
    "pt_sobol_direction_numbers": (C.c_int, [C.c_uint32, _U]),
    "pt_work_division": (C.c_int, [C.POINTER(PtWorkDivisionSettings), C.c_uint32, C.c_uint32, C.c_uint32, _U]),
    "pt_last_batch_phases": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "pt_sort_pairs": (C.c_int, [_P, C.POINTER(C.c_uint64), _U, C.c_uint32, C.c_uint32,
                                C.POINTER(C.c_uint64), _U]),
    "pt_reduce_samples": (C.c_int, [_P, C.POINTER(PtWorkDivisionSettings), _F, C.c_uint32,
                                    C.c_uint32, _F]),
    "pt_scene_bvh": (C.c_int, [_P, _U, _U, _F, _F, _U, _U, _U, _U, _U]),
    "pt_scene_set_accel": (C.c_int, [_P, C.POINTER(PtSettings)]),
    "pt_clip_triangle_bounds": (C.c_int, [_F, _F, _F, _F, _F]),
    "pt_pack_node": (C.c_int, [_F, _F, C.c_uint32, C.POINTER(C.c_int32), C.c_void_p, _F, _F]),
    "pt_last_render_stats": (C.c_int, [_P, C.POINTER(PtRenderStats)]),
    "pt_set_profiling": (C.c_int, [_P, C.c_int]),
    "pt_scene_generate": (C.c_int, [C.c_uint64, C.POINTER(_P), _F, _U]),
    "pt_render_batch": (C.c_int, [_P, C.POINTER(PtBatchItem), C.c_uint32, C.c_uint32,
                                  C.POINTER(PtSettings), C.POINTER(C.c_double)]),
    "pt_set_launch_params": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_int]),
    "pt_autotune_batch": (C.c_int, [_P, C.POINTER(PtBatchItem), C.c_uint32, C.c_uint32,
                                    C.POINTER(PtSettings), _U]),
    "pt_scene_set_material": (C.c_int, [_P, C.c_uint32, C.POINTER(PtMaterial)]),
    "pt_render_adjoint": (C.c_int, [_P, _P, _F, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.POINTER(PtSettings), _F, _F, _F, C.POINTER(C.c_double)]),
    "pt_mlp_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(_P)]),
    "pt_mlp_destroy": (None, [_P]),
    "pt_mlp_param_count": (C.c_int, [C.POINTER(C.c_uint64)]),
    "pt_mlp_get_weights": (C.c_int, [_P, _F]),
    "pt_mlp_set_weights": (C.c_int, [_P, _F]),
    "pt_mlp_forward": (C.c_int, [_P, _F, C.c_uint32, _F]),
    "pt_mlp_set_forward_kernel": (C.c_int, [_P, C.c_int]),
    "pt_set_bvh_builder": (C.c_int, [_P, C.c_int]),
    "pt_set_pipelines": (C.c_int, [_P, C.c_uint32]),
    "pt_set_megakernel_split": (C.c_int, [_P, C.c_int]),
    "pt_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "pt_device_free": (C.c_int, [_P, _P]),
    "pt_device_read": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "pt_device_write": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "pt_l2_loss_gradient": (C.c_int, [_P, _P, _P, C.c_uint64, _P]),
    "pt_last_bvh_build": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
    "pt_mlp_gradient": (C.c_int, [_P, _F, _F, C.c_uint32, _F, _F]),
    "pt_mlp_train_step": (C.c_int, [_P, _F, _F, C.c_uint32, C.c_float, _F]),
    "pt_mlp_set_dataset": (C.c_int, [_P, _F, _F, C.c_uint64]),
    "pt_mlp_train": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _F,
                               C.POINTER(C.c_double)]),
    "pt_mlp_time_forward": (C.c_int, [_P, C.c_uint32, C.c_uint32, _F, _F]),
    "pt_last_render_fp64": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "pt_render_mlp": (C.c_int, [_P, _P, _P, _F, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.POINTER(PtSettings), _F, C.POINTER(C.c_double)]),
    "pt_render_mlp_first_vertex": (C.c_int, [_P, _P, _P, _F, C.c_uint32, C.c_uint32,
                                             C.c_uint32, C.POINTER(PtSettings), _F,
                                             C.POINTER(C.c_double)]),
}

_lib = None


class PathError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing; build it with `python -m path_amd.build`")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(code: int) -> None:
    if code != PT_OK:
        raise PathError(code, lib().pt_last_error().decode())


def fptr(a):
    """float32 numpy array -> POINTER(c_float) (None passes through)."""
    return None if a is None else a.ctypes.data_as(_F)


def uptr(a):
    return None if a is None else a.ctypes.data_as(_U)
