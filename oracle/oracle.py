"""ctypes binding of the CPU checker (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py. The product (path_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from path_amd import _lib as L

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing; build with `python -m path_amd.build --oracle`")
        l = C.CDLL(str(LIB_PATH))
        l.oracle_loopall.argtypes = [_F, C.c_uint32, _F, C.c_uint32, _F, _U, _U]
        l.oracle_bvh_intersect.argtypes = [_F, C.c_uint32, _F, C.c_uint32, _F, _U, _U, _U, _U]
        l.oracle_sobol_table.argtypes = [_U, C.c_uint32, C.c_uint32, _F]
        l.oracle_fnv_hash.argtypes = [C.c_uint32]
        l.oracle_fnv_hash.restype = C.c_uint32
        l.oracle_work_division.argtypes = [C.POINTER(L.PtWorkDivisionSettings), C.c_uint32,
                                           C.c_uint32, C.c_uint32, _U, C.c_uint32]
        l.oracle_render.argtypes = [_F, _F, _U, C.c_uint32, C.POINTER(L.PtMaterial), C.c_uint32,
                                    _F, _F, C.c_uint32, _F, C.POINTER(L.PtSettings), C.c_uint32,
                                    C.c_uint32, C.c_uint32, _F, C.c_int, C.c_int, C.c_uint32,
                                    C.c_uint32, _F, _D, _U64, _D]
        l.oracle_render_ex.argtypes = [_F, _U, C.c_uint32, _F] + l.oracle_render.argtypes
        l.oracle_adjoint.argtypes = [_F, _F, _U, C.c_uint32, C.POINTER(L.PtMaterial), C.c_uint32,
                                     _F, _F, C.c_uint32, _F, C.POINTER(L.PtSettings), C.c_uint32,
                                     C.c_uint32, C.c_uint32, _F, _D, C.c_int, _D, _D, _D, _D]
        l.oracle_max_threads.restype = C.c_int
        l.oracle_bsdf_sample.argtypes = [C.POINTER(L.PtMaterial), _D, _D, _D, C.c_uint32,
                                         C.c_int, _D]
        l.oracle_bsdf_sample.restype = None
        l.oracle_bsdf_sample_f32.argtypes = l.oracle_bsdf_sample.argtypes
        l.oracle_bsdf_sample_f32.restype = None
        l.oracle_bsdf_eval.argtypes = [C.POINTER(L.PtMaterial), _D, _D, _D, _D]
        l.oracle_bsdf_eval.restype = None
        l.oracle_term_prob.argtypes = [C.POINTER(L.PtSettings), C.c_uint32, _D]
        l.oracle_term_prob.restype = C.c_double
        l.oracle_detmath.argtypes = [C.c_int, _F, _F, _F, C.c_uint32]
        l.oracle_detmath.restype = None
        l.oracle_reduce_samples.argtypes = [C.POINTER(L.PtWorkDivisionSettings), _F, C.c_uint32,
                                            C.c_uint32, _F]
        l.oracle_work_division_walk.argtypes = [C.POINTER(L.PtWorkDivisionSettings), C.c_uint32,
                                                C.c_uint32, C.c_uint32]
        l.oracle_work_division_walk.restype = C.c_uint64
        l.oracle_to_bgra.argtypes = [_F, C.c_uint32, C.POINTER(C.c_uint8)]
        l.oracle_to_bgra.restype = None
        l.oracle_hash_rng.argtypes = [C.c_uint32, C.c_uint32, _F]
        l.oracle_hash_rng.restype = None
        l.indep_render.argtypes = [_F, _F, _U, C.c_uint32, C.POINTER(L.PtMaterial), C.c_uint32,
                                   _F, C.POINTER(L.PtSettings), C.c_uint32, C.c_uint32,
                                   C.c_uint32, _F, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                   _D]
        _lib = l
    return _lib


def _f(a):
    return a.ctypes.data_as(_F)


def _u(a):
    return a.ctypes.data_as(_U)


def loopall(tris: np.ndarray, rays: np.ndarray):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    t, idx, back = np.zeros(n, np.float32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    lib().oracle_loopall(_f(tris), tris.shape[0], _f(rays), n, _f(t), _u(idx), _u(back))
    return t, idx, back


def bvh_intersect(tris: np.ndarray, rays: np.ndarray):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    t, idx, back = np.zeros(n, np.float32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    nodes, ntri = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    lib().oracle_bvh_intersect(_f(tris), tris.shape[0], _f(rays), n, _f(t), _u(idx), _u(back),
                               _u(nodes), _u(ntri))
    return t, idx, back, nodes, ntri


def sobol_table(dirs: np.ndarray, dim_bound: int, count: int) -> np.ndarray:
    dirs = np.ascontiguousarray(dirs, np.uint32)
    out = np.zeros((count, dim_bound), np.float32)
    lib().oracle_sobol_table(_u(dirs), dim_bound, count, _f(out))
    return out


def fnv_hash(v: int) -> int:
    return int(lib().oracle_fnv_hash(v))


def thread_infos(wds: L.PtWorkDivisionSettings, spp: int, x: int, y: int, max_threads=1 << 22):
    out = np.zeros((max_threads, 5), np.uint32)
    n = lib().oracle_work_division(C.byref(wds), spp, x, y, _u(out), max_threads)
    if n < 0:
        raise ValueError("invalid work division")
    if n > max_threads:
        raise ValueError("too many threads")
    return out[:n]


def render(flat: dict, film_to_world, settings, samples_per: int, x_dim: int, y_dim: int,
           sobol: np.ndarray | None, precision: int = 32, threads: int = 0,
           pixel_begin: int = 0, pixel_end: int = 0, initial_rays=None, initial_idx=None,
           n_steps: int = 0, gbuffer: bool = False):
    """Returns (image (y, x, 3), info) with info = {"counts": [[rays, nodes,
    tris] for continuation/camera rays, same for light rays], "seconds":
    render time without scene setup}. With initial_rays ((n, 6): one ray per
    "pixel", x_dim = n, y_dim = 1; with initial_idx the first intersection
    is that triangle at distance 1) and / or n_steps (per-step images,
    shape (n_steps, y, x, 3)). With gbuffer (float renders), info["gbuffer"]
    holds the first hit of sample 0 per pixel, (y, x, 12): position, shading
    normal, albedo summary, emission; zeros where the camera ray misses."""
    raw = settings.raw if hasattr(settings, "raw") else settings
    ftw = np.ascontiguousarray(film_to_world, np.float32)
    sob = np.zeros(1, np.float32) if sobol is None else np.ascontiguousarray(sobol, np.float32)
    counts = np.zeros(6, np.uint64)
    secs = C.c_double()
    n_img = max(n_steps, 1)
    out32 = np.zeros((n_img, y_dim, x_dim, 3), np.float32)
    out64 = np.zeros((n_img, y_dim, x_dim, 3), np.float64)
    rays = None if initial_rays is None else np.ascontiguousarray(initial_rays, np.float32)
    idx = None if initial_idx is None else np.ascontiguousarray(initial_idx, np.uint32)
    gb = np.zeros((y_dim, x_dim, 12), np.float32) if gbuffer and precision == 32 else None
    rc = lib().oracle_render_ex(
        _f(rays) if rays is not None else None, _u(idx) if idx is not None else None, n_steps,
        _f(gb) if gb is not None else None, _f(flat["tris"]), _f(flat["normals"]),
        _u(flat["tri_mat"]), flat["tris"].shape[0],
        flat["materials"], flat["n_materials"], _f(flat["light_tris"]), _f(flat["light_cdf"]),
        flat["light_tris"].shape[0], _f(ftw), C.byref(raw), samples_per, x_dim, y_dim, _f(sob),
        precision, threads, pixel_begin, pixel_end, _f(out32), out64.ctypes.data_as(_D),
        counts.ctypes.data_as(_U64), C.byref(secs))
    if rc != 0:
        raise ValueError("oracle render failed")
    info = {"counts": counts.reshape(2, 3), "seconds": secs.value}
    if gb is not None:
        info["gbuffer"] = gb
    img = out64 if precision == 64 else out32
    return (img if n_steps else img[0]), info


def adjoint(flat: dict, film_to_world, settings, samples_per: int, x_dim: int, y_dim: int,
            sobol: np.ndarray | None, d_image: np.ndarray, threads: int = 0):
    """FP64 forward-mode gradients of sum(d_image * image) w.r.t. every
    material's albedo color and emission for the mega-kernel estimator with
    the same samples. Returns (grad_albedo (n, 3), grad_emission (n, 3),
    image (y, x, 3) FP64, seconds)."""
    raw = settings.raw if hasattr(settings, "raw") else settings
    ftw = np.ascontiguousarray(film_to_world, np.float32)
    sob = np.zeros(1, np.float32) if sobol is None else np.ascontiguousarray(sobol, np.float32)
    d_img = np.ascontiguousarray(d_image, np.float64).reshape(y_dim, x_dim, 3)
    nm = flat["n_materials"]
    ga = np.zeros((nm, 3), np.float64)
    ge = np.zeros((nm, 3), np.float64)
    img = np.zeros((y_dim, x_dim, 3), np.float64)
    secs = C.c_double()
    rc = lib().oracle_adjoint(
        _f(flat["tris"]), _f(flat["normals"]), _u(flat["tri_mat"]), flat["tris"].shape[0],
        flat["materials"], nm, _f(flat["light_tris"]), _f(flat["light_cdf"]),
        flat["light_tris"].shape[0], _f(ftw), C.byref(raw), samples_per, x_dim, y_dim, _f(sob),
        d_img.ctypes.data_as(_D), threads, ga.ctypes.data_as(_D), ge.ctypes.data_as(_D),
        img.ctypes.data_as(_D), C.byref(secs))
    if rc != 0:
        raise ValueError(f"oracle adjoint failed ({rc})")
    return ga, ge, img, secs.value


def reduce_samples(wds: L.PtWorkDivisionSettings, values: np.ndarray, num_locations: int,
                   samples_per: int) -> np.ndarray:
    """Work-division reduction of values[sample + location * samples_per]
    (one sample block): per-thread sequential sums, then the lane tree."""
    v = np.ascontiguousarray(values, np.float32).ravel()
    assert v.size == num_locations * samples_per
    out = np.zeros(num_locations, np.float32)
    if lib().oracle_reduce_samples(C.byref(wds), _f(v), num_locations, samples_per, _f(out)):
        raise ValueError("invalid work division or more than one sample block")
    return out


def work_division_walk(wds: L.PtWorkDivisionSettings, spp: int, x: int, y: int) -> tuple:
    """The reference test's coverage walk: (failed check, thread ordinal),
    (0, 0) if every check holds."""
    r = int(lib().oracle_work_division_walk(C.byref(wds), spp, x, y))
    return r >> 32, r & 0xFFFFFFFF


def to_bgra(rgb: np.ndarray) -> np.ndarray:
    """FloatRGB (..., 3) -> bytes (..., 4) in memory order B, G, R, 0."""
    v = np.ascontiguousarray(rgb, np.float32)
    out = np.zeros(v.shape[:-1] + (4,), np.uint8)
    lib().oracle_to_bgra(_f(v), v.size // 3, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def hash_rng(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().oracle_hash_rng(seed, n, _f(out))
    return out


def indep_render(flat: dict, film_to_world, settings, samples_per: int, x_dim: int, y_dim: int,
                 sobol: np.ndarray, pixel_begin: int = 0, pixel_end: int = 0,
                 threads: int = 0, mistake: int = 0) -> np.ndarray:
    """Per-pixel mean radiance (y, x, 3) in FP64 from the independent
    restatement of the reference estimator (oracle/indep.cpp, which shares no
    code with the product). Sobol rng only. `mistake` plants a known error
    in the restatement (tests of the comparison's sensitivity)."""
    raw = settings.raw if hasattr(settings, "raw") else settings
    ftw = np.ascontiguousarray(film_to_world, np.float32)
    sob = np.ascontiguousarray(sobol, np.float32)
    out = np.zeros((y_dim, x_dim, 3), np.float64)
    rc = lib().indep_render(_f(flat["tris"]), _f(flat["normals"]), _u(flat["tri_mat"]),
                            flat["tris"].shape[0], flat["materials"], flat["n_materials"],
                            _f(ftw), C.byref(raw), samples_per, x_dim, y_dim, _f(sob),
                            pixel_begin, pixel_end, threads, mistake,
                            out.ctypes.data_as(_D))
    if rc != 0:
        raise ValueError(f"independent render failed ({rc})")
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def detmath(fn: str, x, y=None):
    code = {"sincos_2pi": 0, "log2": 1, "exp2": 2, "pow": 3}[fn]
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, np.float32)
    out = np.zeros(x.size * (2 if code == 0 else 1), np.float32)
    lib().oracle_detmath(code, _f(x), _f(y), _f(out), x.size)
    return out.reshape(-1, 2) if code == 0 else out


def joe_kuo_directions(dim_bound: int) -> np.ndarray:
    """Joe-Kuo direction numbers as published through rocRAND's host API
    (rocrand_get_direction_vectors32, set ROCRAND_DIRECTION_VECTORS_32_JOEKUO6)."""
    rr = C.CDLL("/opt/rocm/lib/librocrand.so")
    ptr = C.POINTER(C.c_uint32)()
    if rr.rocrand_get_direction_vectors32(C.byref(ptr), 101) != 0:
        raise RuntimeError("rocrand direction numbers unavailable")
    return np.ctypeslib.as_array(ptr, shape=(dim_bound * 32,)).reshape(dim_bound, 32).copy()


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_D)


def bsdf_sample(mat, incoming, normal, u, dir_sampler=1, precision=64):
    """-> (dir, multiplier, is_discrete, n_values_consumed); precision 32
    evaluates the shared code in float."""
    out = np.zeros(8)
    u = np.ascontiguousarray(u, np.float64)
    fn = lib().oracle_bsdf_sample if precision == 64 else lib().oracle_bsdf_sample_f32
    fn(C.byref(mat), _d(incoming), _d(normal), u.ctypes.data_as(_D), u.size, dir_sampler,
       out.ctypes.data_as(_D))
    return out[:3], out[3:6], bool(out[6]), int(out[7])


def bsdf_eval(mat, incoming, outgoing, normal):
    out = np.zeros(4)
    lib().oracle_bsdf_eval(C.byref(mat), _d(incoming), _d(outgoing), _d(normal),
                           out.ctypes.data_as(_D))
    return out[:3], out[3]


def term_prob(settings, iters, mult):
    raw = settings.raw if hasattr(settings, "raw") else settings
    return float(lib().oracle_term_prob(C.byref(raw), iters, _d(mult)))
