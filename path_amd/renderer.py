"""Python mirror of the reference renderer interface over the C ABI.

reference: include/render/renderer.h:43-70 (Renderer::render),
include/render/renderer_from_files.h:22-43 (RendererFromFiles),
src/render/config_io.cpp:14-34 (load_config / print_config).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib as L


class Settings:
    """render::Settings. Attribute access reads/writes the POD fields; enum
    fields accept the reference's enum names (e.g. s.term_prob = "NIters")."""

    _ENUMS = {"kernel_approach": L.KERNEL_APPROACH, "accel": L.ACCEL,
              "light_sampler": L.LIGHT_SAMPLER, "dir_sampler": L.DIR_SAMPLER,
              "term_prob": L.TERM_PROB, "rng": L.RNG}

    def __init__(self, raw: L.PtSettings | None = None):
        object.__setattr__(self, "raw", raw if raw is not None else L.PtSettings())
        if raw is None:
            L.lib().pt_settings_default(C.byref(self.raw))

    @classmethod
    def from_yaml(cls, path: str | Path) -> "Settings":
        s = cls()
        L.check(L.lib().pt_settings_load_yaml(str(path).encode(), C.byref(s.raw)))
        return s

    def to_yaml(self) -> str:
        need = C.c_size_t()
        L.check(L.lib().pt_settings_to_yaml(C.byref(self.raw), None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value + 1)
        L.check(L.lib().pt_settings_to_yaml(C.byref(self.raw), buf, need.value + 1, None))
        return buf.value.decode()

    def __getattr__(self, name):
        return getattr(self.raw, name)

    def __setattr__(self, name, value):
        if name in self._ENUMS and isinstance(value, str):
            value = self._ENUMS[name][value]
        setattr(self.raw, name, value)

    def copy(self) -> "Settings":
        raw = L.PtSettings()
        C.memmove(C.byref(raw), C.byref(self.raw), C.sizeof(raw))
        return Settings(raw)


def material(kind: str, diffuse=(0, 0, 0), specular=(0, 0, 0), shininess=1.0,
             ior=1.0, emission=(0, 0, 0), diffuse_weight=1.0,
             glossy_weight=0.0) -> L.PtMaterial:
    m = L.PtMaterial()
    m.type = L.BSDF[kind]
    m.diffuse[:] = list(map(float, diffuse))
    m.specular[:] = list(map(float, specular))
    m.shininess = shininess
    m.ior = ior
    m.emission[:] = list(map(float, emission))
    m.diffuse_weight = diffuse_weight
    m.glossy_weight = glossy_weight
    return m


class Scene:
    """scene::Scene held by the native library."""

    def __init__(self, handle=None):
        if handle is None:
            handle = C.c_void_p()
            L.check(L.lib().pt_scene_create(C.byref(handle)))
        self.handle = handle

    @classmethod
    def load(cls, path: str | Path, width_height_ratio: float = 1.0, quiet: bool = True):
        """ScenefileLoader::load_scene -> (Scene, film_to_world[12])."""
        h = C.c_void_p()
        ftw = np.zeros(12, np.float32)
        L.check(L.lib().pt_scene_load(str(path).encode(), width_height_ratio, int(quiet),
                                      C.byref(h), L.fptr(ftw)))
        return cls(h), ftw

    def add_mesh(self, verts, tri_mats, materials, normals=None, transform=None):
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
        tm = np.ascontiguousarray(tri_mats, np.uint32)
        mats = (L.PtMaterial * len(materials))(*materials)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        tr = None if transform is None else np.ascontiguousarray(transform, np.float32)
        L.check(L.lib().pt_scene_add_mesh(self.handle, L.fptr(verts), L.fptr(nrm), L.uptr(tm),
                                          len(tm), mats, len(materials), L.fptr(tr)))

    def add_obj(self, path, transform=None):
        tr = None if transform is None else np.ascontiguousarray(transform, np.float32)
        L.check(L.lib().pt_scene_add_obj(self.handle, str(path).encode(), L.fptr(tr)))

    def set_material(self, idx: int, m: L.PtMaterial) -> None:
        """Replace material `idx` (geometry and BVH are kept)."""
        L.check(L.lib().pt_scene_set_material(self.handle, idx, C.byref(m)))

    def materials(self) -> list:
        """Copies of the scene's materials (PtMaterial)."""
        f = self.flat()
        return [L.PtMaterial.from_buffer_copy(f["materials"][i]) for i in range(f["n_materials"])]

    def counts(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        L.check(L.lib().pt_scene_counts(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def flat(self) -> dict:
        """World-space flattened arrays (flat_triangle.h:79-146)."""
        nt, nm, nl = self.counts()
        tris = np.zeros((nt, 9), np.float32)
        normals = np.zeros((nt, 9), np.float32)
        tri_mat = np.zeros(nt, np.uint32)
        mats = (L.PtMaterial * max(nm, 1))()
        lt = np.zeros((nl, 9), np.float32)
        lc = np.zeros(nl, np.float32)
        L.check(L.lib().pt_scene_flat(self.handle, L.fptr(tris), L.fptr(normals), L.uptr(tri_mat),
                                      mats, L.fptr(lt), L.fptr(lc)))
        return {"tris": tris, "normals": normals, "tri_mat": tri_mat,
                "materials": mats, "n_materials": nm, "light_tris": lt, "light_cdf": lc}

    def bvh(self) -> dict:
        n = C.c_uint32()
        nr = C.c_uint32()
        L.check(L.lib().pt_scene_bvh(self.handle, C.byref(n), C.byref(nr), None, None, None,
                                     None, None, None, None))
        nn = n.value
        lo = np.zeros((nn, 3), np.float32)
        hi = np.zeros((nn, 3), np.float32)
        leaf = np.zeros(nn, np.uint32)
        a = np.zeros(nn, np.uint32)
        b = np.zeros(nn, np.uint32)
        perm = np.zeros(nr.value, np.uint32)
        depth = C.c_uint32()
        L.check(L.lib().pt_scene_bvh(self.handle, C.byref(n), C.byref(nr), L.fptr(lo),
                                     L.fptr(hi), L.uptr(leaf), L.uptr(a), L.uptr(b),
                                     L.uptr(perm), C.byref(depth)))
        return {"lo": lo, "hi": hi, "is_leaf": leaf, "a": a, "b": b, "perm": perm,
                "max_depth": depth.value}

    def set_accel(self, settings: "Settings") -> None:
        """Build the tree with the accel fields of settings (spatial splits,
        overlap threshold, traversal cost); renders do the same."""
        L.check(L.lib().pt_scene_set_accel(self.handle, C.byref(settings.raw)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and L._lib is not None:
            L.lib().pt_scene_destroy(h)
            self.handle = None


def camera_from_look(look, up, pos, height_angle_deg: float, width_height_ratio: float):
    out = np.zeros(12, np.float32)
    f = lambda v: L.fptr(np.ascontiguousarray(v, np.float32))
    look = np.asarray(look, np.float32)
    up = np.asarray(up, np.float32)
    pos = np.asarray(pos, np.float32)
    L.check(L.lib().pt_camera_from_look(f(look), f(up), f(pos), height_angle_deg,
                                        width_height_ratio, L.fptr(out)))
    return out


@dataclass
class RenderStats:
    kernel_ms: float
    trace_ms: float
    rays: int
    samples: int
    trace_launches: int
    busy_trace_launches: int
    ext_rays: int
    shadow_rays: int
    mega_ms: float
    mega_launches: int
    profiled: bool
    pipelines: int
    replays: int


class DeviceBuffer:
    """Device memory on a renderer's GPU (pt_device_alloc), e.g. the target
    of a DeviceFloatRGB render."""

    def __init__(self, renderer: "Renderer", nbytes: int):
        self.renderer = renderer
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        L.check(L.lib().pt_device_alloc(renderer.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def data_ptr(self) -> int:
        return self.ptr

    def numpy(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        L.check(L.lib().pt_device_read(self.renderer.handle, C.c_void_p(self.ptr),
                                       out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def write(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        L.check(L.lib().pt_device_write(self.renderer.handle, C.c_void_p(self.ptr),
                                        a.ctypes.data_as(C.c_void_p), a.nbytes))

    def __del__(self):
        # at interpreter exit the modules this needs may already be gone
        ptr = getattr(self, "ptr", None)
        if ptr and L is not None and L._lib is not None and C is not None:
            try:
                L.lib().pt_device_free(self.renderer.handle, C.c_void_p(ptr))
            except TypeError:
                pass
            self.ptr = None


class Renderer:
    """render::Renderer on one GPU (ExecutionModel::GPU)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib().pt_renderer_create(device, C.byref(h)))
        self.handle = h

    def render(self, scene: Scene, film_to_world, x_dim: int, y_dim: int, samples_per: int,
               settings: Settings | None = None, output: str = "FloatRGB",
               show_progress: bool = False, show_times: bool = False,
               rays: np.ndarray | None = None, initial_idx: np.ndarray | None = None,
               n_steps: int = 8, device_out=None):
        """Renderer::render. Sample spec: SquareImage, or InitialRays if `rays`
        (n, 6) is given, or InitialIdxAndDir if `initial_idx` (n,) is given
        too (the first intersection is that triangle at distance 1 along each
        ray). Output "BGRA", "FloatRGB" or "OutputPerStep" (n_steps images,
        shape (n_steps, y, x, 3); MegaKernel) or "DeviceFloatRGB" (into
        `device_out`: a DeviceBuffer of this renderer, or a contiguous float32
        torch tensor of y*x*3 elements on its GPU; the image never leaves
        HBM). Returns (image, seconds)."""
        settings = settings or Settings()
        spec = L.PtSampleSpec()
        keep = []
        if rays is not None:
            rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
            keep.append(rays)
            spec.type = L.SAMPLE_SPEC["InitialRays"]
            spec.rays = L.fptr(rays)
            spec.n_rays = rays.shape[0]
            if initial_idx is not None:
                idx = np.ascontiguousarray(initial_idx, np.uint32).ravel()
                assert idx.size == rays.shape[0]
                keep.append(idx)
                spec.type = L.SAMPLE_SPEC["InitialIdxAndDir"]
                spec.initial_idx = L.uptr(idx)
            x_dim, y_dim = rays.shape[0], 1
        else:
            spec.type = L.SAMPLE_SPEC["SquareImage"]
            spec.x_dim, spec.y_dim = x_dim, y_dim
            spec.film_to_world[:] = [float(v) for v in np.asarray(film_to_world).ravel()]
        out = L.PtOutput()
        out.type = L.OUTPUT[output]
        if output == "BGRA":
            img = np.zeros((y_dim, x_dim), np.uint32)
            out.data = img.ctypes.data
        elif output == "DeviceFloatRGB":
            if isinstance(device_out, DeviceBuffer):
                ok = device_out.nbytes >= y_dim * x_dim * 12
            else:  # a torch tensor on the renderer's GPU
                ok = (device_out is not None and getattr(device_out, "is_cuda", False)
                      and device_out.is_contiguous() and device_out.numel() == y_dim * x_dim * 3
                      and str(device_out.dtype) == "torch.float32")
            if not ok:
                raise ValueError("DeviceFloatRGB needs device_out: a DeviceBuffer of at least "
                                 "y_dim * x_dim * 12 bytes or a contiguous float32 CUDA tensor "
                                 "of y_dim * x_dim * 3 elements")
            img = device_out
            out.data = device_out.data_ptr()
            if not isinstance(device_out, DeviceBuffer):
                # pt_render writes on the renderer's own stream: torch work
                # still reading this tensor must finish first. pt_render
                # returns after its stream synchronised, so the image is
                # complete for any torch stream afterwards.
                import torch
                torch.cuda.current_stream(device_out.device).synchronize()
        elif output == "OutputPerStep":
            img = np.zeros((n_steps, y_dim, x_dim, 3), np.float32)
            ptrs = (C.POINTER(C.c_float) * n_steps)(*[L.fptr(img[i]) for i in range(n_steps)])
            keep.append(ptrs)
            out.per_step = ptrs
            out.n_steps = n_steps
        else:
            img = np.zeros((y_dim, x_dim, 3), np.float32)
            out.data = img.ctypes.data
        secs = C.c_double()
        L.check(L.lib().pt_render(self.handle, C.byref(spec), C.byref(out), scene.handle,
                                  samples_per, C.byref(settings.raw), int(show_progress),
                                  int(show_times), C.byref(secs)))
        return img, secs.value

    def last_render_fp64(self, x_dim: int, y_dim: int) -> np.ndarray:
        """The last Streaming SquareImage render re-summed per pixel in FP64."""
        out = np.zeros((y_dim, x_dim, 3), np.float64)
        L.check(L.lib().pt_last_render_fp64(self.handle,
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def intersect(self, scene: Scene, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        idx = np.zeros(n, np.uint32)
        back = np.zeros(n, np.uint32)
        L.check(L.lib().pt_intersect(self.handle, scene.handle, L.fptr(rays), n, L.fptr(t),
                                     L.uptr(idx), L.uptr(back)))
        return t, idx, back

    def adjoint(self, scene: Scene, film_to_world, x_dim: int, y_dim: int, samples_per: int,
                d_image: np.ndarray, settings: Settings | None = None):
        """Path-replay adjoint: gradients of sum(d_image * image) w.r.t. every
        material's albedo color and emission. Returns (grad_albedo (n, 3),
        grad_emission (n, 3), seconds)."""
        settings = settings or Settings()
        if isinstance(d_image, DeviceBuffer):  # already in HBM
            assert d_image.nbytes >= y_dim * x_dim * 12
            d_ptr = C.cast(C.c_void_p(d_image.ptr), C.POINTER(C.c_float))
        else:
            d = np.ascontiguousarray(d_image, np.float32).reshape(y_dim, x_dim, 3)
            d_ptr = L.fptr(d)
        nm = scene.counts()[1]
        ga = np.zeros((nm, 3), np.float32)
        ge = np.zeros((nm, 3), np.float32)
        ftw = np.ascontiguousarray(np.asarray(film_to_world).ravel(), np.float32)
        secs = C.c_double()
        L.check(L.lib().pt_render_adjoint(self.handle, scene.handle, L.fptr(ftw), x_dim, y_dim,
                                          samples_per, C.byref(settings.raw), d_ptr,
                                          L.fptr(ga), L.fptr(ge), C.byref(secs)))
        return ga, ge, secs.value

    def render_rays_batch(self, items, samples_per: int, n_steps: int = 0,
                          settings: Settings | None = None):
        """Ray sample specs of many scenes in one batched launch
        (pt_render_rays_batch). items: (scene, rays (n, 6), initial_idx (n,)
        or None); all with or all without initial_idx. Returns per item an
        array (n_steps, n, 3) (OutputPerStep), or (n, 3) when n_steps = 0
        (FloatRGB); each equals render() of that item alone with the
        MegaKernel integrator."""
        settings = settings or Settings()
        arr = (L.PtRayBatchItem * max(len(items), 1))()
        keep, outs = [], []
        for i, (scene, rays, idx) in enumerate(items):
            rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
            n = rays.shape[0]
            out = np.zeros((max(n_steps, 1), n, 3), np.float32)
            keep.append(rays)
            arr[i].scene = scene.handle
            arr[i].rays = L.fptr(rays)
            arr[i].n_rays = n
            arr[i].out = L.fptr(out)
            if idx is not None:
                idx = np.ascontiguousarray(idx, np.uint32).ravel()
                assert idx.size == n
                keep.append(idx)
                arr[i].initial_idx = L.uptr(idx)
            outs.append(out if n_steps else out[0])
        secs = C.c_double()
        L.check(L.lib().pt_render_rays_batch(self.handle, arr, len(items), samples_per, n_steps,
                                             C.byref(settings.raw), C.byref(secs)))
        return outs

    def reduce_samples(self, wds: L.PtWorkDivisionSettings, values: np.ndarray,
                       num_locations: int, samples_per: int) -> np.ndarray:
        """The renderer's per-pixel reduction on the GPU applied to
        values[sample + location * samples_per] (one sample block)."""
        v = np.ascontiguousarray(values, np.float32).ravel()
        assert v.size == num_locations * samples_per
        out = np.zeros(num_locations, np.float32)
        L.check(L.lib().pt_reduce_samples(self.handle, C.byref(wds), L.fptr(v), num_locations,
                                          samples_per, L.fptr(out)))
        return out

    def sobol_table(self, dim_bound: int, count: int) -> np.ndarray:
        out = np.zeros((count, dim_bound), np.float32)
        L.check(L.lib().pt_sobol_table(self.handle, dim_bound, count, L.fptr(out)))
        return out

    def sort_pairs(self, keys, vals, end_bit: int = 64):
        """Stable GPU radix sort of (uint64 key, uint32 value) pairs by key
        bits [0, end_bit): the device tree builder's sort."""
        k = np.ascontiguousarray(keys, np.uint64)
        v = np.ascontiguousarray(vals, np.uint32)
        ko = np.zeros_like(k)
        vo = np.zeros_like(v)
        P64 = C.POINTER(C.c_uint64)
        L.check(L.lib().pt_sort_pairs(self.handle, k.ctypes.data_as(P64), L.uptr(v), len(k),
                                      end_bit, ko.ctypes.data_as(P64), L.uptr(vo)))
        return ko, vo

    def last_stats(self) -> RenderStats:
        s = L.PtRenderStats()
        L.check(L.lib().pt_last_render_stats(self.handle, C.byref(s)))
        return RenderStats(s.kernel_ms, s.trace_ms, s.rays, s.samples, s.trace_launches,
                           s.busy_trace_launches, s.ext_rays, s.shadow_rays, s.mega_ms,
                           s.mega_launches, bool(s.profiled), int(s.pipelines),
                           int(s.replays))

    def l2_loss_gradient(self, image: DeviceBuffer, target: DeviceBuffer, n_floats: int,
                         out: DeviceBuffer) -> DeviceBuffer:
        """out = 2 (image - target) on the GPU (device buffers of n_floats)."""
        assert min(image.nbytes, target.nbytes, out.nbytes) >= 4 * n_floats
        L.check(L.lib().pt_l2_loss_gradient(self.handle, C.c_void_p(image.ptr),
                                            C.c_void_p(target.ptr), n_floats,
                                            C.c_void_p(out.ptr)))
        return out

    def set_megakernel_split(self, enable: bool) -> None:
        """Several lanes per work-division thread in the MegaKernel
        integrator (default on; the image does not depend on it)."""
        L.check(L.lib().pt_set_megakernel_split(self.handle, int(enable)))

    def set_profiling(self, enable: bool) -> None:
        """Bracket every traversal / mega-kernel launch with HIP events."""
        L.check(L.lib().pt_set_profiling(self.handle, int(enable)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and L._lib is not None:
            L.lib().pt_renderer_destroy(h)
            self.handle = None


def sobol_direction_numbers(dim_bound: int) -> np.ndarray:
    out = np.zeros((dim_bound, 32), np.uint32)
    L.check(L.lib().pt_sobol_direction_numbers(dim_bound, L.uptr(out)))
    return out


def work_division(settings: L.PtWorkDivisionSettings, samples_per: int, x_dim: int, y_dim: int):
    out = np.zeros(11, np.uint32)
    L.check(L.lib().pt_work_division(C.byref(settings), samples_per, x_dim, y_dim, L.uptr(out)))
    keys = ["block_size", "base_samples_per_thread", "n_threads_per_unit_extra",
            "sample_block_size", "x_block_size", "y_block_size", "num_sample_blocks",
            "num_x_blocks", "num_y_blocks", "x_dim", "y_dim"]
    return dict(zip(keys, map(int, out)))


class RendererFromFiles:
    """RendererFromFiles (include/render/renderer_from_files.h:22-43)."""

    def __init__(self, device: int = 0):
        self.renderer = None
        self.device = device
        self.settings = Settings()
        self.scene = None
        self.film_to_world = None

    def load_scene(self, path, width_height_ratio: float = 1.0, quiet: bool = False):
        self.scene, self.film_to_world = Scene.load(path, width_height_ratio, quiet)

    def load_config(self, path):
        self.settings = Settings.from_yaml(path)

    def print_config(self):
        print(self.settings.to_yaml())

    def render(self, x_dim: int, y_dim: int, samples_per: int, output: str = "BGRA",
               show_progress: bool = False, show_times: bool = False):
        if self.renderer is None:
            self.renderer = Renderer(self.device)
        return self.renderer.render(self.scene, self.film_to_world, x_dim, y_dim, samples_per,
                                    self.settings, output, show_progress, show_times)


def clip_triangle_bounds(tri, lo, hi):
    """Bounds (lo, hi) of the part of triangle tri (3 x 3) inside the box
    [lo, hi]; lo > hi when they do not meet. The clipping a spatial split
    applies to a triangle (include/intersect/accel/detail/clipped_triangle_impl.h)."""
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    blo = np.ascontiguousarray(lo, np.float32).reshape(3)
    bhi = np.ascontiguousarray(hi, np.float32).reshape(3)
    olo = np.zeros(3, np.float32)
    ohi = np.zeros(3, np.float32)
    L.check(L.lib().pt_clip_triangle_bounds(L.fptr(t), L.fptr(blo), L.fptr(bhi), L.fptr(olo),
                                            L.fptr(ohi)))
    return olo, ohi
