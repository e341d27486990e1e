"""Training-data generation: seeded random scenes rendered many per launch.

The reference renders one scene per Renderer::render call
(src/generate_data/full_scene/generate_data.cpp:280-294) and its Python
datagen drives that loop scene by scene. Here a batch of scenes shares one
wavefront path pool: every image is bit-identical to rendering its scene alone
with the Streaming integrator, and the first-hit G-buffer (position, shading
normal, albedo, emission of sample 0) is written on the way.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .renderer import Renderer, Scene, Settings


@dataclass
class GeneratedScene:
    scene: Scene
    film_to_world: np.ndarray  # (12,) float32
    resolution: int
    seed: int


def generate_scene(seed: int) -> GeneratedScene:
    """Deterministic random scene (boxes/spheres, 1-3 quad emitters)."""
    h = C.c_void_p()
    ftw = np.zeros(12, np.float32)
    res = C.c_uint32()
    L.check(L.lib().pt_scene_generate(C.c_uint64(seed), C.byref(h), L.fptr(ftw),
                                      C.byref(res)))
    return GeneratedScene(Scene(h), ftw, int(res.value), seed)


# pt_batch_item as a numpy record (same 80-byte layout): a batch's items are
# filled column by column (filling 1024 ctypes structures field by field took
# ~7 ms of Python per batch).
_ITEM = np.dtype([("scene", "<u8"), ("x_dim", "<u4"), ("y_dim", "<u4"),
                  ("film_to_world", "<f4", (12,)), ("out_rgb", "<u8"), ("out_gbuffer", "<u8")])
assert _ITEM.itemsize == C.sizeof(L.PtBatchItem)
assert [_ITEM.fields[f][1] for f in _ITEM.names] == [
    getattr(L.PtBatchItem, f).offset for f, _ in L.PtBatchItem._fields_]


def _item_records(scenes, out_rgb=None, out_gbuffer=None):
    """Records of `scenes` (scene, size, camera) with the given output
    addresses (sequences of ints, or None for null), and the pointer to pass."""
    a = np.zeros(len(scenes), _ITEM)
    a["scene"] = [g.scene.handle.value or 0 for g in scenes]
    res = np.fromiter((g.resolution for g in scenes), np.uint32, len(scenes))
    a["x_dim"] = res
    a["y_dim"] = res
    if len(scenes):
        a["film_to_world"] = np.stack([np.asarray(g.film_to_world, np.float32) for g in scenes])
    if out_rgb is not None:
        a["out_rgb"] = out_rgb
    if out_gbuffer is not None:
        a["out_gbuffer"] = out_gbuffer
    return a, a.ctypes.data_as(C.POINTER(L.PtBatchItem))


def _items(scenes, gbuffer: bool):
    images = [np.zeros((g.resolution, g.resolution, 3), np.float32) for g in scenes]
    gbufs = [np.zeros((g.resolution, g.resolution, 12), np.float32) if gbuffer else None
             for g in scenes]
    rec, ptr = _item_records(scenes, [im.ctypes.data for im in images],
                             [gb.ctypes.data for gb in gbufs] if gbuffer else None)
    return (rec, ptr), images, gbufs


def render_batch(renderer: Renderer, scenes, samples_per: int,
                 settings: Settings | None = None, gbuffer: bool = True):
    """Render every scene of `scenes` (GeneratedScene list) in one launch
    sequence. Returns (images, gbuffers, seconds)."""
    settings = settings or datagen_settings()
    (rec, items), images, gbufs = _items(scenes, gbuffer)
    secs = C.c_double()
    L.check(L.lib().pt_render_batch(renderer.handle, items, len(scenes), samples_per,
                                    C.byref(settings.raw), C.byref(secs)))
    return images, gbufs, secs.value


def render_batch_device(renderer: Renderer, scenes, samples_per: int,
                        settings: Settings | None = None, gbuffer: bool = True, out=None):
    """render_batch with the results left in HBM: the images (and G-buffers)
    of all scenes packed in batch order in one device buffer each, scene i
    from pixel offsets[i]. `out` (a previous return value) is reused when it
    is large enough. Returns (rgb, gbuffer or None, offsets, seconds)."""
    from .renderer import DeviceBuffer
    settings = settings or datagen_settings()
    px = [g.resolution * g.resolution for g in scenes]
    offsets = np.concatenate([[0], np.cumsum(px)]).astype(np.int64)
    total = int(offsets[-1])
    rgb = gb = None
    if out is not None:
        rgb, gb = out[0], out[1]
        if rgb.nbytes < total * 12 or (gbuffer and (gb is None or gb.nbytes < total * 48)):
            rgb = gb = None
    if rgb is None:
        rgb = DeviceBuffer(renderer, max(total, 1) * 12)
        gb = DeviceBuffer(renderer, max(total, 1) * 48) if gbuffer else None
    rec, items = _item_records(scenes, rgb.ptr + 12 * offsets[:-1].astype(np.uint64),
                               gb.ptr + 48 * offsets[:-1].astype(np.uint64) if gbuffer else None)
    secs = C.c_double()
    L.check(L.lib().pt_render_batch(renderer.handle, items, len(scenes), samples_per,
                                    C.byref(settings.raw), C.byref(secs)))
    return rgb, (gb if gbuffer else None), offsets[:-1], secs.value


def autotune(renderer: Renderer, scenes, samples_per: int, settings: Settings) -> dict:
    """Pick the fastest (path pool, traversal blocks per CU, iterations per
    host sync) for this workload; the renderer and `settings` keep it."""
    # the output arrays must outlive the call: the renders write into them
    (rec, items), images, _ = _items(scenes, False)
    out = (C.c_uint32 * 5)()
    L.check(L.lib().pt_autotune_batch(renderer.handle, items, len(scenes), samples_per,
                                      C.byref(settings.raw), out))
    del images
    return {"max_num_samples_per_launch": out[0], "trace_blocks_per_cu": out[1],
            "iters_per_sync": out[2], "graphs": bool(out[3]), "pipelines": out[4]}


def set_launch_params(renderer: Renderer, trace_blocks_per_cu: int = 0,
                      iters_per_sync: int = 1, use_graphs: bool = False) -> None:
    """Wavefront launch parameters; trace_blocks_per_cu 0 = the defaults (3
    per CU for a render, 6 for a batch)."""
    L.check(L.lib().pt_set_launch_params(renderer.handle, trace_blocks_per_cu,
                                         iters_per_sync, int(use_graphs)))


def set_pipelines(renderer: Renderer, n: int) -> None:
    """Wavefront pipelines of a Streaming render (1-4, default 3)."""
    L.check(L.lib().pt_set_pipelines(renderer.handle, n))


BVH_BUILDERS = {False: 0, "host": 0, "lbvh": 1, "ploc": 2, True: 2}


def set_bvh_builder(renderer: Renderer, gpu) -> None:
    """Trees of render_batch: built on the GPU for the whole batch (True or
    "ploc": locally-ordered clustering; "lbvh": the binary radix tree) or by
    the host SAH builder, cached per scene (False or "host", the default)."""
    if gpu not in BVH_BUILDERS:
        raise ValueError(f"unknown BVH builder {gpu!r}")
    L.check(L.lib().pt_set_bvh_builder(renderer.handle, BVH_BUILDERS[gpu]))


def last_bvh_build(renderer: Renderer) -> tuple:
    """(milliseconds, levels of the deepest 4-wide tree) of the last GPU build."""
    ms, lv = C.c_double(), C.c_uint32()
    L.check(L.lib().pt_last_bvh_build(renderer.handle, C.byref(ms), C.byref(lv)))
    return float(ms.value), int(lv.value)


def last_batch_phases(renderer: Renderer) -> dict:
    """Host wall ms of the phases of the last batch that was not resident:
    prepare (flatten, shading records, host trees), pack, upload, GPU trees."""
    ms = (C.c_double * 4)()
    L.check(L.lib().pt_last_batch_phases(renderer.handle, ms))
    return {"prepare": ms[0], "pack": ms[1], "upload": ms[2], "gpu_trees": ms[3]}


def datagen_settings() -> Settings:
    """Streaming integrator, Sobol sampling, BSDF directions, NEE, and a path
    pool of 16 M paths (about 2 GB of path state): a batch of many scenes
    keeps more paths in flight than one image needs (1024 generated scenes at
    128 spp, batch render ms: 4 M paths 147, 8 M 136, 16 M 134). A render's
    pool never exceeds its own sample count, so small renders are unchanged."""
    s = Settings()
    s.kernel_approach = "Streaming"
    s.max_num_samples_per_launch = 1 << 24
    return s
