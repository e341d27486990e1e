"""Benchmark: Msamples/s of the wavefront path tracer on one MI355X.

Workload (default): Cornell box + seeded level-6 icosphere (81,920 triangles,
radius jitter +-5%, materials diffuse / glossy (Phong n=50) / glass (IOR 1.5)
by triangle index mod 3), 512x512, 64 spp, max depth 8 (NIters 7), Sobol
sampler, FP32, Streaming (wavefront) integrator. One step = one full render
of that image with the scene already resident in HBM (BVH and Sobol table are
built before the timed region, as in the reference's timing convention).

Prints one JSON line with `roofline` (traversal kernel vs HBM peak, bytes
counted with the reference's AoS sizes on the ray set the CPU checker traces
for the same pixels) and `cpu_baseline` (the FP32 CPU checker on the host's
cores, bounded pixel sample of the same workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3]

`--config c4` times one gradient of an L2 image loss (mega-kernel render +
path-replay adjoint) on the Cornell box at 256x256, 32 spp, with respect to
the 8 material albedos and the emitter radiance, and reports its agreement
with central differences (h = 1e-3) and the FP64 CPU adjoint.

`--config c5` builds 50 M training samples with the batched datagen path,
trains the radiance MLP (BF16 MFMA kernels) and compares, on a held-out
generated scene at 512x512, the plain tracer and the MLP-assisted one-bounce
estimator by time to 1% relative RMSE against a 4096-spp render.

`--config c3` is the batched training-data workload instead: 1024 seeded
scenes (2-40 boxes/spheres, 20-20k triangles log-uniform, resolutions
{32,64,96,128}^2), 128 spp, max depth 6, FP32 radiance + G-buffer; one step
renders the whole batch through the shared path pool and the line reports
scenes/s next to the same scenes rendered one render call per scene.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s on 1 GPU at fixed RMSE; MLP-assisted vs plain time-to-1%-RMSE"
HBM_PEAK_GBPS = 8000.0
ENGINE_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
RAY_BYTES = 24 + 16          # ray read + hit write (reference AoS sizes)
NODE_BYTES = 32              # bvh::Node
TRI_BYTES = 36               # intersect::Triangle

CONFIGS = {
    "c2": dict(workload="cornell_icosphere_l6_512x512_64spp_depth8_wavefront",
               level=6, size=512, spp=64, iters=7, approach="Streaming"),
    "c1": dict(workload="cornell_box_36tri_256x256_16spp_depth5_megakernel",
               level=None, size=256, spp=16, iters=4, approach="MegaKernel"),
    "c3": dict(workload="datagen_1024_seeded_scenes_128spp_depth6_gbuffer",
               n_scenes=1024, spp=128, iters=5, approach="Streaming"),
    "c4": dict(workload="cornell_box_adjoint_256x256_32spp_depth5_8albedo_emitter",
               level=None, size=256, spp=32, iters=4, approach="MegaKernel"),
    "c5": dict(workload="radiance_mlp_6x256_50M_datagen_samples_heldout_512x512",
               rows=50_000_000, spp=128, iters=5, size=512, gt_spp=4096),
}
MLP_MACS = 225 * 256 + 5 * 256 * 256 + 256 * 3   # per query, valid (unpadded) dims


def _roofline(name, ms, launches, n_rays, bpr, config, step_bytes=None, step_ms=None,
              pipelines=1):
    """Roofline of the dominant traversal kernel.

    Top level (`achieved`, `frac`): per GPU - the algorithmic bytes of every
    traversal launch of one step (step_bytes) over the measured step time, so
    overlapping launches cannot count the same time twice. `per_launch`: the
    dominant kernel's algorithmic bytes per launch over its average launch
    duration (HIP events on the launch's stream), the figure the rocprofv3
    kernel trace in profiles/ checks. `traffic`: HBM bytes per launch of that
    kernel from the FETCH_SIZE / WRITE_SIZE counters, and `binding`: the
    counters that show what limits it when the algorithmic figure exceeds the
    bandwidth (it is not a bound then)."""
    avg_s = ms / 1e3 / max(launches, 1)
    bytes_per_launch = bpr * n_rays / max(launches, 1)
    per_launch_gbps = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    if step_bytes is None or not step_ms:
        step_bytes, step_ms = bytes_per_launch, avg_s * 1e3
    achieved = step_bytes / (step_ms / 1e3) / 1e9
    prof = {}
    tf = ROOT / "profiles" / f"{name}_{config}_traffic.json"
    if tf.exists():
        prof = json.loads(tf.read_text())
    traffic = prof.get("bytes_per_dispatch")
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
           "traffic": traffic, "kernel": name,
           "scope": "per GPU: algorithmic bytes of every traversal launch of one step / ms_per_step",
           "algorithmic_bytes_per_step": round(step_bytes),
           "per_launch": {"avg_launch_ms": round(avg_s * 1e3, 4),
                          "algorithmic_bytes_per_ray": round(bpr, 1),
                          "rays_per_launch": round(n_rays / max(launches, 1), 1),
                          "achieved": round(per_launch_gbps, 2),
                          "frac": round(per_launch_gbps / HBM_PEAK_GBPS, 5),
                          "concurrent_pipelines": pipelines}}
    if pipelines > 1:
        out["per_launch"]["note"] = (
            f"{pipelines} pipelines run at once and their launches overlap; a launch's "
            "duration covers its share of the rays while the others run")
    if traffic and avg_s > 0:
        # what the memory system actually moved (scaled FETCH + WRITE counters)
        out["measured_traffic_gbps"] = round(traffic / avg_s / 1e9, 2)
        out["measured_frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBPS, 5)
    if prof.get("counters"):
        out["binding"] = dict(prof["counters"])
        sv = prof.get("step_valu")
        if sv and step_ms:
            # VALU issue of all kernels of one step (counter run) over the
            # measured step at the 2.4 GHz peak engine clock: the share of
            # the step every SIMD spends issuing VALU instructions
            out["binding"]["step_valu_busy"] = round(
                sv["valu_cycles_per_simd_per_step"] / (step_ms * 1e-3 * ENGINE_CLOCK_HZ), 3)
            out["binding"]["step_valu_cycles_per_simd"] = sv["valu_cycles_per_simd_per_step"]
    if out["frac"] > 1.0:
        out["note"] = ("frac > 1: the algorithmic bytes (the reference's binary-BVH AoS "
                       "traversal, target_objects = 4) exceed what HBM could deliver in the "
                       "step, because the 4-wide tree visits fewer nodes and its working set "
                       "stays in L2 / Infinity Cache; HBM is not the limit (measured_frac), "
                       "the kernel is bound by issue and latency (binding; step_valu_busy: "
                       "the share of the step the SIMDs spend issuing VALU instructions)")
    return out


def _trace_kernel(st, counts):
    """The wavefront traversal kernel (one launch traces an iteration's
    extension and shadow rays): (name, (ms, launches, rays, algorithmic bytes
    per ray), algorithmic bytes of the step)."""
    bpr_e, bpr_s = _bytes_per_ray(counts, 0), _bytes_per_ray(counts, 1)
    n_rays = st.ext_rays + st.shadow_rays
    step_bytes = st.ext_rays * bpr_e + st.shadow_rays * bpr_s
    return "trace", (st.trace_ms, st.busy_trace_launches, n_rays,
                     step_bytes / max(n_rays, 1)), step_bytes


def _bytes_per_ray(counts, kind):
    n_rays, n_nodes, n_tris_t = (float(x) for x in counts[kind])
    if n_rays == 0:
        return 0.0
    return RAY_BYTES + NODE_BYTES * n_nodes / n_rays + TRI_BYTES * n_tris_t / n_rays


def bench_datagen(a, cfg):
    """Batched datagen: scenes/s of one batched render of every scene vs one
    render call per scene (same scenes, settings and images)."""
    import numpy as np
    from path_amd.datagen import (autotune, datagen_settings, generate_scene,
                                  render_batch, render_batch_device, set_launch_params)
    from path_amd.renderer import Renderer
    n, spp = cfg["n_scenes"], cfg["spp"]
    t0 = time.perf_counter()
    gens = [generate_scene(seed) for seed in range(n)]
    gen_s = time.perf_counter() - t0
    settings = datagen_settings()
    settings.term_prob = "NIters"
    settings.n_iters = cfg["iters"]
    if a.max_paths:
        settings.max_num_samples_per_launch = a.max_paths
    r = Renderer(0)
    if a.pipelines:
        from path_amd.datagen import set_pipelines
        set_pipelines(r, a.pipelines)
    # host BVH builds (thread pool inside the first batched call) + warm-up
    t0 = time.perf_counter()
    render_batch(r, gens, spp, settings)
    first_s = time.perf_counter() - t0
    _progress(f"datagen: first batch {first_s:.1f} s")
    tuned = None
    if not a.no_autotune:
        # tuned on a batch of the same size from other seeds, then applied to
        # the measured batches
        tune_set = [generate_scene(seed) for seed in range(n, 2 * n)]
        tuned = autotune(r, tune_set, spp, settings)
        del tune_set
        _progress(f"datagen: autotuned {tuned}")
    elif a.trace_blocks_per_cu or a.iters_per_sync:
        set_launch_params(r, a.trace_blocks_per_cu, a.iters_per_sync or 1, bool(a.iters_per_sync))
    for _ in range(max(a.warmup - 1, 0)):
        render_batch(r, gens, spp, settings)
    # Timed steps: every step renders a batch of freshly generated scene
    # objects (the same seeds, regenerated outside the timed region), so the
    # step holds what datagen with new scenes costs: flattening, packing,
    # upload, the GPU tree builds of all scenes, the render, and the images
    # and G-buffers left in HBM (one device buffer each, batch order).
    from path_amd.datagen import last_bvh_build, set_bvh_builder
    sync = _sync_fn()
    set_bvh_builder(r, a.gpu_tree)
    dev = render_batch_device(r, [generate_scene(seed) for seed in range(n)], spp, settings)
    sync()
    fresh_s, rays, kernel_ms, gen_step_s = [], 0, 0.0, []
    fresh_mismatch = 0
    for step in range(a.steps):
        t0 = time.perf_counter()
        batch = [generate_scene(seed) for seed in range(n)]
        gen_step_s.append(time.perf_counter() - t0)
        sync()
        t0 = time.perf_counter()
        dev = render_batch_device(r, batch, spp, settings, out=dev)
        sync()
        fresh_s.append(time.perf_counter() - t0)
        st = r.last_stats()
        rays += st.rays
        kernel_ms += st.kernel_ms
        del batch
    tree_ms, tree_levels = last_bvh_build(r)
    from path_amd.datagen import last_batch_phases
    phases = {k: round(v, 2) for k, v in last_batch_phases(r).items()}
    elapsed = sum(fresh_s)
    _progress(f"datagen: {a.steps} new-scene batches in {elapsed:.2f} s")
    value = n * a.steps / elapsed
    total_px = sum(g.resolution ** 2 for g in gens)
    fresh_rgb = dev[0].numpy((total_px, 3)).copy()
    fresh_gb = dev[1].numpy((total_px, 12)).copy()

    # the same scenes kept resident between batches (no packing, upload or
    # tree builds): a re-render rate, reported beside the headline
    set_bvh_builder(r, False)
    render_batch(r, gens, spp, settings)
    dev = render_batch_device(r, gens, spp, settings, out=dev)
    sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dev = render_batch_device(r, gens, spp, settings, out=dev)
    sync()
    resident_value = n * a.steps / (time.perf_counter() - t0)
    t1 = time.perf_counter()
    for _ in range(a.steps):
        images, gbufs, _ = render_batch(r, gens, spp, settings)
    host_value = n * a.steps / (time.perf_counter() - t1)
    dev_rgb = dev[0].numpy((total_px, 3))
    dev_gb = dev[1].numpy((total_px, 12))
    dev_mismatch = 0
    for i, g in enumerate(gens):
        o, px = int(dev[2][i]), g.resolution ** 2
        want_rgb = images[i].reshape(-1, 3).view(np.uint32)
        want_gb = gbufs[i].reshape(-1, 12).view(np.uint32)
        dev_mismatch += not (np.array_equal(dev_rgb[o:o + px].view(np.uint32), want_rgb) and
                             np.array_equal(dev_gb[o:o + px].view(np.uint32), want_gb))
        # fresh batches with GPU-built trees against host trees and host output
        fresh_mismatch += not (np.array_equal(fresh_rgb[o:o + px].view(np.uint32), want_rgb) and
                               np.array_equal(fresh_gb[o:o + px].view(np.uint32), want_gb))
    assert dev_mismatch == 0, f"{dev_mismatch} device-output images differ from host output"
    assert fresh_mismatch == 0, f"{fresh_mismatch} fresh-batch images differ"

    # new scenes with host-built SAH trees (16 host threads) instead
    host_tree_s = []
    for _ in range(2):
        batch = [generate_scene(seed) for seed in range(n)]
        sync()
        t0 = time.perf_counter()
        render_batch_device(r, batch, spp, settings, out=dev)
        sync()
        host_tree_s.append(time.perf_counter() - t0)
        del batch

    _progress("datagen: resident and host-tree batches done")
    # per-scene launches: the same scenes, one render call each
    single = settings.copy()
    set_launch_params(r, 6, 1, False)
    for g in gens[:4]:
        r.render(g.scene, g.film_to_world, g.resolution, g.resolution, spp, single)
    from path_amd.renderer import DeviceBuffer
    one = DeviceBuffer(r, max(g.resolution for g in gens) ** 2 * 12)
    sync()
    t0 = time.perf_counter()
    for g in gens:  # each call uploads its scene, as one Renderer::render per scene does
        r.render(g.scene, g.film_to_world, g.resolution, g.resolution, spp, single,
                 output="DeviceFloatRGB", device_out=one)
    sync()
    per_scene_s = time.perf_counter() - t0
    per_scene_value = n / per_scene_s
    mismatched = 0
    for i, g in enumerate(gens):
        img, _ = r.render(g.scene, g.film_to_world, g.resolution, g.resolution, spp, single)
        if not np.array_equal(img.view(np.uint32), images[i].view(np.uint32)):
            mismatched += 1

    # profiled batched render for the per-kernel split (tuned or default grid)
    set_launch_params(r, tuned["trace_blocks_per_cu"] if tuned else 0, 1, False)
    r.set_profiling(True)
    render_batch(r, gens, spp, settings, gbuffer=False)
    pst = r.last_stats()
    r.set_profiling(False)

    out = {
        "metric": "scenes/s (batched datagen, 1 GPU)",
        "value": round(value, 2),
        "unit": "scenes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "scenes": n, "spp": spp,
                   "max_depth": cfg["iters"] + 1, "pixels": total_px,
                   "triangles_min": min(g.scene.counts()[0] for g in gens),
                   "triangles_max": max(g.scene.counts()[0] for g in gens),
                   "integrator": "Streaming", "sampler": "Sobel", "gbuffer": True},
        "step": "a batch of newly created scene objects every step (generation outside the "
                "timed region, reported below): flattening, packing, upload, GPU tree builds "
                "of all scenes, render; images and G-buffers left in HBM",
        "msamples_per_s": round(total_px * spp * a.steps / elapsed / 1e6, 2),
        "mrays_per_s": round(rays / elapsed / 1e6, 2),
        "gpu_ms_per_batch": round(kernel_ms / a.steps, 2),
        "gpu_tree_builder": a.gpu_tree,
        "gpu_tree_build_ms": round(tree_ms, 2),
        "gpu_tree_levels": tree_levels,
        "fresh_batch_phases_ms": phases,
        "images_differing_from_host_trees": fresh_mismatch,
        "per_scene_launches": {"value": round(per_scene_value, 2), "unit": "scenes/s",
                               "seconds": round(per_scene_s, 3),
                               "images_differing_from_batched": mismatched,
                               "note": "one render call per scene; each call uploads its scene, "
                                       "whose host SAH tree was built once before timing"},
        "speedup_vs_per_scene": round(value / per_scene_value, 2),
        "resident_batch_scenes_per_s": round(resident_value, 2),
        "resident_batch_note": "the same scene objects again: packing, upload and tree builds "
                               "skipped (a re-render, not datagen)",
        "host_output_resident_scenes_per_s": round(host_value, 2),
        "new_scenes_host_sah_trees_scenes_per_s": round(n / min(host_tree_s), 2),
        "autotuned": tuned,
        "autotune_set": (f"{n} generated scenes of seeds {n}..{2 * n - 1}" if tuned else None),
        "host_seconds": {"generate_per_batch": round(min(gen_step_s), 3),
                         "first_batch_incl_host_bvh_builds": round(first_s, 3),
                         "generate_first": round(gen_s, 3)},
    }

    counts = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        tab = O.sobol_table(O.joe_kuo_directions(settings.max_sample_size),
                            settings.max_sample_size, spp)
        done, secs, counts = 0, 0.0, np.zeros((2, 3), np.uint64)
        while done < n and secs < 12.0:
            g = gens[done]
            _, info = O.render(g.scene.flat(), g.film_to_world, settings, spp, g.resolution,
                               g.resolution, tab, 32, 0)
            secs += info["seconds"]
            counts += info["counts"]
            done += 1
        out["cpu_baseline"] = {
            "value": round(done / secs, 4), "unit": "scenes/s", "cores": O.max_threads(),
            "kind": "port",
            "sample": f"first {done} scenes of the batch, full resolution and spp, FP32 CPU "
                      f"checker, {secs:.1f} s"}
        out["cpu_baseline"].update(host_cpu())
    roof = None
    if counts is not None and pst.busy_trace_launches:
        name, vals, step_bytes = _trace_kernel(pst, counts)
        roof = _roofline(name, *vals, config="c3", step_bytes=step_bytes,
                         step_ms=elapsed / a.steps * 1e3, pipelines=pst.pipelines)
    out["roofline"] = roof
    return out


def bench_adjoint(a, cfg):
    """Gradient evaluations of an L2 loss: forward render + path replay."""
    import numpy as np
    from path_amd.diff import DiffRenderer
    from path_amd.renderer import Renderer
    size, spp = cfg["size"], cfg["spp"]
    with tempfile.TemporaryDirectory() as d:
        scene, ftw = build_scene(cfg, d)
    settings = make_settings(cfg)
    r = Renderer(0)
    dr = DiffRenderer(r, scene, ftw, size, size, spp, settings)
    alb0, em0 = dr.parameters()
    target = dr.render().astype(np.float64)
    rng = np.random.default_rng(11)
    alb = np.clip(alb0.numpy() * rng.uniform(0.7, 1.3, alb0.shape), 0.02, 0.98).astype(np.float32)
    em = (em0.numpy() * 0.8).astype(np.float32)
    dr.set_parameters(alb, em)

    def gradient():
        img = dr.render()
        d_img = (2.0 * (img.astype(np.float64) - target)).astype(np.float32)
        return img, d_img, dr.adjoint(d_img)

    # the gradient step in HBM: render into a device buffer, dL/dI = 2 (I - T)
    # on the GPU, adjoint from it; only the 48 gradient floats come back
    from path_amd.renderer import DeviceBuffer
    n_f = size * size * 3
    dev_img, dev_tgt, dev_d = (DeviceBuffer(r, n_f * 4) for _ in range(3))
    dev_tgt.write(target.astype(np.float32))

    def gradient_device():
        r.render(scene, ftw, size, size, spp, settings, output="DeviceFloatRGB",
                 device_out=dev_img)
        fwd = r.last_stats().kernel_ms
        r.l2_loss_gradient(dev_img, dev_tgt, n_f, dev_d)
        ga, ge, _ = r.adjoint(scene, ftw, size, size, spp, dev_d, settings)
        return fwd, ga, ge

    for _ in range(a.warmup):
        gradient()
        gradient_device()
    sync = _sync_fn()
    sync()
    t0 = time.perf_counter()
    fwd_ms = adj_ms = 0.0
    replays = 1
    for _ in range(a.steps):
        f_ms, ga, ge = gradient_device()
        fwd_ms += f_ms
        adj_ms += r.last_stats().kernel_ms
        replays = r.last_stats().replays
    sync()
    elapsed = time.perf_counter() - t0
    samples = size * size * spp
    # the same step through host arrays (DiffRenderer / the autograd Function)
    t1 = time.perf_counter()
    for _ in range(a.steps):
        img, d_img, (ga_h, ge_h) = gradient()
    host_elapsed = time.perf_counter() - t1
    assert np.array_equal(ga_h, ga) and np.array_equal(ge_h, ge), \
        "device and host gradient steps differ"

    # agreement: central differences (h = 1e-3) of the loss through the GPU's
    # FP32 renders and, with the CPU leg, through FP64 renders of the same
    # estimator
    def central(loss):
        h = 1e-3
        fds = []
        for kind, m, c in checks:
            base = alb if kind == "a" else em
            vals, xs = [], []
            for sgn in (+1, -1):
                p_ = base.copy()
                p_[m, c] = np.float32(base[m, c] + sgn * h)
                xs.append(float(p_[m, c]))
                vals.append(loss(p_, em) if kind == "a" else loss(alb, p_))
            fds.append((vals[0] - vals[1]) / (xs[0] - xs[1]))
        dr.set_parameters(alb, em)
        return np.array(fds)

    def rel_err(g, ref):
        return float((np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())).max())

    def loss32(al, e):
        dr.set_parameters(al, e)
        return float(((dr.render().astype(np.float64) - target) ** 2).sum())
    checks = [("a", m, c) for m in range(8) for c in range(3)] + [("e", 7, c) for c in range(3)]
    grads = np.array([float((ga if k == "a" else ge)[m, c]) for k, m, c in checks])
    fd_err = rel_err(grads, central(loss32))

    out = {
        "metric": "Msamples/s of render + path-replay adjoint (1 GPU)",
        "value": round(samples * a.steps / elapsed / 1e6, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "width": size, "height": size, "spp": spp,
                   "max_depth": cfg["iters"] + 1, "parameters": 27,
                   "loss": "sum((I - T)^2)"},
        "forward_ms": round(fwd_ms / a.steps, 3),
        "adjoint_ms": round(adj_ms / a.steps, 3),
        "max_rel_err_vs_central_differences_fp32_renders": fd_err,
        "output": "image, its loss gradient and the adjoint input stay in HBM",
        "host_arrays_msamples_per_s": round(samples * a.steps / host_elapsed / 1e6, 3),
    }
    counts = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        tab = O.sobol_table(O.joe_kuo_directions(settings.max_sample_size),
                            settings.max_sample_size, spp)
        ra, re, _, secs = O.adjoint(scene.flat(), ftw, settings, spp, size, size, tab,
                                    d_img.astype(np.float64))
        g = np.concatenate([ga.ravel(), ge[7]])
        ref = np.concatenate([ra.ravel(), re[7]])
        out["max_rel_err_vs_fp64_adjoint"] = rel_err(g, ref)

        # FD of the functional the adjoint differentiates, sum(d_img * I)
        w = d_img.astype(np.float64)

        def loss64(al, e):
            dr.set_parameters(al, e)
            im, _ = O.render(scene.flat(), ftw, settings, spp, size, size, tab, 64, 0)
            return float((w * im).sum())
        out["max_rel_err_vs_central_differences_fp64_renders"] = rel_err(grads, central(loss64))
        out["cpu_baseline"] = {
            "value": round(samples / secs / 1e6, 4), "unit": "Msamples/s",
            "cores": O.max_threads(), "kind": "port",
            "sample": f"whole workload ({size}x{size} x {spp} spp), FP64 forward-mode adjoint "
                      f"of the CPU checker (no forward render), {secs:.2f} s"}
        out["cpu_baseline"].update(host_cpu())
        _, info = O.render(scene.flat(), ftw, settings, spp, size, size, tab, 32, 0)
        counts = info["counts"]
    roof = None
    if counts is not None:
        n_rays = float(counts[0][0] + counts[1][0])
        bpr = (RAY_BYTES * n_rays + NODE_BYTES * float(counts[0][1] + counts[1][1]) +
               TRI_BYTES * float(counts[0][2] + counts[1][2])) / max(n_rays, 1.0)
        # every path is replayed `replays` times (once when lengths are bounded)
        roof = _roofline("adjoint", adj_ms / a.steps, 1, replays * n_rays, bpr, config="c4",
                         step_bytes=replays * n_rays * bpr, step_ms=elapsed / a.steps * 1e3)
        roof["replays_per_path"] = replays
        roof["note"] = ("36-triangle scene: node and triangle reads are cache hits; the kernel "
                        "is bound by traversal/shading latency, not HBM")
    out["roofline"] = roof
    return out


def bench_mlp(a, cfg):
    """Radiance MLP: datagen -> training -> held-out time-to-RMSE of each
    MLP-assisted estimator against the plain tracer.

    Estimators (all on the same held-out scene and 4096-spp FP64-summed
    ground truth):
      plain            the tracer itself, spp samples per pixel
      vertex2          per sample: camera ray, light sample of the first
                       vertex, one bounce; the model (trained on reflected
                       radiance) stands in for the light reflected at the
                       second vertex (one query per sample)
      first_vertex     the tracer terminated after the first vertex (emission
                       plus light sample, spp samples) plus one query per
                       pixel of a model of the indirect light there
      model_only       emission seen plus one query per pixel of the
                       reflected-radiance model: no sampling at all
    each beside the same estimator with a zero network."""
    import numpy as np
    from path_amd.datagen import datagen_settings, generate_scene
    from path_amd.mlp import RadianceMLP, param_count
    from path_amd.radiance_cache import rel_rmse, training_sets
    from path_amd.renderer import Renderer
    r = Renderer(0)
    st = datagen_settings()
    st.term_prob = "NIters"
    st.n_iters = cfg["iters"]
    rows = cfg["rows"] if not a.max_paths else a.max_paths
    t0 = time.perf_counter()
    feats, refl, ind, seed = [], [], [], 0
    while sum(x.shape[0] for x in feats) < rows:
        f, tr, ti = training_sets(r, range(seed, seed + 1024), samples_per=cfg["spp"],
                                  settings=st)
        feats.append(f)
        refl.append(tr)
        ind.append(ti)
        seed += 1024
        print(f"datagen: {seed} scenes, {sum(x.shape[0] for x in feats)} rows, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    f = np.concatenate(feats)[:rows]
    tr = np.concatenate(refl)[:rows]
    ti = np.concatenate(ind)[:rows]
    del feats, refl, ind
    datagen_s = time.perf_counter() - t0
    batch = 65536
    steps = 2 * rows // batch

    def train(targets, seed_w):
        m = RadianceMLP(0, seed=seed_w)
        m.set_dataset(f, targets)
        losses, secs = m.train(steps, batch, 1e-3, seed=seed_w + 1)
        print(f"trained {steps} steps in {secs:.1f} s", file=sys.stderr, flush=True)
        return m, losses, secs
    m, losses, train_s = train(tr, 1)
    mi, losses_i, train_i_s = train(ti, 3)
    del f, tr, ti
    train_tflops = 3 * 2 * MLP_MACS * batch * steps / train_s / 1e12

    # held-out scene at 512x512
    g = generate_scene(10 ** 9 + 7)
    size = cfg["size"]
    stm = st.copy()
    stm.kernel_approach = "Streaming"
    r.render(g.scene, g.film_to_world, size, size, cfg["gt_spp"], stm)
    gt = r.last_render_fp64(size, size)  # per-sample radiance summed in FP64
    indirect_share = None

    def curve(render, spps):
        out = []
        for spp in spps:
            render(spp)  # warm (buffers, Sobol table)
            sync = _sync_fn()
            sync()
            t1 = time.perf_counter()
            img = render(spp)
            sync()
            out.append({"spp": spp, "seconds": round(time.perf_counter() - t1, 5),
                        "gpu_ms": round(r.last_stats().kernel_ms, 3),
                        "rel_rmse": round(rel_rmse(img, gt), 5)})
        return out
    zero = RadianceMLP(0, seed=1)
    zero.set_weights(np.zeros(param_count(), np.float32))
    fv = lambda net: (lambda spp: net.render_first_vertex(r, g.scene, g.film_to_world, size,
                                                          size, spp, stm)[0])
    curves = {
        "plain": curve(lambda spp: r.render(g.scene, g.film_to_world, size, size, spp, stm)[0],
                       [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]),
        "vertex2": curve(lambda spp: m.render(r, g.scene, g.film_to_world, size, size, spp,
                                              stm)[0], [1, 2, 4, 8, 16, 32, 64]),
        "vertex2_zero_model": curve(lambda spp: zero.render(r, g.scene, g.film_to_world, size,
                                                            size, spp, stm)[0], [1, 4, 16, 64]),
        "first_vertex": curve(fv(mi), [1, 2, 4, 8, 16, 32, 64, 128]),
        "first_vertex_zero_model": curve(fv(zero), [4, 16, 64, 128]),
        "model_only": curve(fv(m), [0]),
    }
    # share of indirect light in the ground truth: the first-vertex render
    # at high spp against the full image
    direct = fv(zero)(1024)
    indirect_share = round(float(1.0 - direct.astype(np.float64).mean() / gt.mean()), 4)

    def time_to(cv, target):
        ok = [c["seconds"] for c in cv if c["rel_rmse"] <= target]
        return min(ok) if ok else None
    tp = time_to(curves["plain"], 0.01)
    per = {k: time_to(v, 0.01) for k, v in curves.items()
           if k in ("vertex2", "first_vertex", "model_only")}
    reached = [t for t in per.values() if t]
    tm = min(reached) if reached else None
    # inference throughput of the network alone (forward of 2^20 queries)
    q = np.random.default_rng(0).uniform(-1, 1, (1 << 20, 9)).astype(np.float32)
    m.forward(q)
    sync = _sync_fn()
    t1 = time.perf_counter()
    m.forward(q)
    sync()
    fwd_s = time.perf_counter() - t1
    out = {
        "metric": "MLP-assisted vs plain time-to-1%-RMSE (held-out scene, 1 GPU)",
        "value": round(tp / tm, 3) if (tp and tm) else None,
        "unit": "x (plain time / fastest MLP-assisted time)",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(train_s / steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (MLP, fp32 accumulate) / f32 (tracer)", "data": "synthetic",
        "config": {"workload": cfg["workload"], "training_rows": int(rows),
                   "training_scenes": seed, "datagen_spp": cfg["spp"], "batch": batch,
                   "epochs": 2, "heldout_seed": 10 ** 9 + 7, "width": size, "height": size,
                   "ground_truth_spp": cfg["gt_spp"], "max_depth": cfg["iters"] + 1},
        "time_to_1pct_plain_s": tp,
        "time_to_1pct_s": per,
        "best_rel_rmse": {k: min(c["rel_rmse"] for c in v) for k, v in curves.items()},
        "indirect_share_of_ground_truth": indirect_share,
        "curves": curves,
        "datagen_s": round(datagen_s, 2), "train_s": round(train_s, 2),
        "train_indirect_s": round(train_i_s, 2),
        "train_loss_first_last": [float(losses[:20].mean()), float(losses[-20:].mean())],
        "train_indirect_loss_first_last": [float(losses_i[:20].mean()),
                                           float(losses_i[-20:].mean())],
        "train_tflops": round(train_tflops, 2),
        "inference_mqueries_per_s": round((1 << 20) / fwd_s / 1e6, 2),
    }
    out.update(mlp_forward_timing(m))
    print(json.dumps(out))
    return 0


def mlp_forward_timing(m, iters=20):
    """Roofline of the dominant inference kernel (k_mlp_forward over 2^20
    queries) and the layer-by-layer GEMM (training forward) beside it."""
    rows_t = 1 << 20
    gemm_ms, fwd_ms = m.time_forward(rows_t, iters)
    gemm_tflops = 2.0 * rows_t * 256 * 256 / (gemm_ms * 1e-3) / 1e12
    fwd_tflops = 2.0 * rows_t * MLP_MACS / (fwd_ms * 1e-3) / 1e12
    return {
        "forward_ms_per_1M_queries": round(fwd_ms, 4),
        "layer_gemm": {"ms": round(gemm_ms, 4), "tflops": round(gemm_tflops, 2),
                       "shape": "2^20 x 256 x 256, BF16 in, FP32 accumulate"},
        "roofline": {"bound": "mfma", "achieved": round(fwd_tflops, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(fwd_tflops / 2500.0, 5),
                     "traffic": None, "kernel": "k_mlp_forward",
                     "avg_launch_ms": round(fwd_ms, 4),
                     "flops_per_launch": 2 * rows_t * MLP_MACS,
                     "note": "whole inference forward of 2^20 queries in one kernel "
                             "(encoding, 6 x 256 hidden, output; activations in LDS), "
                             "valid (unpadded) MACs, HIP events on the MLP stream"},
    }


def bench_mlp_forward(a):
    """Inference forward alone (random weights): the c5 network's dominant
    kernel, for profiling."""
    from path_amd.mlp import RadianceMLP
    m = RadianceMLP(0, seed=1)
    out = {"metric": "radiance-MLP inference forward (1 GPU)", "unit": "M queries/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 (fp32 accumulate)",
           "data": "synthetic", "config": {"workload": "radiance_mlp_6x256_forward_2^20_queries"}}
    out.update(mlp_forward_timing(m, a.steps))
    out["value"] = round((1 << 20) / (out["forward_ms_per_1M_queries"] * 1e-3) / 1e6, 2)
    print(json.dumps(out))
    return 0


def bench_mlp_train(a):
    """Training alone on a synthetic device-resident set (2^20 rows of random
    features and smooth targets): Adam steps of 65,536 rows, the C5 network's
    training kernels, for profiling."""
    import numpy as np
    from path_amd.mlp import RadianceMLP
    rng = np.random.default_rng(0)
    rows = 1 << 20
    f = rng.uniform(-1, 1, (rows, 9)).astype(np.float32)
    t = np.stack([np.sin(f[:, 0] * 2) + 1.5, np.cos(f[:, 1]) + 1.0, f[:, 2] ** 2 + 0.5],
                 1).astype(np.float32)
    m = RadianceMLP(0, seed=1)
    m.set_dataset(f, t)
    batch = 65536
    m.train(max(a.warmup, 1), batch, 1e-3, seed=1, record_losses=False)
    losses, secs = m.train(a.steps, batch, 1e-3, seed=2)
    tflops = 3 * 2 * MLP_MACS * batch * a.steps / secs / 1e12
    out = {"metric": "radiance-MLP training (1 GPU)", "value": round(tflops, 2),
           "unit": "TFLOP/s (3 x forward MACs x 2, valid dims)", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(secs / a.steps * 1e3, 4),
           "higher_is_better": True, "dtype": "bf16 (fp32 accumulate, fp32 Adam)",
           "data": "synthetic",
           "config": {"workload": "radiance_mlp_6x256_train_65536_rows_per_step"},
           "loss_first_last": [float(losses[0]), float(losses[-1])]}
    print(json.dumps(out))
    return 0


def _progress(msg):
    """Progress of the long legs on stderr (the JSON line stays alone on stdout)."""
    print(msg, file=sys.stderr, flush=True)


def _sync_fn():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.synchronize
    except Exception:  # torch is optional plumbing here
        pass
    return lambda: None


def build_scene(cfg, d):
    from path_amd import scenes
    from path_amd.renderer import Scene
    if cfg["level"] is None:
        path = scenes.cornell_box(d)
    else:
        path = scenes.cornell_icosphere(d, level=cfg["level"], seed=0)
    return Scene.load(path, 1.0)


def make_settings(cfg):
    from path_amd.renderer import Settings
    s = Settings()
    s.kernel_approach = cfg["approach"]
    s.term_prob = "NIters"
    s.n_iters = cfg["iters"]
    return s


def host_cpu():
    """CPU model string, usable cores and OMP_NUM_THREADS of this host."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"cpu_model": model, "nproc": nproc,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(scene, ftw, settings, cfg, target_seconds=12.0):
    """FP32 CPU checker (same estimator and samples) on a band of rows,
    all host threads. Returns (baseline dict, per-kind traversal counts,
    (band image, first pixel, end pixel))."""
    import numpy as np
    from oracle import oracle as O
    size, spp = cfg["size"], cfg["spp"]
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    flat = scene.flat()
    threads = O.max_threads()
    # calibrate on a few rows, then size the band for ~target_seconds
    mid = size // 2
    rows = 2
    begin = (mid - rows // 2) * size
    _, info = O.render(flat, ftw, settings, spp, size, size, tab, 32, 0, begin, begin + rows * size)
    per_row = max(info["seconds"] / rows, 1e-4)
    rows = int(min(size, max(2, target_seconds / per_row)))
    begin = max(0, (mid - rows // 2)) * size
    end = min(size * size, begin + rows * size)
    img, info = O.render(flat, ftw, settings, spp, size, size, tab, 32, 0, begin, end)
    n_px = end - begin
    secs = info["seconds"]
    base = {
        "value": round(n_px * spp / secs / 1e6, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_px} pixels (rows {begin // size}-{end // size - 1}) x {spp} spp of the "
                  f"same render, FP32 CPU checker, {secs:.1f} s",
    }
    base.update(host_cpu())
    return base, info["counts"], (img, begin, end)


def band_parity(gpu_img, band, scene, ftw, settings, cfg, fp64_rows=16):
    """The GPU image against the CPU checker on the baseline's band: FP32
    pixels compared bit for bit, and the relative RMSE against the FP64
    checker (same samples) on a sub-band of fp64_rows rows."""
    import numpy as np
    from oracle import oracle as O
    ref, begin, end = band
    size, spp = cfg["size"], cfg["spp"]
    g = gpu_img.reshape(-1, 3)[begin:end]
    c = ref.reshape(-1, 3)[begin:end]
    equal = np.all(g.view(np.uint32) == c.view(np.uint32), axis=1)
    mid = (begin + end) // 2 // size
    b64 = max(begin, (mid - fp64_rows // 2) * size)
    e64 = min(end, b64 + fp64_rows * size)
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    ref64, _ = O.render(scene.flat(), ftw, settings, spp, size, size, tab, 64, 0, b64, e64)
    ind = O.indep_render(scene.flat(), ftw, settings, spp, size, size, tab, b64, e64)
    a64 = gpu_img.reshape(-1, 3)[b64:e64].astype(np.float64)
    r64 = ref64.reshape(-1, 3)[b64:e64]
    i64 = ind.reshape(-1, 3)[b64:e64]

    def rel_rmse(x, ref):
        # RMSE of per-pixel differences relative to the mean of the reference
        return float(np.sqrt(np.mean((x - ref) ** 2)) / max(float(ref.mean()), 1e-12))
    return {"pixels_compared": int(end - begin), "pixels_bit_equal": int(equal.sum()),
            "rows": f"{begin // size}-{end // size - 1}",
            "rel_rmse_vs_fp64": float(f"{rel_rmse(a64, r64):.3e}"),
            "rel_rmse_vs_independent_fp64": float(f"{rel_rmse(a64, i64):.3e}"),
            "fp64_checkers_rel_rmse": float(f"{rel_rmse(r64, i64):.3e}"),
            "fp64_rows": f"{b64 // size}-{e64 // size - 1}",
            "checker": "FP32 CPU checker on the same Sobol samples (bit for bit); FP64 checker "
                       "and the independent FP64 restatement of the reference (relative RMSE)"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--integrator", choices=["Streaming", "MegaKernel"], default=None,
                    help="override the configuration's integrator")
    ap.add_argument("--max-paths", type=int, default=0,
                    help="wavefront path-pool size (max_num_samples_per_launch)")
    ap.add_argument("--trace-blocks-per-cu", type=int, default=0,
                    help="wavefront traversal grid (blocks of 256 per CU)")
    ap.add_argument("--iters-per-sync", type=int, default=0,
                    help="wavefront iterations enqueued per host check (hipGraph replay)")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="default run: skip the short C1 / C4 / C3 runs in other_workloads")
    ap.add_argument("--gpu-tree", choices=["ploc", "lbvh"], default="ploc",
                    help="GPU tree builder of the batched datagen workload (c3)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="c3: keep the default launch parameters")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 256 MiB device-to-device copies (scale for the "
                         "profiler's FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--spatial-splits", action="store_true",
                    help="build the scene's tree with spatial splits (use_spatial_splits)")
    ap.add_argument("--overlap-threshold", type=float, default=0.0,
                    help="with --spatial-splits: overlap_threshold (default 1e-5)")
    ap.add_argument("--pipelines", type=int, default=0,
                    help="wavefront pipelines per render (default 3)")
    ap.add_argument("--mlp-forward", action="store_true",
                    help="only time the radiance MLP's fused inference forward (2^20 queries, "
                         "--steps launches)")
    ap.add_argument("--mlp-train", action="store_true",
                    help="only time radiance-MLP training (Adam steps of 65,536 rows on a "
                         "synthetic resident set, --steps steps)")
    a = ap.parse_args(argv)
    if a.mlp_forward:
        return bench_mlp_forward(a)
    if a.mlp_train:
        return bench_mlp_train(a)
    if a.copy_calibration:
        import torch
        x = torch.ones(256 * 1024 * 1024 // 4, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes_per_launch": x.numel() * 4, "launches": 3}))
        return 0
    if a.gpus != 1:
        print("bench.py measures one GPU only: the path shards by sample range with a single "
              "image sum-reduce, which is not built here; run with --gpus 1", file=sys.stderr)
        return 2

    cfg = dict(CONFIGS[a.config])
    if a.config == "c3":
        print(json.dumps(bench_datagen(a, cfg)))
        return 0
    if a.config == "c4":
        print(json.dumps(bench_adjoint(a, cfg)))
        return 0
    if a.config == "c5":
        return bench_mlp(a, cfg)
    out = bench_render(a, cfg)
    if a.config == "c2" and not a.no_other_workloads:
        out["other_workloads"] = other_workloads(a)
    print(json.dumps(out))
    return 0


def other_workloads(a):
    """Short runs of the other workloads in the same process, so the default
    run measures them too: the Cornell-box mega-kernel (C1), the adjoint
    gradient step (C4) and batched datagen with new scenes every batch (C3),
    each with its own parity checks; their CPU baselines and the datagen
    autotuner are left to the --config runs."""
    import copy
    res = {}
    for name, steps in (("c1", 20), ("c4", 5), ("c3", 3)):
        b = copy.copy(a)
        b.config, b.steps, b.warmup = name, steps, 2
        b.no_cpu_baseline = True
        b.no_autotune = True
        cfg = dict(CONFIGS[name])
        _progress(f"other workloads: {name}")
        o = (bench_render(b, cfg) if name == "c1" else
             bench_adjoint(b, cfg) if name == "c4" else bench_datagen(b, cfg))
        keep = ["metric", "value", "unit", "steps", "warmup", "ms_per_step", "config",
                "mrays_per_s", "msamples_per_s", "speedup_vs_per_scene",
                "images_differing_from_host_trees", "per_scene_launches",
                "max_rel_err_vs_central_differences_fp32_renders", "forward_ms", "adjoint_ms",
                "resident_batch_scenes_per_s", "gpu_tree_builder", "gpu_tree_build_ms",
                "fresh_batch_phases_ms"]
        res[name] = {k: o[k] for k in keep if k in o}
    return res


def bench_render(a, cfg):
    """One SquareImage render of the configuration per step (C1, C2)."""
    from path_amd.renderer import Renderer
    if a.integrator:
        cfg["approach"] = a.integrator
    with tempfile.TemporaryDirectory() as d:
        scene, ftw = build_scene(cfg, d)
    settings = make_settings(cfg)
    if a.max_paths:
        settings.max_num_samples_per_launch = a.max_paths
    if getattr(a, "spatial_splits", False):
        settings.use_spatial_splits = 1
        if a.overlap_threshold:
            settings.overlap_threshold = a.overlap_threshold
    n_tris = scene.counts()[0]
    r = Renderer(0)
    if a.pipelines:
        from path_amd.datagen import set_pipelines
        set_pipelines(r, a.pipelines)
    if a.trace_blocks_per_cu or a.iters_per_sync:
        from path_amd.datagen import set_launch_params
        set_launch_params(r, a.trace_blocks_per_cu, a.iters_per_sync or 1,
                          bool(a.iters_per_sync))
    size, spp = cfg["size"], cfg["spp"]
    # the image stays in HBM (a device buffer of the renderer), so the timed
    # steps hold no PCIe transfer; the host-output rate is reported beside it
    import numpy as np
    from path_amd.renderer import DeviceBuffer
    dev_img = DeviceBuffer(r, size * size * 12)

    def step():
        r.render(scene, ftw, size, size, spp, settings, output="DeviceFloatRGB",
                 device_out=dev_img)
    for _ in range(a.warmup):
        step()
    sync = _sync_fn()
    sync()
    t0 = time.perf_counter()
    rays = 0
    for _ in range(a.steps):
        step()
        rays += r.last_stats().rays
    sync()
    elapsed = time.perf_counter() - t0
    samples = size * size * spp * a.steps
    value = samples / elapsed / 1e6
    host_img = r.render(scene, ftw, size, size, spp, settings)[0]
    assert np.array_equal(host_img, dev_img.numpy((size, size, 3))), \
        "device and host outputs differ"
    t1 = time.perf_counter()
    for _ in range(a.steps):
        r.render(scene, ftw, size, size, spp, settings)
    host_rate = samples / (time.perf_counter() - t1) / 1e6

    # one profiled render: per-kernel durations with HIP events on the
    # renderer's stream
    r.set_profiling(True)
    r.render(scene, ftw, size, size, spp, settings)
    st = r.last_stats()
    r.set_profiling(False)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "output": "float RGB image left in HBM (device buffer)",
        "host_output_msamples_per_s": round(host_rate, 3),
        "config": {"workload": cfg["workload"], "width": size, "height": size, "spp": spp,
                   "max_depth": cfg["iters"] + 1, "triangles": n_tris,
                   "integrator": cfg["approach"], "sampler": "Sobel",
                   "spatial_splits": bool(settings.use_spatial_splits),
                   "tree_references": int(len(scene.bvh()["perm"]))},
        "mrays_per_s": round(rays / elapsed / 1e6, 2),
    }

    counts = None
    if not a.no_cpu_baseline:
        base, counts, band = cpu_baseline(scene, ftw, settings, cfg)
        out["cpu_baseline"] = base
        out["parity"] = band_parity(host_img, band, scene, ftw, settings, cfg)
        assert out["parity"]["pixels_bit_equal"] == out["parity"]["pixels_compared"], \
            f"GPU image differs from the CPU checker: {out['parity']}"

    # roofline of the dominant kernel
    roof = None
    if counts is not None:
        step_ms = elapsed / a.steps * 1e3
        if cfg["approach"] == "Streaming" and st.busy_trace_launches:
            name, vals, step_bytes = _trace_kernel(st, counts)
            kernels = {name: vals}
        else:
            # mega-kernel: every ray of the render
            tot_rays = float(counts[0][0] + counts[1][0])
            bpr = (RAY_BYTES * tot_rays + NODE_BYTES * float(counts[0][1] + counts[1][1]) +
                   TRI_BYTES * float(counts[0][2] + counts[1][2])) / max(tot_rays, 1.0)
            kernels = {"megakernel": (st.mega_ms, st.mega_launches, st.rays, bpr)}
            step_bytes = st.rays * bpr
        name, vals = max(kernels.items(), key=lambda kv: kv[1][0])
        roof = _roofline(name, *vals, config=a.config, step_bytes=step_bytes, step_ms=step_ms,
                         pipelines=st.pipelines if cfg["approach"] == "Streaming" else 1)
    out["roofline"] = roof
    return out




if __name__ == "__main__":
    sys.exit(main())
