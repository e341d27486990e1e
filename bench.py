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

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1]
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
RAY_BYTES = 24 + 16          # ray read + hit write (reference AoS sizes)
NODE_BYTES = 32              # bvh::Node
TRI_BYTES = 36               # intersect::Triangle

CONFIGS = {
    "c2": dict(workload="cornell_icosphere_l6_512x512_64spp_depth8_wavefront",
               level=6, size=512, spp=64, iters=7, approach="Streaming"),
    "c1": dict(workload="cornell_box_36tri_256x256_16spp_depth5_megakernel",
               level=None, size=256, spp=16, iters=4, approach="MegaKernel"),
}


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


def cpu_baseline(scene, ftw, settings, cfg, target_seconds=12.0):
    """FP32 CPU checker (same estimator and samples) on a band of rows,
    all host threads. Returns (baseline dict, per-kind traversal counts)."""
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
    _, info = O.render(flat, ftw, settings, spp, size, size, tab, 32, 0, begin, end)
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
    return base, info["counts"]


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
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print("bench.py measures one GPU only: the path shards by sample range with a single "
              "image sum-reduce, which is not built here; run with --gpus 1", file=sys.stderr)
        return 2

    from path_amd.renderer import Renderer
    cfg = dict(CONFIGS[a.config])
    if a.integrator:
        cfg["approach"] = a.integrator
    with tempfile.TemporaryDirectory() as d:
        scene, ftw = build_scene(cfg, d)
    settings = make_settings(cfg)
    if a.max_paths:
        settings.max_num_samples_per_launch = a.max_paths
    n_tris = scene.counts()[0]
    r = Renderer(0)
    size, spp = cfg["size"], cfg["spp"]
    for _ in range(a.warmup):
        r.render(scene, ftw, size, size, spp, settings)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is optional plumbing here
        sync = lambda: None
    sync()
    t0 = time.perf_counter()
    rays = 0
    for _ in range(a.steps):
        r.render(scene, ftw, size, size, spp, settings)
        rays += r.last_stats().rays
    sync()
    elapsed = time.perf_counter() - t0
    samples = size * size * spp * a.steps
    value = samples / elapsed / 1e6

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
        "config": {"workload": cfg["workload"], "width": size, "height": size, "spp": spp,
                   "max_depth": cfg["iters"] + 1, "triangles": n_tris,
                   "integrator": cfg["approach"], "sampler": "Sobel"},
        "mrays_per_s": round(rays / elapsed / 1e6, 2),
    }

    counts = None
    if not a.no_cpu_baseline:
        base, counts = cpu_baseline(scene, ftw, settings, cfg)
        out["cpu_baseline"] = base

    # roofline of the dominant kernel
    roof = None
    if counts is not None:
        def bytes_per_ray(kind):
            n_rays, n_nodes, n_tris_t = (float(x) for x in counts[kind])
            if n_rays == 0:
                return 0.0
            return RAY_BYTES + NODE_BYTES * n_nodes / n_rays + TRI_BYTES * n_tris_t / n_rays
        if cfg["approach"] == "Streaming" and st.ext_launches:
            kernels = {
                "trace_ext": (st.ext_ms, st.ext_launches, st.ext_rays, bytes_per_ray(0)),
                "trace_shadow": (st.shadow_ms, st.shadow_launches, st.shadow_rays,
                                 bytes_per_ray(1)),
            }
        else:
            # mega-kernel: every ray of the render
            tot_rays = float(counts[0][0] + counts[1][0])
            bpr = (RAY_BYTES * tot_rays + NODE_BYTES * float(counts[0][1] + counts[1][1]) +
                   TRI_BYTES * float(counts[0][2] + counts[1][2])) / max(tot_rays, 1.0)
            kernels = {"megakernel": (st.mega_ms, st.mega_launches, st.rays, bpr)}
        name, (ms, launches, n_rays, bpr) = max(kernels.items(), key=lambda kv: kv[1][0])
        avg_s = ms / 1e3 / max(launches, 1)
        bytes_per_launch = bpr * n_rays / max(launches, 1)
        achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
        traffic = None
        tf = ROOT / "profiles" / f"{name}_{a.config}_traffic.json"
        if tf.exists():
            traffic = json.loads(tf.read_text()).get("bytes_per_dispatch")
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 5),
                "traffic": traffic, "kernel": name,
                "avg_launch_ms": round(avg_s * 1e3, 4),
                "algorithmic_bytes_per_ray": round(bpr, 1),
                "rays_per_launch": round(n_rays / max(launches, 1), 1)}
    out["roofline"] = roof
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
