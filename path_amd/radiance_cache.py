"""Learned radiance cache: training data from the batched datagen render and
the one-bounce estimator that queries the radiance MLP.

Training pairs come from the first camera hit of each pixel of the generated
scenes: the G-buffer gives the point's features (position, shading normal,
albedo) and its emission, the pixel's radiance minus that emission is the
radiance reflected toward the camera (view-independent on the diffuse
surfaces of these scenes).
"""
from __future__ import annotations

import numpy as np

from .datagen import datagen_settings, generate_scene, render_batch


def first_vertex_settings(settings):
    """The same settings terminated after the first vertex: emission seen by
    the camera plus the light sample there (direct light only)."""
    d = settings.copy()
    d.term_prob = "NIters"
    d.n_iters = 0  # NIters stops at the vertex whose bounce count reaches it
    return d


def training_set(renderer, seeds, samples_per: int = 128, settings=None, batch: int = 1024,
                 indirect: bool = False):
    """(features [n, 9], targets [n, 3]) from the scenes of `seeds`.

    indirect=False: targets are the radiance reflected toward the camera
    (pixel minus the emission seen). indirect=True: only its indirect part,
    the full render minus the render terminated after the first vertex on the
    same samples (their first-vertex light samples coincide, so the
    difference holds no direct-light noise)."""
    settings = settings or datagen_settings()
    feats, targs = [], []
    seeds = list(seeds)
    for b0 in range(0, len(seeds), batch):
        gens = [generate_scene(s) for s in seeds[b0:b0 + batch]]
        images, gbufs, _ = render_batch(renderer, gens, samples_per, settings, gbuffer=True)
        if indirect:
            direct, _, _ = render_batch(renderer, gens, samples_per,
                                        first_vertex_settings(settings), gbuffer=False)
        for k, (img, gb) in enumerate(zip(images, gbufs)):
            gb = gb.reshape(-1, 12)
            rgb = img.reshape(-1, 3)
            hit = np.abs(gb[:, 3:6]).sum(1) > 0
            feats.append(gb[hit, 0:9])
            sub = direct[k].reshape(-1, 3)[hit] if indirect else gb[hit, 9:12]
            targs.append(np.maximum(rgb[hit] - sub, 0.0))
    return np.concatenate(feats).astype(np.float32), np.concatenate(targs).astype(np.float32)


def training_sets(renderer, seeds, samples_per: int = 128, settings=None, batch: int = 1024):
    """(features [n, 9], reflected [n, 3], indirect [n, 3]) from one pass
    over the scenes of `seeds`: the targets of training_set for both
    indirect=False and indirect=True on the same first hits."""
    settings = settings or datagen_settings()
    feats, refl, ind = [], [], []
    seeds = list(seeds)
    for b0 in range(0, len(seeds), batch):
        gens = [generate_scene(s) for s in seeds[b0:b0 + batch]]
        images, gbufs, _ = render_batch(renderer, gens, samples_per, settings, gbuffer=True)
        direct, _, _ = render_batch(renderer, gens, samples_per, first_vertex_settings(settings),
                                    gbuffer=False)
        for img, gb, dimg in zip(images, gbufs, direct):
            gb = gb.reshape(-1, 12)
            rgb = img.reshape(-1, 3)
            hit = np.abs(gb[:, 3:6]).sum(1) > 0
            feats.append(gb[hit, 0:9])
            refl.append(np.maximum(rgb[hit] - gb[hit, 9:12], 0.0))
            ind.append(np.maximum(rgb[hit] - dimg.reshape(-1, 3)[hit], 0.0))
    cat = lambda xs: np.concatenate(xs).astype(np.float32)
    return cat(feats), cat(refl), cat(ind)


def rel_rmse(img, ref) -> float:
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(((img - ref) ** 2).mean()) / np.sqrt((ref ** 2).mean()))
