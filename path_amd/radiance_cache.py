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


def training_set(renderer, seeds, samples_per: int = 128, settings=None, batch: int = 1024):
    """(features [n, 9], targets [n, 3]) from the scenes of `seeds`."""
    settings = settings or datagen_settings()
    feats, targs = [], []
    seeds = list(seeds)
    for b0 in range(0, len(seeds), batch):
        gens = [generate_scene(s) for s in seeds[b0:b0 + batch]]
        images, gbufs, _ = render_batch(renderer, gens, samples_per, settings, gbuffer=True)
        for img, gb in zip(images, gbufs):
            gb = gb.reshape(-1, 12)
            rgb = img.reshape(-1, 3)
            hit = np.abs(gb[:, 3:6]).sum(1) > 0
            feats.append(gb[hit, 0:9])
            targs.append(np.maximum(rgb[hit] - gb[hit, 9:12], 0.0))
    return np.concatenate(feats).astype(np.float32), np.concatenate(targs).astype(np.float32)


def rel_rmse(img, ref) -> float:
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(((img - ref) ** 2).mean()) / np.sqrt((ref ** 2).mean()))
