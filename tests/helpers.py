"""Shared helpers for the test-suite (random rays, image metrics)."""
import numpy as np


def random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    return np.concatenate([o, d.astype(np.float32)], axis=1).astype(np.float32)


def rel_rmse(a, ref):
    """Root-mean-square of per-pixel differences relative to the mean of the
    reference image."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((a - ref) ** 2)) / max(np.mean(ref), 1e-12))
