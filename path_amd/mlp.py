"""Radiance MLP on the GPU (hand-written BF16 MFMA kernels, FP32 accumulate).

Inputs are 9 features per surface point: position (3), shading normal (3),
albedo (3). The network predicts the point's reflected radiance (RGB): a
12-frequency encoding, 6 ReLU layers of width 256 and a linear output layer.
Training uses a relative L2 loss and Adam, with the training set resident in
device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FEATURES = 9
OUTPUTS = 3


def param_count() -> int:
    n = C.c_uint64()
    L.check(L.lib().pt_mlp_param_count(C.byref(n)))
    return int(n.value)


def layer_shapes():
    """[(out, in)] of the 7 weight matrices, in storage order."""
    return [(256, 256)] * 6 + [(16, 256)]


class RadianceMLP:
    def __init__(self, device: int = 0, seed: int = 0):
        h = C.c_void_p()
        L.check(L.lib().pt_mlp_create(device, C.c_uint64(seed), C.byref(h)))
        self.handle = h

    def get_weights(self) -> np.ndarray:
        w = np.zeros(param_count(), np.float32)
        L.check(L.lib().pt_mlp_get_weights(self.handle, L.fptr(w)))
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == param_count()
        L.check(L.lib().pt_mlp_set_weights(self.handle, L.fptr(w)))

    def forward(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        out = np.zeros((f.shape[0], OUTPUTS), np.float32)
        L.check(L.lib().pt_mlp_forward(self.handle, L.fptr(f), f.shape[0], L.fptr(out)))
        return out

    def set_forward_kernel(self, fused: bool) -> None:
        """Inference through the fused kernel (default) or the layer GEMMs."""
        L.check(L.lib().pt_mlp_set_forward_kernel(self.handle, int(bool(fused))))

    def gradient(self, features: np.ndarray, targets: np.ndarray):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, OUTPUTS)
        g = np.zeros(param_count(), np.float32)
        loss = C.c_float()
        L.check(L.lib().pt_mlp_gradient(self.handle, L.fptr(f), L.fptr(t), f.shape[0],
                                        L.fptr(g), C.byref(loss)))
        return g, float(loss.value)

    def train_step(self, features, targets, lr: float) -> float:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, OUTPUTS)
        loss = C.c_float()
        L.check(L.lib().pt_mlp_train_step(self.handle, L.fptr(f), L.fptr(t), f.shape[0],
                                          C.c_float(lr), C.byref(loss)))
        return float(loss.value)

    def set_dataset(self, features: np.ndarray, targets: np.ndarray) -> None:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, OUTPUTS)
        L.check(L.lib().pt_mlp_set_dataset(self.handle, L.fptr(f), L.fptr(t),
                                           C.c_uint64(f.shape[0])))

    def train(self, steps: int, batch: int, lr: float, seed: int = 0, record_losses=True):
        """Adam steps on batches drawn on the device from the resident set.
        Returns (losses or None, seconds)."""
        losses = np.zeros(steps, np.float32) if record_losses else None
        secs = C.c_double()
        L.check(L.lib().pt_mlp_train(self.handle, steps, batch, C.c_float(lr), C.c_uint64(seed),
                                     L.fptr(losses), C.byref(secs)))
        return losses, secs.value

    def time_forward(self, rows: int, iters: int = 10):
        """(ms of one hidden-layer GEMM, ms of a whole forward) on `rows`
        device-resident rows, HIP-event timed."""
        g, f = C.c_float(), C.c_float()
        L.check(L.lib().pt_mlp_time_forward(self.handle, rows, iters, C.byref(g), C.byref(f)))
        return float(g.value), float(f.value)

    def render(self, renderer, scene, film_to_world, x_dim: int, y_dim: int, samples_per: int,
               settings=None):
        """Radiance-cache render (one traced bounce + this model). Returns
        (image (y, x, 3), seconds)."""
        from .renderer import Settings
        settings = settings or Settings()
        img = np.zeros((y_dim, x_dim, 3), np.float32)
        ftw = np.ascontiguousarray(np.asarray(film_to_world).ravel(), np.float32)
        secs = C.c_double()
        L.check(L.lib().pt_render_mlp(renderer.handle, self.handle, scene.handle, L.fptr(ftw),
                                      x_dim, y_dim, samples_per, C.byref(settings.raw),
                                      L.fptr(img), C.byref(secs)))
        return img, secs.value

    def render_first_vertex(self, renderer, scene, film_to_world, x_dim: int, y_dim: int,
                            samples_per: int, settings=None):
        """Radiance-cache render with one query per pixel at the first camera
        hit: samples_per > 0 adds the tracer's first-vertex render (emission
        plus light sample, a model of indirect light); 0 uses the model alone
        (a model of all reflected light). Returns (image (y, x, 3), seconds)."""
        from .renderer import Settings
        settings = settings or Settings()
        img = np.zeros((y_dim, x_dim, 3), np.float32)
        ftw = np.ascontiguousarray(np.asarray(film_to_world).ravel(), np.float32)
        secs = C.c_double()
        L.check(L.lib().pt_render_mlp_first_vertex(
            renderer.handle, self.handle, scene.handle, L.fptr(ftw), x_dim, y_dim, samples_per,
            C.byref(settings.raw), L.fptr(img), C.byref(secs)))
        return img, secs.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and L._lib is not None:
            L.lib().pt_mlp_destroy(h)
            self.handle = None
