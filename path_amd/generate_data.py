"""Training-data surface of the reference's `neural_render_generate_data_full_scene`
module (src/generate_data/full_scene/python_module.cpp:9-46), on the path_amd
renderer: `generate_data`, `generate_data_for_image`, `deinit_renderers`,
`Constants`, and the `NetworkInputs` / `StandardData` / `ImageData` records
with the reference's field names, tensor shapes and feature layouts
(src/generate_data/full_scene/generate_data.cpp:36-323).

The radiance values come from InitialIdxAndDir sample specs (each sample
starts on a given triangle point with a given direction) with OutputPerStep
output (one value per path step), as in the reference, but all scenes of a
call are rendered in one batched launch (pt_render_rays_batch) where the
reference calls Renderer::render once per scene
(src/generate_data/full_scene/generate_data.cpp:280-294); each scene's values
equal its own render.

Differences (each deliberate): scenes come from path_amd's seeded
box/sphere generator (`datagen.generate_scene`, seeds base_seed, base_seed+1,
...) instead of the reference's random-triangle generator; sample positions and
directions are drawn with numpy's PCG64 seeded by base_seed (the reference's
CPU `Uniform` stream is not reproducible here); the start rays are built in
the scene's own coordinates (the reference builds them on the normalised,
vertex-sorted triangles used for the features).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from .datagen import generate_scene
from .renderer import Renderer, Settings, camera_from_look

SCALES = (1e-2, 1e0, 1e4)  # value_adder.h:18


class Constants:
    """generate_data/constants.cpp + full_scene/constants.cpp:7-13."""

    def __init__(self):
        self.n_multiscale = len(SCALES)
        self.n_all = self.n_multiscale + 1
        self.n_dims = 3
        self.n_baryo_dims = 2
        self.n_rgb_dims = 3
        points_per_tri = 3
        self.n_tri_values = (self.n_dims * points_per_tri * self.n_all +
                             self.n_dims * self.n_all + self.n_dims + self.n_all)
        self.n_bsdf_values = self.n_rgb_dims * 4 + 2
        self.n_coords_feature_values = self.n_baryo_dims + self.n_dims


constants = Constants()


def _remap_all(v):
    """add_remap_all_value(s): the value, then tanh(v / scale) per scale."""
    v = np.asarray(v, np.float64)
    return np.stack([v] + [np.tanh(v / s) for s in SCALES], -1)


@dataclass
class NetworkInputs:
    triangle_features: torch.Tensor
    mask: torch.Tensor
    bsdf_features: torch.Tensor
    emissive_values: torch.Tensor
    baryocentric_coords: torch.Tensor
    triangle_idxs_for_coords: torch.Tensor
    total_tri_count: int
    n_samples_per: list = field(default_factory=list)

    def to(self, example: torch.Tensor) -> "NetworkInputs":
        d, t = example.device, example.dtype
        return NetworkInputs(self.triangle_features.to(d, t), self.mask.to(d),
                             self.bsdf_features.to(d, t), self.emissive_values.to(d, t),
                             self.baryocentric_coords.to(d, t),
                             self.triangle_idxs_for_coords.to(d), self.total_tri_count,
                             list(self.n_samples_per))


@dataclass
class StandardData:
    inputs: NetworkInputs
    values: torch.Tensor

    def to(self, example: torch.Tensor) -> "StandardData":
        return StandardData(self.inputs.to(example), self.values.to(example.device, example.dtype))


@dataclass
class ImageData:
    standard: StandardData
    image_indexes: torch.Tensor

    def to(self, example: torch.Tensor) -> "ImageData":
        return ImageData(self.standard.to(example), self.image_indexes.to(example.device))


_renderers: list = []


def _renderer() -> Renderer:
    if not _renderers:
        _renderers.append(Renderer(0))
    return _renderers[0]


def deinit_renderers() -> None:
    _renderers.clear()


def triangle_features(tris: np.ndarray):
    """[n, 55] features of flattened triangles [n, 3, 3] (v0, v1, v2) after
    normalising the scene (centre on the vertex mean, scale the largest
    extent from it to 1), sorting each triangle's vertices by (z, y, x) and
    turning its normal to +z (generate_data.cpp:105-156). Also returns the
    normalised, sorted triangles."""
    t = np.asarray(tris, np.float64).reshape(-1, 3, 3)
    verts = t.reshape(-1, 3)
    avg = verts.mean(0)
    scale = 1.0 / max((avg - verts.min(0)).max(), (verts.max(0) - avg).max())
    t = (t - avg) * scale
    order = np.lexsort((t[:, :, 0], t[:, :, 1], t[:, :, 2]), axis=1)
    t = np.take_along_axis(t, order[:, :, None], 1)
    raw = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    flip = raw[:, 2] < 0
    t[flip, 1], t[flip, 2] = t[flip, 2].copy(), t[flip, 1].copy()
    raw = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    scaled = 0.5 * raw
    area = np.linalg.norm(scaled, axis=1)
    normal = scaled / np.maximum(area, 1e-300)[:, None]
    feats = np.concatenate([
        _remap_all(t[:, 0]).reshape(len(t), -1), _remap_all(t[:, 1]).reshape(len(t), -1),
        _remap_all(t[:, 2]).reshape(len(t), -1), _remap_all(scaled).reshape(len(t), -1),
        normal, _remap_all(area)], 1)
    assert feats.shape[1] == constants.n_tri_values
    return feats, t


def bsdf_features(materials, tri_mat):
    """[n, 14]: diffuse, glossy, mirror, dielectric colours, shininess, ior
    (generate_data.cpp:158-206)."""
    from ._lib import BSDF
    rows = []
    for m in materials:
        diff = np.zeros(3)
        gloss = np.zeros(3)
        mirror = np.zeros(3)
        diel = np.zeros(3)
        shin, ior = 40.0, 1.6
        if m.type == BSDF["Diffuse"]:
            diff = np.array(list(m.diffuse))
        elif m.type == BSDF["Glossy"]:
            gloss, shin = np.array(list(m.specular)), m.shininess
        elif m.type == BSDF["DiffuseGlossy"]:
            w = m.diffuse_weight / (m.diffuse_weight + m.glossy_weight)
            diff, gloss = w * np.array(list(m.diffuse)), (1 - w) * np.array(list(m.specular))
            shin = m.shininess
        elif m.type == BSDF["Mirror"]:
            mirror = np.array(list(m.specular))
        elif m.type == BSDF["DielectricRefractive"]:
            diel, ior = np.array(list(m.specular)), m.ior
        else:
            raise ValueError("DiffuseMirror materials are not part of the datagen scenes")
        rows.append(np.concatenate([diff, gloss, mirror, diel, [shin, ior]]))
    table = np.asarray(rows, np.float64)
    return table[np.asarray(tri_mat)]


def coord_features(s, t, tris_norm, tri_idx, dirs):
    """[n, 5]: barycentric (s, t) and the direction in the triangle's frame
    (normal, edge 0, edge 1 components, normalised) (generate_data.cpp:224-239)."""
    tri = tris_norm[tri_idx]
    e0 = tri[:, 1] - tri[:, 0]
    e1 = tri[:, 2] - tri[:, 0]
    n = np.cross(e0, e1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = np.asarray(dirs, np.float64)
    v = np.stack([(n * d).sum(1), (e0 * d).sum(1) / np.linalg.norm(e0, axis=1),
                  (e1 * d).sum(1) / np.linalg.norm(e1, axis=1)], 1)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return np.concatenate([np.stack([s, t], 1), v], 1)


def uniform_sphere(rng, n):
    """Directions uniform on the sphere (uniform_direction_sample.h:10-25 with
    need_whole_sphere)."""
    phi = 2 * np.pi * rng.random(n)
    ct = 2 * rng.random(n) - 1
    st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
    return np.stack([st * np.cos(phi), st * np.sin(phi), ct], 1)


def _settings():
    s = Settings()
    s.kernel_approach = "MegaKernel"
    s.back_cull_emission = False  # amend_config.h:9-12
    return s


def sample_specs(max_tris, forced_n_scenes, per, base_seed, image):
    """The scenes and sample specifications of one generate_data(_for_image)
    call: a list of (GeneratedScene, flat scene, vertices, coords, triangle
    indexes, rays (n, 6), image indexes or None) per scene."""
    rng = np.random.default_rng(base_seed)
    r = _renderer()
    scenes = []
    tri_count = 0
    seed = base_seed
    while (len(scenes) < forced_n_scenes) if forced_n_scenes is not None else \
            (tri_count < max_tris):
        g = generate_scene(seed)
        seed += 1
        flat = g.scene.flat()
        tris = flat["tris"].reshape(-1, 3, 3).astype(np.float64)
        # v0, e1, e2 -> vertices
        verts = np.stack([tris[:, 0], tris[:, 0] + tris[:, 1], tris[:, 0] + tris[:, 2]], 1)
        tri_count += len(verts)
        scenes.append((g, flat, verts))
    out = []
    for g, flat, verts in scenes:
        n = len(verts)
        _, norm = triangle_features(verts)
        if image:
            dim = per
            ftw = camera_from_look((0, 0, -1), (0, 1, 0), (0, 0, 10), 45.0, 1.0)
            ys, xs = np.mgrid[0:dim, 0:dim]
            xs, ys = xs.ravel(), ys.ravel()
            cam = ftw.reshape(3, 4).astype(np.float64)
            fx = 2.0 * xs / dim - 1.0
            fy = -2.0 * ys / dim + 1.0
            film = np.stack([fx, fy, -np.ones_like(fx)], 1)
            o = cam[:, 3]
            w = film @ cam[:, :3].T + o
            d = w - o
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            rays = np.concatenate([np.broadcast_to(o, d.shape), d], 1).astype(np.float32)
            t_hit, idx, _ = r.intersect(g.scene, rays)
            hit = idx != 0xFFFFFFFF
            idx = idx[hit].astype(np.int64)
            p = rays[hit, :3].astype(np.float64) + t_hit[hit, None] * d[hit]
            v0 = verts[idx, 0]
            e1 = verts[idx, 1] - v0
            e2 = verts[idx, 2] - v0
            # barycentric values of the hit point on the scene's triangle
            g11, g12, g22 = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
            q = p - v0
            b1, b2 = (q * e1).sum(1), (q * e2).sum(1)
            det = g11 * g22 - g12 * g12
            s = (g22 * b1 - g12 * b2) / det
            tt = (g11 * b2 - g12 * b1) / det
            dirs = d[hit]
            origins = p - dirs
            image_idx = np.stack([xs[hit], ys[hit]], 1)
        else:
            idx = np.repeat(np.arange(n), per)
            u, v = rng.random(len(idx)), rng.random(len(idx))
            over = u + v > 1
            s = np.where(over, 1 - u, u)
            tt = np.where(over, 1 - v, v)
            dirs = uniform_sphere(rng, len(idx))
            # the same (s, t) on the normalised, sorted triangle and on the scene's
            # triangle: map through the inverse of the normalisation
            pn = norm[idx, 0] + s[:, None] * (norm[idx, 1] - norm[idx, 0]) + \
                tt[:, None] * (norm[idx, 2] - norm[idx, 0])
            vv = verts.reshape(-1, 3)
            avg = vv.mean(0)
            sc = 1.0 / max((avg - vv.min(0)).max(), (vv.max(0) - avg).max())
            p = pn / sc + avg
            origins = p - dirs
            image_idx = None
        coords = coord_features(s, tt, norm, idx, dirs)
        rays = np.concatenate([origins, dirs], 1).astype(np.float32)
        out.append((g, flat, verts, coords, idx, rays, image_idx))
    return out


def _generate(max_tris, forced_n_scenes, per, n_samples, n_steps, base_seed, image):
    r = _renderer()
    specs = sample_specs(max_tris, forced_n_scenes, per, base_seed, image)
    # every scene's samples in one batched launch
    values = r.render_rays_batch([(g.scene, rays, idx.astype(np.uint32))
                                  for g, _, _, _, idx, rays, _ in specs],
                                 n_samples, n_steps, _settings())
    tri_count = sum(len(v) for _, _, v, _, _, _, _ in specs)
    max_t = max(len(v) for _, _, v, _, _, _, _ in specs)
    S = len(specs)
    tf = np.zeros((S, max_t, constants.n_tri_values), np.float32)
    bf = np.zeros((S, max_t, constants.n_bsdf_values), np.float32)
    ev = np.zeros((S, max_t, 3), np.float32)
    mask = np.ones((S, max_t), bool)
    for si, (g, flat, verts, _, _, _, _) in enumerate(specs):
        n = len(verts)
        feats, _ = triangle_features(verts)
        tf[si, :n] = feats
        mats = [flat["materials"][i] for i in range(flat["n_materials"])]
        bf[si, :n] = bsdf_features(mats, flat["tri_mat"])
        ev[si, :n] = np.array([list(mats[m].emission) for m in flat["tri_mat"]])
        mask[si, :n] = False
    N = max(len(c) for _, _, _, c, _, _, _ in specs)
    bc = np.zeros((S, N, constants.n_coords_feature_values), np.float32)
    ti = np.zeros((S, N), np.int64)
    vals = np.zeros((S, n_steps, N, 3), np.float32)
    ii = np.zeros((S, N, 2), np.int64)
    for si, ((_, _, _, coords, idx, _, image_idx), v) in enumerate(zip(specs, values)):
        k = len(coords)
        bc[si, :k] = coords
        ti[si, :k] = idx
        vals[si, :, :k] = v
        if image_idx is not None:
            ii[si, :k] = image_idx
    inputs = NetworkInputs(torch.from_numpy(tf), torch.from_numpy(mask), torch.from_numpy(bf),
                           torch.from_numpy(ev), torch.from_numpy(bc), torch.from_numpy(ti),
                           int(tri_count), [len(c) for _, _, _, c, _, _, _ in specs])
    std = StandardData(inputs, torch.from_numpy(vals))
    return ImageData(std, torch.from_numpy(ii)) if image else std


def generate_data(max_tris: int, forced_n_scenes, n_samples_per_tri: int, n_samples: int,
                  n_steps: int, base_seed: int) -> StandardData:
    """Scenes until max_tris triangles (or exactly forced_n_scenes), with
    n_samples_per_tri random (point, direction) samples per triangle, each
    rendered with n_samples samples into n_steps per-step values."""
    return _generate(max_tris, forced_n_scenes, n_samples_per_tri, n_samples, n_steps,
                     base_seed, image=False)


def generate_data_for_image(max_tris: int, forced_n_scenes, dim: int, n_samples: int,
                            n_steps: int, base_seed: int) -> ImageData:
    """As generate_data, with the samples at the first hits of a dim x dim
    camera (position (0, 0, 10), looking down -z, 45 degrees) and their pixel
    indexes."""
    return _generate(max_tris, forced_n_scenes, dim, n_samples, n_steps, base_seed, image=True)
