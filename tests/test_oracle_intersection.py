"""Closest-hit oracle: reference known-answer tests and brute-force
equivalence (tests/src/intersection.cu:18-199 of the reference), plus an
independent float64 plane/barycentric solve that pins the Moller-Trumbore
restatement."""
import numpy as np

from oracle import oracle as O
from tests.helpers import random_rays


def tri_rows(tris):
    """(n, 3, 3) vertices -> (n, 9) rows of v0, e1, e2 (float32)."""
    t = np.asarray(tris, np.float32)
    return np.concatenate([t[:, 0], t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]], axis=1).astype(np.float32)


def ray(o, d):
    d = np.asarray(d, np.float64)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    return np.concatenate([np.asarray(o, np.float32), d]).astype(np.float32)


def test_single_triangle_known_answers():
    # reference tests/src/intersection.cu:119-141
    tris = tri_rows([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]])
    rays = np.stack([
        ray([0.1, 0.1, -1], [0, 0, 1]),
        ray([0.8, 0.7, -1], [0, 0, 1]),
        ray([0.3, 0.1, -1], [0, 0, 1]),
        ray([0.1, 0.8, -1], [0, -0.7, 1]),
        ray([0.1, 0.1, -1], [0, 0, 1]),
    ])
    expected = [0, None, 0, 0, 0]
    for fn in (O.loopall, lambda t, r: O.bvh_intersect(t, r)[:3]):
        t, idx, back = fn(tris, rays)
        for i, e in enumerate(expected):
            assert (idx[i] == 0xFFFFFFFF) == (e is None), i
            if e is not None:
                assert idx[i] == e
                assert abs(t[i] - 1.0) < 1e-5 or i == 3


def test_cornell_triangles_bvh_equals_loopall():
    # reference tests/src/intersection.cu:143-155 (short-box face + light quad)
    tris = tri_rows([
        [[0.13, 0, 0], [0.13, 0.6, 0], [0.7, 0.6, 0.17]],
        [[0.13, 0, 0], [0.7, 0.6, 0.17], [0.7, 0, 0.17]],
        [[-0.24, 1.98, 0.16], [-0.24, 1.98, -0.22], [0.23, 1.98, -0.22]],
        [[-0.24, 1.98, 0.16], [0.23, 1.98, -0.22], [0.23, 1.98, 0.16]],
    ])
    rays = random_rays(512, -2, 2, seed=3)
    a = O.loopall(tris, rays)
    b = O.bvh_intersect(tris, rays)
    assert np.array_equal(a[1], b[1])
    assert (a[1] != 0xFFFFFFFF).any()


def test_random_triangle_sets_bvh_equals_loopall():
    # reference tests/src/intersection.cu:157-177: 16 trials, 1-100 triangles
    rng = np.random.default_rng(11)
    for trial in range(16):
        n = int(rng.integers(1, 101))
        tris = tri_rows(rng.uniform(-1, 1, size=(n, 3, 3)))
        rays = random_rays(256, -1, 1, seed=100 + trial)
        a = O.loopall(tris, rays)
        b = O.bvh_intersect(tris, rays)
        assert np.array_equal(a[1], b[1]), trial
        hit = a[1] != 0xFFFFFFFF
        assert np.array_equal(a[0][hit], b[0][hit])


def test_moller_trumbore_against_float64_plane_solve():
    """Hit/miss, distance and side agree with an independent float64 solve
    (ray-plane distance, then barycentric coordinates by Cramer's rule) for
    rays not grazing an edge."""
    rng = np.random.default_rng(5)
    n = 3000
    v = rng.uniform(-1, 1, size=(n, 3, 3))
    o = rng.uniform(-2, 2, size=(n, 3))
    target = v.mean(axis=1) + rng.uniform(-0.6, 0.6, size=(n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    v32 = v.astype(np.float32)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    checked = 0
    for i in range(n):
        tri = tri_rows(v32[i:i + 1])
        t, idx, back = O.loopall(tri, rays[i:i + 1])
        a, b, c = v32[i].astype(np.float64)
        oo = rays[i, :3].astype(np.float64)
        dd = rays[i, 3:].astype(np.float64)
        nrm = np.cross(b - a, c - a)
        denom = nrm @ dd
        if abs(denom) < 1e-3:
            continue
        tt = nrm @ (a - oo) / denom
        p = oo + tt * dd
        m = np.stack([b - a, c - a, nrm], axis=1)
        u, w, _ = np.linalg.solve(m, p - a)
        margin = min(u, w, 1 - u - w)
        if abs(margin) < 1e-4 or abs(tt - 1e-5) < 1e-4:
            continue
        inside = margin > 0 and tt > 1e-5
        assert (idx[0] == 0) == inside, i
        if inside:
            checked += 1
            assert abs(t[0] - tt) < 1e-4 * max(1.0, abs(tt))
            # back side iff the ray travels along the geometric normal
            assert bool(back[0]) == (denom > 0)
    assert checked > 300
