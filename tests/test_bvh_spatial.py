"""Spatial splits in the host SAH builder (no GPU): the triangle clipping they
use, restating the reference's known-answer tests
(tests/src/clipped_triangle.cpp:26-118), checked against brute force on
random triangles, and the trees they build: every part of every triangle
stays reachable, leaves partition the reference array, and with spatial
splits off the tree is the plain object-split tree.
"""
import itertools

import numpy as np
import pytest

from path_amd.renderer import Scene, Settings, clip_triangle_bounds

INF = np.float32(np.inf)
PERMS = list(itertools.permutations(range(3)))
BASIC = np.array([[0, 0, 0], [1, 2, 0], [0, 2, 0]], np.float32)


def tri_bounds(tri):
    return tri.min(0), tri.max(0)


def chop(tri, bounds, a, b, axis):
    """ClippedTriangle::new_bounds(a, b, axis) of a triangle whose current
    bounds are `bounds`: the part inside bounds with the axis cut to [a, b]."""
    lo, hi = (np.array(x, np.float32) for x in bounds)
    lo[axis] = max(lo[axis], a)
    hi[axis] = min(hi[axis], b)
    return clip_triangle_bounds(tri, lo, hi)


def expect(actual, lo, hi):
    np.testing.assert_array_equal(actual[0], np.float32(lo))
    np.testing.assert_array_equal(actual[1], np.float32(hi))


@pytest.mark.parametrize("perm", PERMS)
def test_clipped_triangle_basic_2d(perm):
    """tests/src/clipped_triangle.cpp:26-118, every vertex order. The values
    are dyadic, so the clipping lands on them exactly (the reference allows 4
    ulps)."""
    tri = BASIC[list(perm)]
    start = tri_bounds(tri)
    uncut = ((0, 0, 0), (1, 2, 0))
    for axis in range(3):
        expect(chop(tri, start, -INF, INF, axis), *uncut)
        expect(chop(tri, start, 0.0, 2.0, axis), *uncut)
    expect(chop(tri, start, 0.25, 1.0, 0), (0.25, 0.5, 0), (1, 2, 0))
    expect(chop(tri, start, 0.0, 0.75, 0), (0, 0, 0), (0.75, 2, 0))
    expect(chop(tri, start, 0.25, 0.75, 0), (0.25, 0.5, 0), (0.75, 2, 0))
    b = chop(tri, start, 0.25, 0.5, 0)
    expect(chop(tri, b, 1.25, 1.5, 1), (0.25, 1.25, 0), (0.5, 1.5, 0))
    b = chop(tri, start, 1.25, 1.5, 1)
    expect(chop(tri, b, 0.25, 0.5, 0), (0.25, 1.25, 0), (0.5, 1.5, 0))


def test_clip_outside_box_is_empty():
    lo, hi = clip_triangle_bounds(BASIC, (2, 0, -1), (3, 2, 1))
    assert (lo > hi).any()
    # touching the box at a vertex keeps that vertex
    expect(clip_triangle_bounds(BASIC, (1, 2, 0), (3, 3, 1)), (1, 2, 0), (1, 2, 0))


def _sample_points(tri, n=160):
    """Barycentric grid on the triangle, edges and vertices included."""
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = i + j <= n
    u = (i[keep] / n)[:, None]
    v = (j[keep] / n)[:, None]
    t = tri.astype(np.float64)
    return t[0] + u * (t[1] - t[0]) + v * (t[2] - t[0])


def test_clip_matches_brute_force_on_random_triangles():
    """The clipped bounds contain every sampled point of the triangle inside
    the box, and exceed the sampled points' bounds by no more than the
    sampling step: a wrong plane side, a dropped crossing vertex or a wrong
    interpolation fails one of the two."""
    rng = np.random.default_rng(5)
    n_grid = 160
    checked = 0
    for _ in range(300):
        tri = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        c = rng.uniform(-0.6, 0.6, 3)
        half = rng.uniform(0.05, 0.8, 3)
        blo, bhi = (c - half).astype(np.float32), (c + half).astype(np.float32)
        lo, hi = clip_triangle_bounds(tri, blo, bhi)
        p = _sample_points(tri, n_grid)
        inside = np.all((p >= blo) & (p <= bhi), axis=1)
        if not inside.any():
            continue
        checked += 1
        assert (lo <= hi).all()
        plo, phi = p[inside].min(0), p[inside].max(0)
        eps = 1e-5
        assert (lo <= plo + eps).all() and (hi >= phi - eps).all()
        step = np.abs(tri.astype(np.float64)[[1, 2]] - tri[0]).max(0) * 2.0 / n_grid
        assert (lo >= plo - step - eps).all() and (hi <= phi + step + eps).all()
        assert (lo >= blo).all() and (hi <= bhi).all()
    assert checked > 150


def _settings(spatial):
    s = Settings()
    s.use_spatial_splits = spatial
    return s


def _scene_with_long_triangles(seed=3, n=400):
    """Long thin triangles crossing a cloud of small ones: the object splits
    of such a scene overlap and spatial splits pay."""
    from path_amd.renderer import material
    rng = np.random.default_rng(seed)
    tris = []
    for _ in range(n):
        c = rng.uniform(-1, 1, 3)
        tris.append(c + rng.normal(0, 0.03, (3, 3)))
    for _ in range(24):
        a, b = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        tris.append(np.stack([a, b, b + rng.normal(0, 0.05, 3)]))
    tris = np.array(tris, np.float32).reshape(-1, 9)
    sc = Scene()
    sc.add_mesh(tris, np.zeros(len(tris), np.uint32), [material("Diffuse", diffuse=(0.5, 0.5, 0.5))])
    return sc, tris.reshape(-1, 3, 3)


def _check_tree(b, tris):
    lo, hi, leaf, a, bb, perm = b["lo"], b["hi"], b["is_leaf"].astype(bool), b["a"], b["b"], b["perm"]
    n_nodes = len(lo)
    # leaf ranges partition perm in order; inner nodes' boxes contain children
    ranges = sorted((a[i], bb[i]) for i in range(n_nodes) if leaf[i])
    assert ranges[0][0] == 0 and ranges[-1][1] == len(perm)
    assert all(x[1] == y[0] for x, y in zip(ranges, ranges[1:]))
    for i in range(n_nodes):
        if not leaf[i]:
            for c in (a[i], bb[i]):
                assert (lo[c] >= lo[i]).all() and (hi[c] <= hi[i]).all()
    # every triangle referenced at least once
    assert set(perm.tolist()) == set(range(len(tris)))
    # every sampled point of every triangle lies in a leaf box that
    # references that triangle
    boxes = {}
    for i in np.nonzero(leaf)[0]:
        for k in range(a[i], bb[i]):
            boxes.setdefault(int(perm[k]), []).append(i)
    for t in range(len(tris)):
        p = _sample_points(tris[t], 12)
        ids = boxes[t]
        ok = np.zeros(len(p), bool)
        for i in ids:
            ok |= np.all((p >= lo[i] - 1e-6) & (p <= hi[i] + 1e-6), axis=1)
        assert ok.all(), t


def test_spatial_splits_keep_every_triangle_reachable():
    sc, tris = _scene_with_long_triangles()
    sc.set_accel(_settings(1))
    b = sc.bvh()
    assert len(b["perm"]) > len(tris)  # some triangles were split
    _check_tree(b, tris)


def test_spatial_splits_off_is_the_object_split_tree():
    sc, tris = _scene_with_long_triangles()
    sc.set_accel(_settings(0))
    b0 = sc.bvh()
    assert len(b0["perm"]) == len(tris)
    assert sorted(b0["perm"].tolist()) == list(range(len(tris)))
    _check_tree(b0, tris)
    # toggling rebuilds; back off gives the same tree again
    sc.set_accel(_settings(1))
    b1 = sc.bvh()
    sc.set_accel(_settings(0))
    b2 = sc.bvh()
    assert len(b1["perm"]) > len(b0["perm"])
    for k in ("lo", "hi", "is_leaf", "a", "b", "perm"):
        np.testing.assert_array_equal(b0[k], b2[k])


def test_spatial_split_tree_costs_less():
    """The SAH cost of the tree (sum over inner nodes of area x traversal
    cost plus leaves' area x count, over the root's area) drops with spatial
    splits on a scene they target: they are chosen only when cheaper."""
    sc, tris = _scene_with_long_triangles()

    def sah(b):
        lo, hi = b["lo"].astype(np.float64), b["hi"].astype(np.float64)
        d = np.maximum(hi - lo, 0)
        area = 2 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
        leaf = b["is_leaf"].astype(bool)
        cnt = (b["b"] - b["a"]).astype(np.float64)
        return (0.5 * area[~leaf].sum() + (area[leaf] * cnt[leaf]).sum()) / area[0]

    sc.set_accel(_settings(0))
    c0 = sah(sc.bvh())
    sc.set_accel(_settings(1))
    c1 = sah(sc.bvh())
    assert c1 < c0


def test_overlap_threshold_disables_spatial_splits():
    sc, tris = _scene_with_long_triangles()
    s = _settings(1)
    s.overlap_threshold = 1e9  # no overlap is that large
    sc.set_accel(s)
    assert len(sc.bvh()["perm"]) == len(tris)
