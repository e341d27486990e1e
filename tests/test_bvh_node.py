"""The GPU tree's 64-byte node (no GPU): child refs survive packing
unchanged at the extremes the reference's node test uses
(tests/src/node.cpp:7-54: indices 0, 82, 2,838,382 and 2^31 - 1 for inner
nodes and for leaf ranges), empty slots are marked, and every decoded child
box contains its input box and exceeds it by at most one grid step per side
(random boxes, degenerate and far-from-origin ones)."""
import ctypes as C

import numpy as np
import pytest

from path_amd import _lib as L

INT32_MIN = -(2 ** 31)


def pack(lo, hi, child):
    lo = np.ascontiguousarray(lo, np.float32).reshape(-1, 3)
    hi = np.ascontiguousarray(hi, np.float32).reshape(-1, 3)
    n = lo.shape[0]
    ch = (C.c_int32 * 4)(*[int(c) for c in child])
    node = np.zeros(64, np.uint8)
    dlo = np.zeros((4, 3), np.float32)
    dhi = np.zeros((4, 3), np.float32)
    rc = L.lib().pt_pack_node(L.fptr(lo), L.fptr(hi), n, ch, node.ctypes.data_as(C.c_void_p),
                              L.fptr(dlo), L.fptr(dhi))
    return rc, node, dlo, dhi


def fields(node):
    origin = node[0:12].view(np.float32)
    step = node[12:24].view(np.float32)
    child = node[24:40].view(np.int32)
    return origin, step, child


REF_INDICES = [0, 82, 2838382, 2 ** 31 - 1]


@pytest.mark.parametrize("leaf", [False, True])
def test_child_refs_round_trip(leaf):
    """Inner refs are node indices, leaf refs ~first triangle; both are
    stored as given, at the reference test's extremes."""
    rng = np.random.default_rng(1)
    for a in REF_INDICES:
        for b in REF_INDICES:
            refs = [a, b]
            if leaf:
                refs = [~min(r, 2 ** 31 - 2) for r in refs]  # ~first, first < 2^31 - 1
            lo = rng.uniform(-1, 0, (2, 3))
            hi = lo + rng.uniform(0, 1, (2, 3))
            rc, node, _, _ = pack(lo, hi, refs + [0, 0])
            assert rc == L.PT_OK
            _, _, child = fields(node)
            assert child.tolist() == refs + [INT32_MIN, INT32_MIN]


def test_decoded_boxes_contain_and_stay_tight():
    rng = np.random.default_rng(7)
    for trial in range(2000):
        n = int(rng.integers(1, 5))
        scale = 10.0 ** rng.uniform(-3, 4)
        center = rng.uniform(-1, 1, 3) * 10.0 ** rng.uniform(0, 5)
        lo = center + rng.uniform(-1, 1, (n, 3)) * scale
        ext = rng.uniform(0, 1, (n, 3)) * scale
        ext[rng.uniform(size=(n, 3)) < 0.1] = 0.0  # flat boxes
        hi = lo + ext
        lo32, hi32 = lo.astype(np.float32), hi.astype(np.float32)
        rc, node, dlo, dhi = pack(lo32, hi32, list(range(1, n + 1)) + [0] * (4 - n))
        assert rc == L.PT_OK, trial
        origin, step, child = fields(node)
        assert (step > 0).all()
        assert all(float(s) == 2.0 ** np.round(np.log2(float(s))) for s in step)  # powers of two
        assert (dlo[:n] <= lo32).all() and (dhi[:n] >= hi32).all(), trial
        assert (lo32 - dlo[:n] <= step * 1.0001).all(), trial
        assert (dhi[:n] - hi32 <= step * 1.0001).all(), trial
        # the grid spans the union in at most 255 steps
        assert ((hi32.max(0) - origin) <= 255 * step).all()
        # empty slots: marked by their ref (what the traversal tests), with
        # lo code 255 and hi code 0, so on every axis of nonzero extent the
        # decoded box is inverted (on a flat axis the grid step is tiny and
        # both codes decode to the origin)
        if n < 4:
            assert (child[n:] == INT32_MIN).all()
            wide = (hi32.max(0) - lo32.min(0)) > 0
            assert (dlo[n:][:, wide] > dhi[n:][:, wide]).all()


def test_bad_arguments():
    lo = np.zeros((5, 3), np.float32)
    rc, _, _, _ = pack(lo, lo, [0, 0, 0, 0])
    assert rc == L.PT_ERR_INVALID
