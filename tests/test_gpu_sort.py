"""The device tree builder's radix sort against numpy's stable sort: same
permutation bit for bit, at tile edges, digit counts with odd and even
numbers of passes, heavy duplication (stability) and the builder's own
(scene, Morton code) key shape at three million triangles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _expect(renderer, keys, end_bit):
    vals = np.arange(len(keys), dtype=np.uint32)
    ko, vo = renderer.sort_pairs(keys, vals, end_bit)
    masked = keys & np.uint64((1 << end_bit) - 1) if end_bit < 64 else keys
    order = np.argsort(masked, kind="stable")
    assert np.array_equal(vo, vals[order])
    assert np.array_equal(ko, keys[order])


@pytest.mark.parametrize("n", [1, 2, 63, 64, 255, 4095, 4096, 4097, 8192 + 300, 100_003])
@pytest.mark.parametrize("end_bit", [1, 7, 8, 9, 33, 64])
def test_sort_matches_stable_argsort(renderer, n, end_bit):
    rng = np.random.default_rng(n * 131 + end_bit)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n).astype(np.uint64)
    _expect(renderer, keys, end_bit)


def test_sort_is_stable_under_heavy_duplication(renderer):
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 5, 50_000).astype(np.uint64) << np.uint64(20)
    _expect(renderer, keys, 40)
    _expect(renderer, np.zeros(20_000, np.uint64), 64)  # one digit everywhere


def test_sort_tree_builder_keys_at_full_size(renderer):
    """(scene id << 32 | 30-bit Morton code) for 3,000,000 triangles in 700
    scenes: 42 key bits, six passes (an even count: the result is copied
    back)."""
    rng = np.random.default_rng(11)
    n = 3_000_000
    scene = np.sort(rng.integers(0, 700, n)).astype(np.uint64)
    morton = rng.integers(0, 2**30, n, dtype=np.uint64)
    keys = (scene << np.uint64(32)) | morton
    _expect(renderer, keys, 42)


def test_sort_rejects_bad_end_bit(renderer):
    from path_amd import _lib as L
    with pytest.raises(L.PathError):
        renderer.sort_pairs(np.zeros(4, np.uint64), np.zeros(4, np.uint32), 0)
    with pytest.raises(L.PathError):
        renderer.sort_pairs(np.zeros(4, np.uint64), np.zeros(4, np.uint32), 65)
    ko, vo = renderer.sort_pairs(np.zeros(0, np.uint64), np.zeros(0, np.uint32), 64)
    assert len(ko) == 0 and len(vo) == 0
