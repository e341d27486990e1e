"""Sobol table, FNV hash, deterministic float math: pinned by closed forms
and textbook values, and GPU-independent."""
import struct

import numpy as np
import pytest

from oracle import oracle as O


def fnv1a32(data: bytes) -> int:
    h = 0x811C9DC5
    for byte in data:
        h ^= byte
        h = (h * 0x01000193) & 0xFFFFFFFF
    return h


def test_fnv1a_reference_vectors():
    # published FNV-1a 32-bit test vectors pin the byte-wise definition above
    assert fnv1a32(b"") == 0x811C9DC5
    assert fnv1a32(b"a") == 0xE40C292C
    assert fnv1a32(b"foobar") == 0xBF9CF968
    # reference include/lib/fnv.h:5-14 hashes the 4 little-endian bytes
    for v in [0, 1, 255, 256, 65535, 123456789, 0xFFFFFFFF]:
        assert O.fnv_hash(v) == fnv1a32(struct.pack("<I", v))


@pytest.fixture(scope="module")
def dirs():
    return O.joe_kuo_directions(256)


def test_direction_numbers_dims_0_and_1(dirs):
    # dim 0: van der Corput; dim 1: primitive polynomial x + 1, m_k = 2m_{k-1} ^ m_{k-1}
    assert all(int(dirs[0, k]) == 1 << (31 - k) for k in range(32))
    m = 1
    for k in range(32):
        assert int(dirs[1, k]) == (m << (31 - k)) & 0xFFFFFFFF, k
        m = (2 * m) ^ m


def test_direction_numbers_are_valid(dirs):
    # v_k = m_k / 2^(k+1) with m_k odd and m_k < 2^(k+1): as a 32-bit word the
    # lowest set bit is bit 31 - k
    for d in range(256):
        for k in range(32):
            v = int(dirs[d, k])
            assert (v >> (31 - k)) & 1 == 1, (d, k)
            assert v & ((1 << (31 - k)) - 1) == 0, (d, k)


def test_sobol_table_stratification(dirs):
    """Each dimension of the scrambled sequence is a (0, m, 1)-net: the first
    2^m points put exactly one point in each interval [j/2^m, (j+1)/2^m)."""
    tab = O.sobol_table(dirs, 256, 1024)
    assert tab.min() >= 0.0 and tab.max() < 1.0
    for m in (1, 4, 10):
        n = 1 << m
        cells = np.floor(tab[:n] * n).astype(np.int64)
        for d in range(256):
            assert np.array_equal(np.sort(cells[:, d]), np.arange(n)), (m, d)


def test_sobol_dims_0_1_closed_form(dirs):
    """Point i of dimension 0 is the bit reversal of gray(i), xor-scrambled."""
    tab = O.sobol_table(dirs, 2, 64)
    scr0 = None
    for i in range(64):
        g = i ^ (i >> 1)
        rev = int("{:032b}".format(g)[::-1], 2)
        x = tab[i, 0]
        if scr0 is None:
            scr0 = int(x * (1 << 24)) ^ (rev >> 8)
        assert int(x * (1 << 24)) == (rev >> 8) ^ scr0


def test_detmath_accuracy():
    rng = np.random.default_rng(0)
    x = rng.random(20000).astype(np.float32)
    sc = O.detmath("sincos_2pi", x)
    ref = np.stack([np.sin(2 * np.pi * x.astype(np.float64)), np.cos(2 * np.pi * x.astype(np.float64))], 1)
    assert np.abs(sc - ref).max() < 4e-7
    y = rng.uniform(1e-30, 1e4, 20000).astype(np.float32)
    l2 = O.detmath("log2", y)
    assert np.abs(l2 - np.log2(y.astype(np.float64))).max() < 2e-6
    e = rng.uniform(-60, 60, 20000).astype(np.float32)
    ex = O.detmath("exp2", e)
    ref = np.exp2(e.astype(np.float64))
    assert (np.abs(ex - ref) / ref).max() < 8e-7
    b = rng.random(20000).astype(np.float32)
    p = rng.uniform(0.01, 60, 20000).astype(np.float32)
    pw = O.detmath("pow", b, p)
    ref = np.power(b.astype(np.float64), p.astype(np.float64))
    big = ref > 1e-30
    assert (np.abs(pw[big] - ref[big]) / ref[big]).max() < 2e-5
    assert O.detmath("pow", np.float32([0, 1, 0.5]), np.float32([3, 7, 0]))[0] == 0
