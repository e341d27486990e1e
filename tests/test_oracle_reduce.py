"""The reference's reduction known-answer test (tests/src/reduce.cu:49-158)
on the CPU checker: work-division threads sum their samples in order, the
threads of a location are combined by the lane reduction tree, and the result
equals the brute-force sum over the location's samples - exactly for small
integers, to 4 float ulps for uniform values."""
import numpy as np
import pytest

from oracle import oracle as O
from path_amd import _lib as L
from path_amd.renderer import work_division

LOCATIONS = [1, 2, 3, 7, 8, 17, 32, 256]
SAMPLES = [1, 2, 3, 7, 8, 32, 37, 49, 128, 189, 256]
# the reference also runs 32-thread blocks; wave64 needs multiples of 64
BLOCKS = [64, 128, 256, 1024]


def one_sample_block(block_size, spp, n_loc, base):
    """Double the target samples per thread until the division has one
    sample block (reduce.cu:79-86)."""
    target = base
    while True:
        s = L.PtWorkDivisionSettings()
        L.lib()
        s.block_size = block_size
        s.target_x_block_size = block_size
        s.force_target_samples = 1
        s.forced_target_samples_per_thread = target
        s.base_num_threads = 16384
        s.samples_per_thread_scaling_power = 0.5
        s.max_samples_per_thread = 32
        if work_division(s, spp, n_loc, 1)["num_sample_blocks"] == 1:
            return s
        target *= 2


def cases():
    for n_loc in LOCATIONS:
        for spp in SAMPLES:
            if n_loc * spp > 4096:
                continue
            for bs in BLOCKS:
                for base in (1, 2, 3, 5):
                    yield n_loc, spp, bs, base


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("kind", ["int", "uniform"])
def test_reduce_known_answers(kind):
    rng = np.random.default_rng(1234)
    n = 0
    for n_loc, spp, bs, base in cases():
        s = one_sample_block(bs, spp, n_loc, base)
        if kind == "int":
            vals = rng.integers(-100, 101, n_loc * spp).astype(np.float32)
        else:
            vals = rng.uniform(0.0, 1.0, n_loc * spp).astype(np.float32)
        got = O.reduce_samples(s, vals, n_loc, spp)
        expect = vals.astype(np.float64).reshape(n_loc, spp).sum(1)
        if kind == "int":
            assert np.array_equal(got, expect.astype(np.float32)), (n_loc, spp, bs, base)
        else:
            assert ulps(got, expect.astype(np.float32)).max() <= 4, (n_loc, spp, bs, base)
        n += 1
    assert n > 300

