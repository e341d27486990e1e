"""Work division invariants (reference tests/src/work_division.cpp:24-110)
on the CPU checker, and agreement of the library's planner."""
import numpy as np
import pytest

from oracle import oracle as O
from path_amd import _lib as L
from path_amd.renderer import work_division


def settings(**kw):
    s = L.PtWorkDivisionSettings()
    L.lib()
    base = dict(block_size=256, target_x_block_size=32, force_target_samples=0,
                forced_target_samples_per_thread=8, base_num_threads=16384,
                samples_per_thread_scaling_power=0.5, max_samples_per_thread=32)
    base.update(kw)
    for k, v in base.items():
        setattr(s, k, v)
    return s


def check(s, spp, x, y):
    wd = work_division(s, spp, x, y)
    info = O.thread_infos(s, spp, x, y)
    sbs, bs = wd["sample_block_size"], wd["block_size"]
    # division covers exactly samples_per
    assert wd["num_sample_blocks"] * sbs * wd["base_samples_per_thread"] + \
        wd["n_threads_per_unit_extra"] == spp
    ex, ey = wd["x_block_size"] * wd["num_x_blocks"], wd["y_block_size"] * wd["num_y_blocks"]
    assert x <= ex < x + wd["x_block_size"] and y <= ey < y + wd["y_block_size"]
    # every (pixel, sample) covered exactly once, in increasing order per pixel
    covered = {}
    for (px, py, st, en, exit_) in info:
        assert bool(exit_) == (px >= x or py >= y)
        if exit_:
            continue
        covered.setdefault((px, py), []).append((st, en))
    assert len(covered) == x * y
    for ranges in covered.values():
        flat = []
        for st, en in sorted(ranges):
            flat.extend(range(st, en))
        assert flat == list(range(spp))


@pytest.mark.parametrize("x", [1, 3, 17, 72])
@pytest.mark.parametrize("y", [1, 3, 17])
@pytest.mark.parametrize("spp", [1, 3, 17, 72, 2048])
def test_combination(x, y, spp):
    check(settings(block_size=64, target_x_block_size=16, force_target_samples=1,
                   forced_target_samples_per_thread=2), spp, x, y)


def test_default_settings():
    check(settings(), 2048, 8, 8)
    wd = work_division(settings(), 16, 256, 256)
    # configuration of the 256^2, 16 spp render: 2 threads per pixel, 512 blocks
    assert wd["sample_block_size"] == 2 and wd["base_samples_per_thread"] == 8
    assert wd["num_x_blocks"] * wd["num_y_blocks"] * wd["num_sample_blocks"] == 512
    wd = work_division(settings(), 64, 512, 512)
    assert wd["sample_block_size"] == 2 and wd["base_samples_per_thread"] == 32


@pytest.mark.parametrize("x", [1, 3, 17, 72, 128])
@pytest.mark.parametrize("y", [1, 3, 17, 72, 128])
def test_reference_coverage_walk(x, y):
    """The reference's own walk (tests/src/work_division.cpp:24-80, grid
    :93-105): threads in launch order advance through a pixel's samples, then
    x inside the x block, then y inside the y block, then the next block;
    exit is set exactly outside the image; the division covers samples_per
    and the image with less than one block of padding (:82-91)."""
    s = settings(block_size=64, target_x_block_size=16, force_target_samples=1,
                 forced_target_samples_per_thread=2)
    for spp in (1, 3, 17, 72, 128, 2048):
        assert O.work_division_walk(s, spp, x, y) == (0, 0), (spp, x, y)


def test_reference_coverage_walk_default_settings():
    assert O.work_division_walk(settings(), 2048, 8, 8) == (0, 0)
    # the render configurations of the benchmark workloads
    for spp, size in ((16, 256), (64, 512), (128, 128), (128, 32), (32, 256)):
        assert O.work_division_walk(settings(), spp, size, size) == (0, 0), (spp, size)

