"""GPU parity against the CPU checker (oracle/), through the C ABI."""
import numpy as np
import pytest

from tests.helpers import random_rays, rel_rmse

pytestmark = pytest.mark.gpu


def test_sobol_table_bit_exact(renderer):
    from oracle import oracle as O
    dirs = O.joe_kuo_directions(256)
    for count in (1, 16, 63, 64, 65, 1000):
        gpu = renderer.sobol_table(256, count)
        cpu = O.sobol_table(dirs, 256, count)
        assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32)), count


@pytest.mark.parametrize("which", ["cornell", "cornell_sphere"])
def test_intersect_matches_loopall(renderer, request, which):
    from oracle import oracle as O
    scene, _ = request.getfixturevalue(which)
    flat = scene.flat()
    rays = random_rays(20000, -1.2, 2.2, seed=7)
    t, idx, back = renderer.intersect(scene, rays)
    rt, ridx, rback = O.loopall(flat["tris"], rays)
    assert np.array_equal(idx, ridx)
    hit = ridx != 0xFFFFFFFF
    assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
    assert np.array_equal(back[hit], rback[hit])


def _settings(approach="MegaKernel", iters=4):
    from path_amd.renderer import Settings
    s = Settings()
    s.kernel_approach = approach
    s.term_prob = "NIters"
    s.n_iters = iters
    return s


@pytest.mark.parametrize("which,spp,size", [("cornell", 16, 64), ("cornell_sphere", 8, 48)])
def test_megakernel_image_bit_exact(renderer, request, which, spp, size):
    from oracle import oracle as O
    scene, ftw = request.getfixturevalue(which)
    s = _settings()
    img, _ = renderer.render(scene, ftw, size, size, spp, s)
    dirs = O.joe_kuo_directions(256)
    tab = O.sobol_table(dirs, 256, spp)
    ref32, _info = O.render(scene.flat(), ftw, s, spp, size, size, tab, precision=32)
    ref64, _info64 = O.render(scene.flat(), ftw, s, spp, size, size, tab, precision=64)
    diff = np.abs(img.astype(np.float64) - ref32)
    assert np.array_equal(img.view(np.uint32), ref32.view(np.uint32)), (
        f"max |diff| {diff.max()} at {np.unravel_index(diff.argmax(), diff.shape)}")
    # FP64 oracle on the same samples: per-pixel relative RMSE bound
    assert rel_rmse(img, ref64) < 2e-2


@pytest.mark.parametrize("block_size,target_x,spp,per_thread", [
    (64, 8, 16, 2),       # one wave per block, 8 threads per pixel
    (512, 32, 512, 1),    # 512 threads per pixel: reduction through LDS
    (1024, 32, 512, 1),   # 16 waves per block
    (512, 32, 1024, 1),   # two sample blocks per pixel: second reduction pass
])
def test_megakernel_block_sizes_bit_exact(renderer, cornell_sphere, block_size, target_x, spp,
                                          per_thread):
    """Mega-kernel blocks of 64..1024 threads (traversal stacks in LDS per
    wave) against the CPU checker, with the reference's work division."""
    from oracle import oracle as O
    scene, ftw = cornell_sphere
    s = _settings()
    wd = s.render_work_division
    wd.block_size = block_size
    wd.target_x_block_size = target_x
    wd.force_target_samples = 1
    wd.forced_target_samples_per_thread = per_thread
    size = 24
    img, _ = renderer.render(scene, ftw, size, size, spp, s)
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    ref, _ = O.render(scene.flat(), ftw, s, spp, size, size, tab, precision=32)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("block_size,per_thread,spp,x,y", [
    (256, 8, 16, 64, 64),    # the default shape: 2 threads per pixel, 8 samples each
    (256, 3, 13, 37, 19),    # uneven samples per thread, ragged image
    (128, 32, 64, 40, 24),   # 32 samples per thread (the largest run), 2 blocks per 256 lanes
    (1024, 5, 7, 33, 17),    # one thread per pixel, 7 samples over 4 lanes, 16-wave blocks
    (64, 2, 64, 16, 16),     # 32 threads per pixel: the reduction spans half a wave
])
def test_megakernel_split_bit_exact(renderer, cornell_sphere, block_size, per_thread, spp, x, y):
    """Sample-parallel mega-kernel (four lanes per work-division thread) equals
    the one-lane-per-thread kernel and the CPU checker bit for bit."""
    from oracle import oracle as O
    scene, ftw = cornell_sphere
    s = _settings()
    wd = s.render_work_division
    wd.block_size = block_size
    wd.force_target_samples = 1
    wd.forced_target_samples_per_thread = per_thread
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    ref, _ = O.render(scene.flat(), ftw, s, spp, x, y, tab, precision=32)
    try:
        renderer.set_megakernel_split(False)
        one, _ = renderer.render(scene, ftw, x, y, spp, s)
        rays_one = renderer.last_stats().rays
    finally:
        renderer.set_megakernel_split(True)
    split, _ = renderer.render(scene, ftw, x, y, spp, s)
    assert renderer.last_stats().rays == rays_one
    assert np.array_equal(one.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(split.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("which,spp,size", [("cornell", 16, 64), ("cornell_sphere", 8, 48)])
def test_streaming_image_equals_megakernel(renderer, request, which, spp, size):
    """The wavefront integrator produces the mega-kernel's image bit for bit
    (same paths, same per-pixel summation order)."""
    scene, ftw = request.getfixturevalue(which)
    mk, _ = renderer.render(scene, ftw, size, size, spp, _settings("MegaKernel"))
    s = _settings("Streaming")
    wf, _ = renderer.render(scene, ftw, size, size, spp, s)
    assert np.array_equal(mk.view(np.uint32), wf.view(np.uint32))
    # small path pools force many regenerations per slot; below 512 slots the
    # pool runs as one pipeline instead of two
    for pool in (1000, 300):
        s.max_num_samples_per_launch = pool
        wf2, _ = renderer.render(scene, ftw, size, size, spp, s)
        assert np.array_equal(mk.view(np.uint32), wf2.view(np.uint32)), pool
    # any number of pipelines
    from path_amd.datagen import set_pipelines
    s.max_num_samples_per_launch = 5000
    try:
        for n in (1, 2, 4):
            set_pipelines(renderer, n)
            wf3, _ = renderer.render(scene, ftw, size, size, spp, s)
            assert np.array_equal(mk.view(np.uint32), wf3.view(np.uint32)), n
    finally:
        set_pipelines(renderer, 3)


@pytest.mark.parametrize("which", ["cornell_sphere", "benchmark_scene"])
def test_spatial_split_tree_same_hits_and_images(renderer, request, which):
    """With spatial splits a triangle can sit in several leaves, each under a
    box of only its clipped part. Closest hits equal the loop-all oracle's
    (index, t and side bit for bit) and images equal the CPU checker's (whose
    tree has no spatial splits) for both integrators."""
    from oracle import oracle as O
    from path_amd.renderer import Settings
    scene, ftw = request.getfixturevalue(which)
    flat = scene.flat()
    s = _settings()
    s.use_spatial_splits = 1
    try:
        scene.set_accel(s)
        assert len(scene.bvh()["perm"]) > flat["tris"].shape[0]  # triangles were split
        rays = random_rays(20000, -1.2, 2.2, seed=11)
        t, idx, back = renderer.intersect(scene, rays)
        rt, ridx, rback = O.loopall(flat["tris"], rays)
        assert np.array_equal(idx, ridx)
        hit = ridx != 0xFFFFFFFF
        assert hit.mean() > 0.3
        assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
        assert np.array_equal(back[hit], rback[hit])
        spp, size = 4, 40
        tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
        ref32, _ = O.render(flat, ftw, s, spp, size, size, tab, precision=32)
        for approach in ("MegaKernel", "Streaming"):
            s.kernel_approach = approach
            img, _ = renderer.render(scene, ftw, size, size, spp, s)
            assert np.array_equal(img.view(np.uint32), ref32.view(np.uint32)), approach
    finally:
        scene.set_accel(Settings())


def test_streaming_default_settings_multiplier_func(renderer, cornell):
    """Reference default settings (MultiplierFunc Russian roulette): wavefront
    == mega-kernel == CPU checker."""
    from oracle import oracle as O
    from path_amd.renderer import Settings
    scene, ftw = cornell
    s = Settings()
    mk, _ = renderer.render(scene, ftw, 40, 40, 16, s)
    s.kernel_approach = "Streaming"
    wf, _ = renderer.render(scene, ftw, 40, 40, 16, s)
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, 16)
    ref, _ = O.render(scene.flat(), ftw, s, 16, 40, 40, tab, precision=32)
    assert np.array_equal(mk.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(wf.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("n_steps", [3, 12])
def test_output_per_step_bit_exact(renderer, cornell, n_steps):
    """OutputPerStep (rendering_equation.h:56-123): each step image equals the
    CPU checker's bit for bit."""
    from oracle import oracle as O
    scene, ftw = cornell
    s = _settings("MegaKernel")
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, 16)
    gpu, _ = renderer.render(scene, ftw, 48, 40, 16, s, output="OutputPerStep", n_steps=n_steps)
    cpu, _ = O.render(scene.flat(), ftw, s, 16, 48, 40, tab, 32, 0, n_steps=n_steps)
    assert gpu.shape == cpu.shape == (n_steps, 40, 48, 3)
    assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32))
    full, _ = renderer.render(scene, ftw, 48, 40, 16, s)
    if n_steps >= 8:
        assert np.allclose(gpu.sum(0), full, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("spp", [8, 64])  # 64: several samples per thread (split kernel)
def test_initial_rays_and_idx_bit_exact(renderer, cornell_sphere, spp):
    from oracle import oracle as O
    scene, _ = cornell_sphere
    flat = scene.flat()
    tris = flat["tris"].reshape(-1, 3, 3)
    rng = np.random.default_rng(3)
    pick = rng.choice(len(tris), 2000, replace=False).astype(np.uint32)
    v0, e1, e2 = tris[pick, 0], tris[pick, 1], tris[pick, 2]
    c = v0 + (e1 + e2) / 3.0
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    rays = np.concatenate([c + n, -n], 1).astype(np.float32)
    s = _settings("MegaKernel")
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    for idx in (None, pick):
        gpu, _ = renderer.render(scene, None, 0, 0, spp, s, rays=rays, initial_idx=idx)
        cpu, _ = O.render(flat, np.zeros(12, np.float32), s, spp, len(rays), 1, tab, 32, 0,
                          initial_rays=rays, initial_idx=idx)
        assert np.array_equal(gpu.view(np.uint32), cpu.view(np.uint32)), idx is None


def test_fp64_resum_of_streaming_render(renderer, cornell):
    scene, ftw = cornell
    s = _settings("Streaming")
    img, _ = renderer.render(scene, ftw, 40, 32, 64, s)
    img64 = renderer.last_render_fp64(40, 32)
    assert img64.dtype == np.float64
    assert np.allclose(img64, img, rtol=2e-6, atol=1e-7)


SETTINGS_MATRIX = [
    # term prob, light sampler, dir sampler, rng, back-face culled emission
    ("MultiplierFunc", "RandomTriangle", "BSDF", "Sobel", 1),
    ("MultiplierFunc", "NoLightSampling", "BSDF", "Sobel", 1),
    ("Constant", "RandomTriangle", "Uniform", "Sobel", 1),
    ("Normalize", "RandomTriangle", "BSDF", "Uniform", 0),
    ("NIters", "NoLightSampling", "Uniform", "Uniform", 0),
    ("Constant", "NoLightSampling", "BSDF", "Uniform", 1),
]


@pytest.fixture(scope="module")
def all_bsdf_scene():
    """The Cornell box with every BSDF of union_bsdf.h (diffuse, glossy,
    DiffuseGlossy, DiffuseMirror, mirror, dielectric) and a quad emitter."""
    from tests.test_oracle_independent import all_bsdf_scene as make
    return make()


@pytest.mark.parametrize("term,light,dirs,rng,cull", SETTINGS_MATRIX)
def test_settings_matrix_bit_exact(renderer, all_bsdf_scene, term, light, dirs, rng, cull):
    """Every termination rule, both light samplers, both direction samplers,
    both generators and back-face culling on and off, on a scene holding
    every BSDF: both integrators equal the FP32 checker bit for bit."""
    from oracle import oracle as O
    from path_amd.renderer import Settings
    scene, ftw = all_bsdf_scene
    s = Settings()
    s.term_prob = term
    s.n_iters = 5
    s.constant_prob = 0.4
    s.light_sampler = light
    s.dir_sampler = dirs
    s.rng = rng
    s.back_cull_emission = cull
    spp, size = 8, 40
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp) if rng == "Sobel" else None
    ref, _ = O.render(scene.flat(), ftw, s, spp, size, size, tab, precision=32)
    assert ref.mean() > 0
    for approach in ("MegaKernel", "Streaming"):
        s.kernel_approach = approach
        img, _ = renderer.render(scene, ftw, size, size, spp, s)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), approach
