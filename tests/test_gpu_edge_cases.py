"""Edge cases through the C ABI on the GPU: the reference's intersection
known-answer sets, scenes without triangles or without emitters, one-pixel and
ragged images, empty ray arrays and argument errors."""
import numpy as np
import pytest

from tests.helpers import random_rays

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _settings(approach="MegaKernel", iters=4):
    from path_amd.renderer import Settings
    s = Settings()
    s.kernel_approach = approach
    s.term_prob = "NIters"
    s.n_iters = iters
    return s


def _camera():
    from path_amd.renderer import camera_from_look
    return camera_from_look((0.0, 1.0, 0.0), (0.0, 1.0, 0.0), (0.0, 1.0, 3.4), 45.0, 1.0)


def _triangle_scene(verts):
    """One diffuse mesh from (n, 9) vertex coordinates (v0, v1, v2 per row)."""
    from path_amd.renderer import Scene, material
    verts = np.asarray(verts, np.float32).reshape(-1, 9)
    s = Scene()
    s.add_mesh(verts, np.zeros(len(verts), np.uint32),
               [material("Diffuse", diffuse=(0.5, 0.5, 0.5))])
    return s


def _rays(origins, dirs):
    o = np.asarray(origins, np.float32).reshape(-1, 3)
    d = np.asarray(dirs, np.float32).reshape(-1, 3)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return np.concatenate([o, d], 1).astype(np.float32)


def _intersect_vs_loopall(renderer, verts, rays):
    from oracle import oracle as O
    scene = _triangle_scene(verts)
    t, idx, back = renderer.intersect(scene, rays)
    rt, ridx, rback = O.loopall(scene.flat()["tris"], rays)
    assert np.array_equal(idx, ridx)
    hit = ridx != MISS
    assert np.array_equal(t[hit].view(np.uint32), rt[hit].view(np.uint32))
    assert np.array_equal(back[hit], rback[hit])
    return idx


def test_reference_intersection_known_answers(renderer):
    """The sets of tests/src/intersection.cu:119-177: one triangle with five
    fixed rays (expected ids from the reference) and 64 random rays; a
    four-triangle scene with 512 random rays; 16 soups of 1-100 random
    triangles with 256 random rays each. Every answer also equals the
    brute-force closest hit (LoopAll) of the CPU checker."""
    rng = np.random.default_rng(11)
    one = [[0, 0, 0, 1, 0, 0, 0, 1, 0]]
    rays = _rays([[0.1, 0.1, -1], [0.8, 0.7, -1], [0.3, 0.1, -1], [0.1, 0.8, -1],
                  [0.1, 0.1, -1]],
                 [[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, -0.7, 1], [0, 0, 1]])
    idx = _intersect_vs_loopall(renderer, one, rays)
    assert idx.tolist() == [0, MISS, 0, 0, 0]

    def random_set(n):
        return _rays(rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3)))

    _intersect_vs_loopall(renderer, one, random_set(64))
    four = [[0.13, 0, 0, 0.13, 0.6, 0, 0.7, 0.6, 0.17],
            [0.13, 0, 0, 0.7, 0.6, 0.17, 0.7, 0, 0.17],
            [-0.24, 1.98, 0.16, -0.24, 1.98, -0.22, 0.23, 1.98, -0.22],
            [-0.24, 1.98, 0.16, 0.23, 1.98, -0.22, 0.23, 1.98, 0.16]]
    four_rays = _rays(rng.uniform(-2, 2, (512, 3)), rng.uniform(-2, 2, (512, 3)))
    assert (_intersect_vs_loopall(renderer, four, four_rays) != MISS).any()
    for _ in range(16):
        n = int(rng.integers(1, 101))
        soup = rng.uniform(-1, 1, (n, 9))
        _intersect_vs_loopall(renderer, soup, random_set(256))


def test_empty_ray_array(renderer, cornell):
    scene, _ = cornell
    t, idx, back = renderer.intersect(scene, np.zeros((0, 6), np.float32))
    assert t.size == idx.size == back.size == 0


def test_scene_without_triangles(renderer):
    """A scene with no geometry: every ray misses and every image is black,
    on both integrators."""
    from path_amd.renderer import Scene
    scene = Scene()
    _, idx, _ = renderer.intersect(scene, random_rays(500, -2, 2, seed=5))
    assert (idx == MISS).all()
    for approach in ("MegaKernel", "Streaming"):
        img, _ = renderer.render(scene, _camera(), 7, 5, 4, _settings(approach))
        assert img.shape == (5, 7, 3) and not img.any(), approach


def test_batch_with_an_empty_scene(renderer):
    """A scene without triangles inside a batch renders black and leaves the
    other scenes' images as their single renders, with host-built and with
    GPU-built trees."""
    from path_amd.datagen import (GeneratedScene, datagen_settings, generate_scene,
                                  render_batch, set_bvh_builder, set_launch_params)
    from path_amd.renderer import Scene
    set_launch_params(renderer, 6, 1, False)
    st = datagen_settings()
    scenes = [generate_scene(3), GeneratedScene(Scene(), _camera(), 32, -1), generate_scene(4)]
    singles = [renderer.render(g.scene, g.film_to_world, g.resolution, g.resolution, 4,
                               settings=st)[0] for g in scenes]
    for gpu_tree in ("host", "ploc", "lbvh"):
        set_bvh_builder(renderer, gpu_tree)
        try:
            images, gbufs, _ = render_batch(renderer, scenes, 4, st)
        finally:
            set_bvh_builder(renderer, False)
        assert not images[1].any(), gpu_tree
        for a, b in zip(images, singles):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), gpu_tree
        assert all(np.isfinite(g).all() for g in gbufs)


def test_scene_without_emitters(renderer, tmp_path):
    """No emissive triangle: the light sampler returns nothing (no dimensions
    consumed, random_triangle.h:40-42) and the image is black."""
    from oracle import oracle as O
    from path_amd import scenes
    from path_amd.renderer import Scene
    xml = scenes.cornell_box(tmp_path)
    mtl = tmp_path / "cornell.mtl"
    lines = mtl.read_text().splitlines()
    mtl.write_text("\n".join(ln for ln in lines if not ln.strip().startswith("Ke")) + "\n")
    scene, ftw = Scene.load(xml, 1.0)
    assert scene.counts()[2] == 0
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, 4)
    ref, _ = O.render(scene.flat(), ftw, _settings(), 4, 16, 16, tab)
    assert not ref.any()
    for approach in ("MegaKernel", "Streaming"):
        img, _ = renderer.render(scene, ftw, 16, 16, 4, _settings(approach))
        assert not img.any(), approach


@pytest.mark.parametrize("x,y,spp", [(1, 1, 1), (37, 19, 3), (1, 64, 2), (64, 1, 5)])
def test_tiny_and_ragged_images(renderer, cornell, x, y, spp):
    """Images whose sides are not multiples of the work-division tiles: the
    mega-kernel equals the CPU checker bit for bit and the wavefront
    integrator equals the mega-kernel."""
    from oracle import oracle as O
    scene, ftw = cornell
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    ref, _ = O.render(scene.flat(), ftw, _settings(), spp, x, y, tab)
    mk, _ = renderer.render(scene, ftw, x, y, spp, _settings("MegaKernel"))
    wf, _ = renderer.render(scene, ftw, x, y, spp, _settings("Streaming"))
    assert np.array_equal(mk.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(wf.view(np.uint32), mk.view(np.uint32))


def test_argument_errors(renderer, cornell):
    """Invalid arguments come back as errors (the reference asserts)."""
    from path_amd._lib import PathError
    scene, ftw = cornell
    with pytest.raises(PathError, match="samples_per"):
        renderer.render(scene, ftw, 8, 8, 0, _settings())
    with pytest.raises(PathError, match="empty image"):
        renderer.render(scene, ftw, 0, 8, 1, _settings())
    s = _settings()
    s.max_sample_size = 0
    with pytest.raises(PathError, match="max_sample_size"):
        renderer.render(scene, ftw, 8, 8, 1, s)
    rays = random_rays(4, -0.5, 0.5, seed=1)
    bad = np.array([0, 1, 2, scene.counts()[0]], np.uint32)
    with pytest.raises(PathError, match="out of range"):
        renderer.render(scene, None, 0, 0, 1, _settings(), rays=rays, initial_idx=bad)
    with pytest.raises(PathError, match="MegaKernel"):
        renderer.render(scene, None, 0, 0, 1, _settings("Streaming"), rays=rays,
                        initial_idx=bad[:3].tolist() + [0])
    # the renderer still works after the errors
    img, _ = renderer.render(scene, ftw, 8, 8, 1, _settings())
    assert img.any()


def test_device_output(renderer, cornell):
    """DEVICE_FLOAT_RGB leaves the image in HBM: the same bits as the host
    output, for both integrators; host memory and short buffers are rejected."""
    import ctypes as C
    from path_amd import _lib as L
    from path_amd.renderer import DeviceBuffer
    scene, ftw = cornell
    dev = DeviceBuffer(renderer, 24 * 40 * 12)
    for approach in ("MegaKernel", "Streaming"):
        s = _settings(approach)
        host, _ = renderer.render(scene, ftw, 40, 24, 4, s)
        out, _ = renderer.render(scene, ftw, 40, 24, 4, s, output="DeviceFloatRGB",
                                 device_out=dev)
        assert out is dev
        got = dev.numpy((24, 40, 3))
        assert np.array_equal(got.view(np.uint32), host.view(np.uint32)), approach
    with pytest.raises(ValueError):
        renderer.render(scene, ftw, 48, 48, 1, _settings(), output="DeviceFloatRGB",
                        device_out=dev)
    # the C ABI checks the pointer itself
    spec = L.PtSampleSpec()
    spec.type = L.SAMPLE_SPEC["SquareImage"]
    spec.x_dim = spec.y_dim = 8
    spec.film_to_world[:] = [float(v) for v in np.asarray(ftw).ravel()]
    host_buf = np.zeros((8, 8, 3), np.float32)
    out = L.PtOutput()
    out.type = L.OUTPUT["DeviceFloatRGB"]
    out.data = host_buf.ctypes.data
    secs = C.c_double()
    rc = L.lib().pt_render(renderer.handle, C.byref(spec), C.byref(out), scene.handle, 1,
                           C.byref(_settings().raw), 0, 0, C.byref(secs))
    assert rc != 0 and b"device memory" in L.lib().pt_last_error()
