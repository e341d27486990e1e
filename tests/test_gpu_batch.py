"""Batched multi-scene rendering (training-data generation) on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = list(range(6))
SPP = 4


@pytest.fixture(scope="module")
def gen_scenes():
    from path_amd.datagen import generate_scene
    return [generate_scene(s) for s in SEEDS]


def _single(renderer, g, spp, settings):
    img, _ = renderer.render(g.scene, g.film_to_world, g.resolution, g.resolution, spp,
                             settings=settings, output="FloatRGB")
    return img


def test_batch_equals_single_scene_renders(renderer, gen_scenes):
    from path_amd.datagen import datagen_settings, render_batch, set_launch_params
    set_launch_params(renderer, 8, 1, False)
    st = datagen_settings()
    images, gbufs, secs = render_batch(renderer, gen_scenes, SPP, st)
    stats = renderer.last_stats()
    assert stats.samples == sum(g.resolution ** 2 for g in gen_scenes) * SPP
    assert stats.ext_rays > 0 and stats.shadow_rays > 0
    for g, img in zip(gen_scenes, images):
        ref = _single(renderer, g, SPP, st)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), g.seed
        assert np.isfinite(img).all() and img.max() > 0


def test_batch_small_pool_and_graphs_bit_exact(renderer, gen_scenes):
    """Results are independent of the path pool, the traversal grid, the
    number of iterations between host syncs and hipGraph replay."""
    from path_amd.datagen import datagen_settings, render_batch, set_launch_params
    st = datagen_settings()
    set_launch_params(renderer, 8, 1, False)
    base, _, _ = render_batch(renderer, gen_scenes, SPP, st, gbuffer=False)
    for blocks, iters, graphs, pool in [(4, 4, True, 3000), (16, 2, True, 1 << 20),
                                        (8, 3, False, 777)]:
        s2 = st.copy()
        s2.max_num_samples_per_launch = pool
        set_launch_params(renderer, blocks, iters, graphs)
        imgs, _, _ = render_batch(renderer, gen_scenes, SPP, s2, gbuffer=False)
        for a, b in zip(base, imgs):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (blocks, iters, graphs, pool)
    set_launch_params(renderer, 8, 1, False)


@pytest.mark.parametrize("builder", ["ploc", "lbvh"])
def test_gpu_tree_build_bit_exact(renderer, gen_scenes, builder):
    """Trees built on the GPU (locally-ordered clustering or the binary radix
    tree) render the same images and G-buffers as the host SAH trees: closest
    hits do not depend on the tree, and every triangle is in exactly one
    leaf."""
    from path_amd.datagen import (datagen_settings, generate_scene, last_bvh_build,
                                  render_batch, set_bvh_builder, set_launch_params)
    set_launch_params(renderer, 6, 1, False)
    st = datagen_settings()
    scenes = gen_scenes + [generate_scene(s) for s in (40, 41, 42)]
    host, host_g, _ = render_batch(renderer, scenes, SPP, st)
    set_bvh_builder(renderer, builder)
    try:
        gpu, gpu_g, _ = render_batch(renderer, scenes, SPP, st)
        ms, levels = last_bvh_build(renderer)
    finally:
        set_bvh_builder(renderer, False)
    assert 1 <= levels <= 21 and ms > 0
    for a, b, ga, gb in zip(host, gpu, host_g, gpu_g):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))


def test_batch_with_spatial_split_trees(renderer, gen_scenes):
    """Host trees built with spatial splits (some triangles referenced by
    several leaves) give the batch the same images and G-buffers, bit for bit,
    as the object-split trees; the scenes' trees are rebuilt when the setting
    changes and again when it changes back."""
    from path_amd.datagen import datagen_settings, render_batch, set_launch_params
    from path_amd.renderer import Settings
    set_launch_params(renderer, 8, 1, False)
    st = datagen_settings()
    base, base_g, _ = render_batch(renderer, gen_scenes, SPP, st)
    refs0 = sum(len(g.scene.bvh()["perm"]) for g in gen_scenes)
    sp = st.copy()
    sp.use_spatial_splits = 1
    try:
        imgs, gbufs, _ = render_batch(renderer, gen_scenes, SPP, sp)
        refs1 = sum(len(g.scene.bvh()["perm"]) for g in gen_scenes)
    finally:
        for g in gen_scenes:
            g.scene.set_accel(Settings())
    assert refs1 > refs0  # some triangles were split
    for a, b, ga, gb in zip(base, imgs, base_g, gbufs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    again, _, _ = render_batch(renderer, gen_scenes, SPP, st, gbuffer=False)
    for a, b in zip(base, again):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_gbuffer_first_hit(renderer, gen_scenes):
    """The G-buffer position is the closest hit along the camera ray through
    it, its normal is the hit triangle's (unit, facing the camera side), and
    albedo/emission are the hit material's."""
    from path_amd.datagen import datagen_settings, render_batch
    _, gbufs, _ = render_batch(renderer, gen_scenes[:3], SPP, datagen_settings())
    for g, gb in zip(gen_scenes[:3], gbufs):
        flat = g.scene.flat()
        pos = gb[..., 0:3].reshape(-1, 3)
        nrm = gb[..., 3:6].reshape(-1, 3)
        alb = gb[..., 6:9].reshape(-1, 3)
        emi = gb[..., 9:12].reshape(-1, 3)
        hit = np.abs(nrm).sum(1) > 0
        assert hit.mean() > 0.05
        assert np.allclose(np.linalg.norm(nrm[hit], axis=1), 1.0, atol=1e-4)
        o = g.film_to_world[[3, 7, 11]].astype(np.float64)
        d = pos[hit].astype(np.float64) - o
        dist = np.linalg.norm(d, axis=1)
        rays = np.concatenate([np.broadcast_to(o, d.shape), d / dist[:, None]], 1)
        t, idx, back = renderer.intersect(g.scene, rays.astype(np.float32))
        ok = (idx != 0xFFFFFFFF) & (np.abs(t - dist) <= 1e-3 * np.maximum(dist, 1))
        assert ok.mean() > 0.97
        mats = flat["materials"]
        tri_mat = flat["tri_mat"][idx[ok]]
        want_alb = np.array([list(mats[m].diffuse) for m in tri_mat], np.float32)
        want_emi = np.array([list(mats[m].emission) for m in tri_mat], np.float32)
        # grazing rays may pick a neighbouring triangle of another object
        agree = np.all(np.abs(alb[hit][ok] - want_alb) < 1e-6, axis=1)
        assert agree.mean() > 0.97
        assert np.allclose(emi[hit][ok][agree], want_emi[agree])
        assert (gb[~hit.reshape(gb.shape[:2])] == 0).all()


def test_batch_device_output_and_resident_scenes(renderer, gen_scenes):
    """Images and G-buffers left in HBM equal the host output bit for bit; a
    repeated batch reuses the scenes already on the GPU, and a material change
    invalidates them."""
    from path_amd.datagen import (datagen_settings, generate_scene, render_batch,
                                  render_batch_device, set_launch_params)
    set_launch_params(renderer, 6, 1, False)
    st = datagen_settings()
    scenes = gen_scenes[:3] + [generate_scene(50)]
    host, host_g, _ = render_batch(renderer, scenes, SPP, st)
    total = sum(g.resolution ** 2 for g in scenes)

    def check(dev, imgs, gbs):
        rgb = dev[0].numpy((total, 3))
        gb = dev[1].numpy((total, 12))
        for i, g in enumerate(scenes):
            o, px = int(dev[2][i]), g.resolution ** 2
            assert np.array_equal(rgb[o:o + px].view(np.uint32),
                                  imgs[i].reshape(-1, 3).view(np.uint32)), i
            assert np.array_equal(gb[o:o + px].view(np.uint32),
                                  gbs[i].reshape(-1, 12).view(np.uint32)), i

    dev = render_batch_device(renderer, scenes, SPP, st)
    check(dev, host, host_g)
    dev = render_batch_device(renderer, scenes, SPP, st, out=dev)  # resident scenes
    check(dev, host, host_g)
    # a material change must reach the GPU copy of the batch
    from path_amd import _lib as L
    mats = scenes[3].scene.materials()
    k = next(i for i, m in enumerate(mats) if m.type == L.BSDF["Diffuse"] and any(m.diffuse))
    mats[k].diffuse[:] = [0.5 * v for v in mats[k].diffuse]
    scenes[3].scene.set_material(k, mats[k])
    dev = render_batch_device(renderer, scenes, SPP, st, out=dev)
    host2, host2_g, _ = render_batch(renderer, [generate_scene(0)] + scenes, SPP, st)
    check(dev, host2[1:], host2_g[1:])
    assert not np.array_equal(host2[4], host[3])


def test_resident_batch_with_moved_cameras(renderer, gen_scenes):
    """The same scenes rendered again from other viewpoints: the batch stays
    resident (no repacking), the new cameras reach the GPU, and every image
    equals rendering that scene alone from its new camera."""
    import dataclasses
    from path_amd.datagen import datagen_settings, render_batch, set_launch_params
    set_launch_params(renderer, 6, 1, False)
    st = datagen_settings()
    scenes = gen_scenes[:4]
    first, _, _ = render_batch(renderer, scenes, SPP, st)
    moved = []
    for k, g in enumerate(scenes):
        ftw = g.film_to_world.copy()
        ftw[[3, 7, 11]] += np.float32(0.05 * (k + 1))  # translate the camera
        moved.append(dataclasses.replace(g, film_to_world=ftw))
    again, _, _ = render_batch(renderer, moved, SPP, st)
    for g, a, b in zip(moved, first, again):
        ref = _single(renderer, g, SPP, st)
        assert np.array_equal(b.view(np.uint32), ref.view(np.uint32)), g.seed
        assert not np.array_equal(a.view(np.uint32), b.view(np.uint32)), g.seed
    # and back to the original cameras
    back, _, _ = render_batch(renderer, scenes, SPP, st)
    for a, b in zip(first, back):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _ray_items(gen_scenes, n_each, with_idx, seed=4):
    """Rays starting 1 unit in front of random triangles of each scene, aimed at
    them (an InitialIdxAndDir spec names that triangle as the first hit)."""
    rng = np.random.default_rng(seed)
    items = []
    for k, g in enumerate(gen_scenes):
        tris = g.scene.flat()["tris"].reshape(-1, 3, 3)
        n = 0 if k == 1 else n_each + 7 * k  # one empty item, uneven sizes
        pick = rng.integers(0, len(tris), n).astype(np.uint32)
        c = tris[pick, 0] + (tris[pick, 1] + tris[pick, 2]) / 3.0
        nrm = np.cross(tris[pick, 1], tris[pick, 2])
        nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
        rays = np.concatenate([c + nrm, -nrm], 1).astype(np.float32)
        items.append((g.scene, rays, pick if with_idx else None))
    return items


@pytest.mark.parametrize("with_idx,n_steps,spp", [(True, 5, 16), (False, 0, 16),
                                                  (True, 3, 4096)])
def test_ray_batch_equals_single_renders(renderer, gen_scenes, with_idx, n_steps, spp):
    """pt_render_rays_batch (the reference datagen's per-scene InitialIdxAndDir
    + OutputPerStep renders, batched) equals pt_render of each item alone,
    bit for bit; 4096 spp over few rays spans several sample blocks per ray
    (the second reduction pass)."""
    from path_amd.renderer import Settings
    s = Settings()
    s.kernel_approach = "MegaKernel"
    s.back_cull_emission = 0
    items = _ray_items(gen_scenes[:4], 9 if spp > 1000 else 120, with_idx)
    outs = renderer.render_rays_batch(items, spp, n_steps, s)
    assert renderer.last_stats().samples == sum(len(r) for _, r, _ in items) * spp
    for (scene, rays, idx), got in zip(items, outs):
        if len(rays) == 0:
            assert got.size == 0
            continue
        if n_steps:
            ref, _ = renderer.render(scene, None, 0, 0, spp, s, output="OutputPerStep",
                                     rays=rays, initial_idx=idx, n_steps=n_steps)
            ref = ref[:, 0]
        else:
            ref, _ = renderer.render(scene, None, 0, 0, spp, s, rays=rays, initial_idx=idx)
            ref = ref[0]
        assert got.shape == ref.shape
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        assert got.sum() > 0


def test_ray_batch_rejects_mixed_specs(renderer, gen_scenes):
    from path_amd._lib import PathError
    items = _ray_items(gen_scenes[:3], 10, True)
    items[2] = (items[2][0], items[2][1], None)
    with pytest.raises(PathError):
        renderer.render_rays_batch(items, 4, 2)


def _mesh_scene(tris, emitter=True):
    """Scene from an (n, 9) triangle array (one diffuse material) plus, when
    asked, a quad emitter above the origin; with a camera looking at it."""
    from path_amd.datagen import GeneratedScene
    from path_amd.renderer import Scene, camera_from_look, material
    sc = Scene()
    tris = np.asarray(tris, np.float32).reshape(-1, 9)[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]]  # face +y
    mats = [material("Diffuse", diffuse=(0.6, 0.5, 0.4))]
    tm = np.zeros(len(tris), np.uint32)
    if emitter:
        q = np.array([[-0.3, 1.5, -0.3, 0.3, 1.5, -0.3, 0.3, 1.5, 0.3],
                      [-0.3, 1.5, -0.3, 0.3, 1.5, 0.3, -0.3, 1.5, 0.3]], np.float32)
        tris = np.concatenate([tris, q])
        mats.append(material("Diffuse", diffuse=(0.5, 0.5, 0.5), emission=(10.0, 10.0, 10.0)))
        tm = np.concatenate([tm, np.ones(2, np.uint32)])
    sc.add_mesh(tris, tm, mats)
    cam = camera_from_look((0.0, -0.4, -1.0), (0.0, 1.0, 0.0), (0.0, 1.2, 3.0), 50.0, 1.0)
    return GeneratedScene(sc, cam, 32, -1)


def test_ploc_degenerate_scenes(renderer):
    """GPU trees of scenes that stress the clustering's tie rule and small
    cases (1, 2 and 3 triangles, 300 coincident triangles whose candidate
    areas all tie, a regular grid of equal squares, a batch mixing them)
    render the same images as host SAH trees with both GPU builders."""
    from path_amd.datagen import (datagen_settings, render_batch, set_bvh_builder,
                                  set_launch_params)
    set_launch_params(renderer, 6, 1, False)
    st = datagen_settings()
    tri = np.array([-1, 0, -1, 1, 0, -1, 0, 0, 1], np.float32)
    grid = []
    for i in range(12):
        for j in range(12):
            x, z = -1.2 + 0.2 * i, -1.2 + 0.2 * j
            grid += [[x, 0, z, x + 0.2, 0, z, x + 0.2, 0, z + 0.2],
                     [x, 0, z, x + 0.2, 0, z + 0.2, x, 0, z + 0.2]]
    scenes = [_mesh_scene(tri[None], emitter=False),  # one triangle, no light
              _mesh_scene(tri[None]),                  # 3 with the emitter
              _mesh_scene(np.stack([tri, tri + np.float32(0.01)]), emitter=False),
              _mesh_scene(np.repeat(tri[None], 300, axis=0)),
              _mesh_scene(np.array(grid, np.float32))]
    host, host_g, _ = render_batch(renderer, scenes, 8, st)
    assert any(h.max() > 0 for h in host)
    for builder in ("ploc", "lbvh"):
        set_bvh_builder(renderer, builder)
        try:
            gpu, gpu_g, _ = render_batch(renderer, scenes, 8, st)
        finally:
            set_bvh_builder(renderer, False)
        for k, (a, b, ga, gb) in enumerate(zip(host, gpu, host_g, gpu_g)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (builder, k)
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), (builder, k)
