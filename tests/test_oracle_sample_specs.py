"""OutputPerStep and the initial-ray sample specs in the CPU checker."""
import numpy as np
import pytest


def _settings():
    from path_amd.renderer import Settings
    s = Settings()
    s.term_prob = "NIters"
    s.n_iters = 4
    return s


@pytest.fixture(scope="module")
def box(tmp_path_factory):
    from path_amd import scenes
    from path_amd.renderer import Scene
    d = tmp_path_factory.mktemp("spec_box")
    return Scene.load(scenes.cornell_box(d), 1.0)


def test_per_step_images_sum_to_the_image(box):
    from oracle import oracle as O
    scene, ftw = box
    st = _settings()
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, 8)
    img, _ = O.render(scene.flat(), ftw, st, 8, 24, 24, tab, 64, 0)
    steps, _ = O.render(scene.flat(), ftw, st, 8, 24, 24, tab, 64, 0, n_steps=16)
    assert steps.shape == (16, 24, 24, 3)
    assert np.allclose(steps.sum(0), img, rtol=1e-12, atol=1e-14)
    # step 0 is the emission seen directly; no path runs past NIters + 2 steps
    assert np.all(steps[7:] == 0)
    assert steps[1].sum() > 0
    # fewer steps cut the paths: a prefix of the same images
    short, _ = O.render(scene.flat(), ftw, st, 8, 24, 24, tab, 64, 0, n_steps=3)
    assert np.array_equal(short, steps[:3])


def test_initial_idx_and_dir_matches_traced_first_hit(box):
    """Rays aimed at triangle centroids from one unit away: giving the first
    intersection (InitialIdxAndDir) or tracing it (InitialRays) renders the
    same radiance where the traced hit is that triangle at distance ~1."""
    from oracle import oracle as O
    scene, ftw = box
    st = _settings()
    flat = scene.flat()
    tris = flat["tris"].reshape(-1, 3, 3)
    v0, e1, e2 = tris[:, 0], tris[:, 1], tris[:, 2]
    c = v0 + (e1 + e2) / 3.0
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    rays = np.concatenate([c + n, -n], 1).astype(np.float32)   # from the front side
    idx = np.arange(len(tris), dtype=np.uint32)
    t, hit, _ = O.loopall(flat["tris"], rays)
    ok = (hit == idx) & (np.abs(t - 1.0) < 1e-5)
    assert ok.sum() >= len(tris) // 2
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, 16)
    a, _ = O.render(flat, ftw, st, 16, len(rays), 1, tab, 64, 0, initial_rays=rays)
    b, _ = O.render(flat, ftw, st, 16, len(rays), 1, tab, 64, 0, initial_rays=rays,
                    initial_idx=idx)
    a, b = a[0], b[0]
    assert np.allclose(a[ok], b[ok], rtol=1e-3, atol=1e-5)
