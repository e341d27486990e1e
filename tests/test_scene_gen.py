"""Seeded scene generator for training-data generation (host side, no GPU)."""
import numpy as np


def test_generator_deterministic_and_in_range():
    from path_amd.datagen import generate_scene
    tri_counts = []
    for seed in range(40):
        a = generate_scene(seed)
        b = generate_scene(seed)
        fa, fb = a.scene.flat(), b.scene.flat()
        assert np.array_equal(fa["tris"], fb["tris"])
        assert np.array_equal(a.film_to_world, b.film_to_world)
        assert a.resolution == b.resolution and a.resolution in (32, 64, 96, 128)
        nt, nm, nl = a.scene.counts()
        # objects: >= 2 boxes (12 tris) + 1-3 emitter quads
        assert 2 * 12 + 2 <= nt <= 25000
        assert nl in (2, 4, 6)
        assert 3 <= nm <= 43
        assert np.isfinite(fa["tris"]).all()
        # emitters face down and carry radiance in [5, 30]
        mats = fa["materials"]
        emit = [m for m in list(mats)[:nm] if max(m.emission) > 0]
        assert 1 <= len(emit) <= 3
        assert all(5 <= m.emission[0] <= 30 for m in emit)
        tri_counts.append(nt)
    assert len(set(tri_counts)) > 30  # seeds give different scenes
    assert min(tri_counts) < 200 and max(tri_counts) > 2000


def test_generated_bvh_fits_stack():
    from path_amd.datagen import generate_scene
    for seed in range(0, 200, 17):
        g = generate_scene(seed)
        assert g.scene.bvh()["max_depth"] < 60
