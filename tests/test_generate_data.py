"""The generate_data surface (reference python_module.cpp / generate_data.cpp
layouts)."""
import numpy as np
import pytest


def test_constants_match_reference_layout():
    from path_amd.generate_data import constants as c
    assert (c.n_dims, c.n_baryo_dims, c.n_rgb_dims, c.n_all) == (3, 2, 3, 4)
    assert c.n_tri_values == 3 * 3 * 4 + 3 * 4 + 3 + 4 == 55
    assert c.n_bsdf_values == 14 and c.n_coords_feature_values == 5


def test_triangle_features():
    from path_amd.generate_data import triangle_features, SCALES
    rng = np.random.default_rng(0)
    tris = rng.uniform(-3, 5, (50, 3, 3))
    f, t = triangle_features(tris)
    assert f.shape == (50, 55)
    verts = t.reshape(-1, 3)
    # normalised: centred on the vertex mean, largest extent 1
    assert np.allclose(verts.mean(0), 0, atol=1e-12)
    assert np.isclose(np.abs(verts).max(), 1.0)
    # vertices sorted by (z, y, x) up to the winding flip that puts the normal on +z
    raw = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.all(raw[:, 2] >= 0)
    assert np.all(t[:, 0, 2] <= t[:, 1:, 2].min(1) + 1e-15)
    # first vertex x: value then tanh(x / scale)
    x = t[:, 0, 0]
    assert np.allclose(f[:, 0:4], np.stack([x] + [np.tanh(x / s) for s in SCALES], 1))
    # unit normal, area
    n = f[:, 48:51]
    assert np.allclose(np.linalg.norm(n, axis=1), 1)
    area = 0.5 * np.linalg.norm(raw, axis=1)
    assert np.allclose(f[:, 51], area)


def test_coord_features_unit_direction():
    from path_amd.generate_data import coord_features, triangle_features, uniform_sphere
    rng = np.random.default_rng(1)
    _, t = triangle_features(rng.uniform(-1, 1, (10, 3, 3)))
    d = uniform_sphere(rng, 100)
    assert np.allclose(np.linalg.norm(d, axis=1), 1)
    idx = rng.integers(0, 10, 100)
    c = coord_features(rng.random(100), rng.random(100), t, idx, d)
    assert c.shape == (100, 5)
    assert np.allclose(np.linalg.norm(c[:, 2:], axis=1), 1)


@pytest.mark.gpu
def test_generate_data_shapes_and_values(renderer):
    from path_amd.generate_data import generate_data, generate_data_for_image
    d = generate_data(0, 3, 2, 8, 4, 7)
    x = d.inputs
    S, T = x.mask.shape
    assert S == 3
    assert x.triangle_features.shape == (S, T, 55)
    assert x.bsdf_features.shape == (S, T, 14)
    assert x.emissive_values.shape == (S, T, 3)
    N = x.baryocentric_coords.shape[1]
    assert d.values.shape == (S, 4, N, 3)
    assert x.n_samples_per == [2 * int((~x.mask[i]).sum()) for i in range(S)]
    assert np.isfinite(d.values.numpy()).all()
    assert d.values.numpy().sum() > 0
    img = generate_data_for_image(0, 2, 32, 4, 3, 9)
    ii = img.image_indexes.numpy()
    assert ii.shape[0] == 2 and ii.shape[2] == 2 and ii.max() < 32
    assert img.standard.values.shape[1] == 3
    moved = d.to(__import__("torch").zeros(1, dtype=__import__("torch").float64))
    assert moved.values.dtype == __import__("torch").float64


@pytest.mark.gpu
@pytest.mark.parametrize("image", [False, True])
def test_generate_data_values_equal_cpu_checker(renderer, image):
    """Every scene's InitialIdxAndDir + OutputPerStep values, rendered in one
    batched launch, equal the CPU checker's per-step values for the same
    sample specification bit for bit."""
    from oracle import oracle as O
    from path_amd.generate_data import (_settings, generate_data, generate_data_for_image,
                                        sample_specs)
    n_samples, n_steps = 16, 4
    if image:
        d = generate_data_for_image(0, 3, 24, n_samples, n_steps, 5).standard
        specs = sample_specs(0, 3, 24, 5, True)
    else:
        d = generate_data(0, 3, 3, n_samples, n_steps, 11)
        specs = sample_specs(0, 3, 3, 11, False)
    st = _settings()
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, n_samples)
    vals = d.values.numpy()
    for si, (g, flat, _, coords, idx, rays, _) in enumerate(specs):
        k = len(rays)
        ref, _ = O.render(flat, np.zeros(12, np.float32), st, n_samples, k, 1, tab, 32, 0,
                          initial_rays=rays, initial_idx=idx.astype(np.uint32), n_steps=n_steps)
        got = np.ascontiguousarray(vals[si, :, :k])
        assert np.array_equal(got.view(np.uint32), ref[:, 0].view(np.uint32)), si
        assert np.array_equal(d.inputs.triangle_idxs_for_coords.numpy()[si, :k], idx)
    assert vals.sum() > 0
