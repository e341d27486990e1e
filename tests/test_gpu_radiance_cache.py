"""Radiance-cache estimator: the trained MLP against the same one-bounce
estimator without it, on a held-out generated scene."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_one_bounce_estimator_without_model_is_direct_light(renderer):
    """With a zero network the cache estimator is the path tracer cut after
    the first vertex's light sample: it equals a NIters render that stops
    there (same samples)."""
    from path_amd.datagen import datagen_settings, generate_scene
    from path_amd.mlp import RadianceMLP, param_count
    g = generate_scene(123)
    st = datagen_settings()
    st.kernel_approach = "MegaKernel"
    st.term_prob = "NIters"
    st.n_iters = 5
    m = RadianceMLP(0, seed=1)
    m.set_weights(np.zeros(param_count(), np.float32))
    img, _ = m.render(renderer, g.scene, g.film_to_world, 64, 64, 8, st)
    st1 = st.copy()
    st1.n_iters = 1   # second vertex: light sample only, no further bounce
    ref, _ = renderer.render(g.scene, g.film_to_world, 64, 64, 8, st1)
    # the one-bounce render adds the second vertex's light sample; the cache
    # estimator with a zero model does not
    st0 = st.copy()
    st0.n_iters = 0
    ref0, _ = renderer.render(g.scene, g.film_to_world, 64, 64, 8, st0)
    assert np.allclose(img, ref0, rtol=1e-5, atol=1e-6)
    assert ref.mean() >= img.mean()


def test_trained_model_beats_no_model(renderer):
    from path_amd.datagen import datagen_settings, generate_scene
    from path_amd.mlp import RadianceMLP, param_count
    from path_amd.radiance_cache import rel_rmse, training_set
    st = datagen_settings()
    st.term_prob = "NIters"
    st.n_iters = 5
    f, t = training_set(renderer, range(64), samples_per=64, settings=st)
    m = RadianceMLP(0, seed=2)
    m.set_dataset(f, t)
    losses, secs = m.train(600, 16384, 1e-3, seed=3)
    print("training rows", f.shape[0], "loss", losses[:10].mean(), "->", losses[-20:].mean(),
          "in", secs, "s")
    g = generate_scene(10 ** 6 + 7)
    stm = st.copy()
    stm.kernel_approach = "MegaKernel"
    gt, _ = renderer.render(g.scene, g.film_to_world, 128, 128, 1024, stm)
    img, _ = m.render(renderer, g.scene, g.film_to_world, 128, 128, 16, stm)
    zero = RadianceMLP(0, seed=2)
    zero.set_weights(np.zeros(param_count(), np.float32))
    base, _ = zero.render(renderer, g.scene, g.film_to_world, 128, 128, 16, stm)
    e_m, e_0 = rel_rmse(img, gt), rel_rmse(base, gt)
    print("held-out relative RMSE: with model", e_m, "without", e_0)
    assert e_m < e_0
