"""FP64 forward-mode adjoint of the CPU checker (the reference for the GPU
path-replay adjoint), pinned by central finite differences of the FP64
render with the same samples (CPU only)."""
import numpy as np
import pytest

X = Y = 24
SPP = 4


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
    d = tmp_path_factory.mktemp("adj_box")
    return Scene.load(scenes.cornell_box(d), 1.0)


def _render64(O, scene, ftw, st, tab):
    img, _ = O.render(scene.flat(), ftw, st, SPP, X, Y, tab, 64, 0)
    return img


def _fd(O, scene, ftw, st, tab, w, mat_idx, field, ch, h=1e-3):
    """Central difference of sum(w * I) in one material parameter, with the
    step actually representable in float32."""
    base = scene.materials()[mat_idx]
    x0 = float(getattr(base, field)[ch])
    vals = []
    steps = []
    for sgn in (+1, -1):
        m = scene.materials()[mat_idx]
        arr = getattr(m, field)
        arr[ch] = np.float32(x0 + sgn * h)
        steps.append(float(np.float32(x0 + sgn * h)))
        scene.set_material(mat_idx, m)
        vals.append(float((w * _render64(O, scene, ftw, st, tab)).sum()))
    scene.set_material(mat_idx, base)
    return (vals[0] - vals[1]) / (steps[0] - steps[1])


def test_oracle_adjoint_image_is_the_fp64_render(box):
    from oracle import oracle as O
    scene, ftw = box
    st = _settings()
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, SPP)
    w = np.ones((Y, X, 3))
    _, _, img, _ = O.adjoint(scene.flat(), ftw, st, SPP, X, Y, tab, w)
    ref = _render64(O, scene, ftw, st, tab)
    assert np.allclose(img, ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("mat_idx,field,ch", [
    (0, "diffuse", 0), (0, "diffuse", 2),   # floor
    (3, "diffuse", 0),                      # red wall
    (4, "diffuse", 1),                      # green wall
    (6, "diffuse", 2),                      # tall box
    (7, "diffuse", 1),                      # light fixture albedo
    (7, "emission", 0), (7, "emission", 2),  # emitter radiance
])
def test_oracle_adjoint_matches_central_differences(box, mat_idx, field, ch):
    from oracle import oracle as O
    scene, ftw = box
    st = _settings()
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, SPP)
    rng = np.random.default_rng(5)
    w = rng.normal(size=(Y, X, 3))
    ga, ge, _, _ = O.adjoint(scene.flat(), ftw, st, SPP, X, Y, tab, w)
    grad = (ga if field == "diffuse" else ge)[mat_idx, ch]
    fd = _fd(O, scene, ftw, st, tab, w, mat_idx, field, ch)
    assert abs(grad - fd) <= 1e-5 * max(abs(fd), 1e-3), (grad, fd)


def test_oracle_adjoint_emission_gradient_closed_form(box):
    """One emitting material: every path term carries its radiance exactly
    once, so sum_c Le_c * dL/dLe_c equals the loss itself (Euler's theorem
    for a degree-1 homogeneous function)."""
    from oracle import oracle as O
    scene, ftw = box
    st = _settings()
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, SPP)
    w = np.random.default_rng(2).normal(size=(Y, X, 3))
    ga, ge, img, _ = O.adjoint(scene.flat(), ftw, st, SPP, X, Y, tab, w)
    le = np.array(list(scene.materials()[7].emission))
    lhs = float((ge[7] * le).sum())
    rhs = float((w * img).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs)
