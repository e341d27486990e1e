"""BSDF samplers, light sampling and the path estimator against closed
forms (moments of the sampled distributions, Snell's law, furnace)."""
import numpy as np
import pytest

from oracle import oracle as O
from path_amd.renderer import Scene, Settings, material


def sobol2(n):
    d = O.joe_kuo_directions(2)
    return O.sobol_table(d, 2, n).astype(np.float64)


N = np.array([0.0, 0.0, 1.0])
IN = -np.array([0.3, 0.2, 0.9]) / np.linalg.norm([0.3, 0.2, 0.9])


def test_diffuse_cosine_sampling_moments():
    m = material("Diffuse", diffuse=(0.5, 0.6, 0.7))
    u = sobol2(4096)
    cos = []
    for a, b in u:
        d, mult, disc, used = O.bsdf_sample(m, IN, N, [a, b])
        assert used == 2 and not disc
        assert np.allclose(mult, [0.5, 0.6, 0.7], atol=1e-7)
        assert abs(np.linalg.norm(d) - 1) < 1e-12
        cos.append(d @ N)
    cos = np.array(cos)
    assert cos.min() >= -1e-12
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2
    assert abs(cos.mean() - 2 / 3) < 2e-3
    assert abs((cos ** 2).mean() - 0.5) < 2e-3


def test_diffuse_flips_to_incoming_side():
    m = material("Diffuse", diffuse=(1, 1, 1))
    d, _, _, _ = O.bsdf_sample(m, -IN, N, [0.3, 0.6])  # ray arriving from below
    assert d @ N < 0


@pytest.mark.parametrize("shininess", [5.0, 50.0])
def test_glossy_phong_lobe_moments(shininess):
    m = material("Glossy", specular=(0.8, 0.8, 0.8), shininess=shininess)
    refl = IN - 2 * (IN @ N) * N
    cosa = []
    for a, b in sobol2(4096):
        d, mult, disc, used = O.bsdf_sample(m, IN, N, [a, b])
        cosa.append(d @ refl)
    cosa = np.array(cosa)
    n = shininess
    # pdf of cos(alpha) ~ (n+1) c^n: E[c] = (n+1)/(n+2)
    assert abs(cosa.mean() - (n + 1) / (n + 2)) < 2e-3
    # eval integrates the lobe normalization (n+2)/(2 pi): f(peak) = ks (n+2)/(2pi)
    f, pc = O.bsdf_eval(m, IN, refl, N)
    assert np.allclose(f, 0.8 * (n + 2) / (2 * np.pi), rtol=1e-6) and pc == 1.0


def test_dielectric_fresnel_and_snell():
    ior = 1.5
    m = material("DielectricRefractive", specular=(1, 1, 1), ior=ior)
    r0 = ((ior - 1) / (ior + 1)) ** 2
    # normal incidence: reflect iff u < R0
    d, _, disc, used = O.bsdf_sample(m, -N, N, [r0 * 0.99])
    assert disc and used == 1 and np.allclose(d, N)
    d, _, _, _ = O.bsdf_sample(m, -N, N, [r0 * 1.01])
    assert np.allclose(d, -N)
    # oblique refraction obeys Snell's law, entering and exiting
    for theta in [0.2, 0.6, 1.0]:
        inc = np.array([np.sin(theta), 0, -np.cos(theta)])
        d, _, _, _ = O.bsdf_sample(m, inc, N, [0.999999])
        assert abs(np.linalg.norm(np.cross(d, N)) - np.sin(theta) / ior) < 1e-9
        assert d @ N < 0 and d[0] > 0
    inc = np.array([np.sin(0.3), 0, np.cos(0.3)])  # exiting the medium
    d, _, _, _ = O.bsdf_sample(m, inc, N, [0.999999])
    assert abs(np.linalg.norm(np.cross(d, N)) - ior * np.sin(0.3)) < 1e-9
    # total internal reflection beyond the critical angle
    inc = np.array([np.sin(1.0), 0, np.cos(1.0)])
    d, _, _, _ = O.bsdf_sample(m, inc, N, [0.999999])
    assert np.allclose(d, inc - 2 * (inc @ N) * N)


def test_combined_lobe_draws():
    # DiffuseGlossy draws the lobe index before the lobe's two values
    dg = material("DiffuseGlossy", diffuse=(0.5, 0.5, 0.5), specular=(0.5, 0.5, 0.5),
                  shininess=10, diffuse_weight=0.6, glossy_weight=0.4)
    _, mult, disc, used = O.bsdf_sample(dg, IN, N, [0.5, 0.3, 0.3])
    assert used == 3 and not disc
    # DiffuseMirror: continuous-vs-discrete coin, then the lobe draw
    dm = material("DiffuseMirror", diffuse=(0.5, 0.5, 0.5), specular=(0.5, 0.5, 0.5),
                  shininess=1000, diffuse_weight=0.6, glossy_weight=0.4)
    _, _, disc, used = O.bsdf_sample(dm, IN, N, [0.9, 0.2])  # 0.6 < 0.9: discrete
    assert disc and used == 2
    _, _, disc, used = O.bsdf_sample(dm, IN, N, [0.1, 0.2, 0.3, 0.4])
    assert not disc and used == 4


def test_term_probs():
    s = Settings()
    s.term_prob = "NIters"
    s.n_iters = 3
    assert O.term_prob(s, 2, [1, 1, 1]) == 0.0 and O.term_prob(s, 3, [1, 1, 1]) == 1.0
    s.term_prob = "MultiplierFunc"
    # reference defaults exp 50, min 0.1: q = 0.57 m/(m+1)
    m = np.array([0.5, 0.3, 0.1])
    q = 0.57 * m / (m + 1)
    expect = max(abs((1 - min(q @ q, 1)) ** 50), 0.1)
    assert abs(O.term_prob(s, 0, m) - expect) < 1e-7  # settings are float32
    s.term_prob = "Constant"
    s.constant_prob = 0.25
    assert O.term_prob(s, 5, m) == 0.25


def _furnace_scene(albedo, le):
    """Closed cube [-1,1]^3 with inward faces, every face diffuse + emissive."""
    q = []
    c = [-1.0, 1.0]
    faces = [
        [(-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1)],  # y=-1
        [(-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)],      # y=+1
        [(-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1)],  # z=-1
        [(-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)],      # z=+1
        [(-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1)],  # x=-1
        [(1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)],      # x=+1
    ]
    for a, b, cc, d in faces:  # listed with outward normals: reverse the winding
        q.append([a, cc, b])
        q.append([a, d, cc])
    verts = np.array(q, np.float32).reshape(-1, 9)
    sc = Scene()
    sc.add_mesh(verts, np.zeros(len(q), np.uint32),
                [material("Diffuse", diffuse=(albedo,) * 3, emission=(le,) * 3)])
    return sc


@pytest.mark.parametrize("light", ["RandomTriangle", "NoLightSampling"])
def test_white_furnace(light):
    """Inside a closed emissive diffuse box, NIters{N} gives N+1 surface
    vertices. With NEE every vertex also gathers the next bounce's emission:
    L = Le * sum_{k=0}^{N+1} rho^k; with BSDF sampling only, emission is
    picked up at each of the N+1 hits: L = Le * sum_{k=0}^{N} rho^k (and the
    estimator has zero variance). FP64 oracle."""
    from path_amd.renderer import camera_from_look
    rho, le, iters = 0.6, 1.0, 3
    sc = _furnace_scene(rho, le)
    ftw = camera_from_look([0, 0, -1], [0, 1, 0], [0, 0, 0.2], 60.0, 1.0)
    s = Settings()
    s.light_sampler = light
    s.term_prob = "NIters"
    s.n_iters = iters
    spp = 512
    tab = O.sobol_table(O.joe_kuo_directions(256), 256, spp)
    img, _ = O.render(sc.flat(), ftw, s, spp, 8, 8, tab, precision=64)
    terms = iters + 2 if light == "RandomTriangle" else iters + 1
    expected = le * sum(rho ** k for k in range(terms))
    assert abs(img.mean() - expected) / expected < 0.01, (img.mean(), expected)
