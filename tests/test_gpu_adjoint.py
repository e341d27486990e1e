"""Path-replay adjoint on the GPU: against the FP64 forward-mode adjoint of
the CPU checker on the same samples, against central finite differences of
GPU renders (h = 1e-3), and through torch.autograd."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# the differentiable-render workload: Cornell box (8 materials), 256^2, 32 spp
X = Y = 256
SPP = 32


def _settings(n_iters=4):
    from path_amd.renderer import Settings
    s = Settings()
    s.term_prob = "NIters"
    s.n_iters = n_iters
    return s


@pytest.fixture(scope="module")
def dbox(tmp_path_factory):
    from path_amd import scenes
    from path_amd.renderer import Scene
    d = tmp_path_factory.mktemp("dbox")
    return Scene.load(scenes.cornell_box(d), 1.0)


def _target_and_start(dr):
    """Target = the box's own materials; start = albedos scaled per channel
    and the emitter 20% dimmer (same emissive set)."""
    alb0, em0 = dr.parameters()
    dr.set_parameters(alb0.numpy(), em0.numpy())
    target = dr.render().astype(np.float64)
    rng = np.random.default_rng(11)
    alb = np.clip(alb0.numpy() * rng.uniform(0.7, 1.3, alb0.shape), 0.02, 0.98).astype(np.float32)
    em = (em0.numpy() * 0.8).astype(np.float32)
    return target, alb, em, (alb0.numpy(), em0.numpy())


# 4 iterations: one replay per path (at most 16 vertices); 20: radiance replay
# + back-propagation replay
@pytest.mark.parametrize("n_iters", [4, 20])
def test_adjoint_matches_fp64_oracle(renderer, dbox, n_iters):
    from oracle import oracle as O
    from path_amd.diff import DiffRenderer
    scene, ftw = dbox
    st = _settings(n_iters)
    dr = DiffRenderer(renderer, scene, ftw, X, Y, SPP, st)
    target, alb, em, orig = _target_and_start(dr)
    dr.set_parameters(alb, em)
    img = dr.render().astype(np.float64)
    d_img = (2.0 * (img - target)).astype(np.float32)
    ga, ge = dr.adjoint(d_img)
    ga2, ge2 = dr.adjoint(d_img)
    assert np.array_equal(ga, ga2) and np.array_equal(ge, ge2)  # deterministic
    # one replay per path while paths fit the vertex list (n_iters + 1 <= 16)
    assert renderer.last_stats().replays == (1 if n_iters + 1 <= 16 else 2)
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, SPP)
    ra, re, img64, _ = O.adjoint(scene.flat(), ftw, st, SPP, X, Y, tab, d_img.astype(np.float64))
    # the FP32 image is the same estimator as the FP64 one (rare FP32/FP64
    # decision flips along a path aside)
    assert np.abs(img - img64).max() <= 1e-3 * np.abs(img64).max()
    assert np.sqrt(((img - img64) ** 2).mean()) <= 1e-5 * np.abs(img64).max()
    g = np.concatenate([ga.ravel(), ge[7]])
    ref = np.concatenate([ra.ravel(), re[7]])
    scale = np.abs(ref).max()
    err = np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * scale)
    print("max rel err vs FP64 adjoint", err.max())
    assert err.max() <= 1e-3
    dr.set_parameters(*orig)


# Work divisions the 256^2 x 32 spp case does not reach: 131 x 97 at 5 spp is
# 993 wave rows of gradients (ranges of 4 rows, the last of 1; 48 parameters
# put 5 rows in a step, so some threads sum none); 256^2 at 67 spp runs 2
# samples per thread, a pixel's last thread 1.
@pytest.mark.parametrize("x,y,spp", [(131, 97, 5), (256, 256, 67)])
def test_adjoint_ragged_work_division(renderer, dbox, x, y, spp):
    from oracle import oracle as O
    from path_amd.diff import DiffRenderer
    scene, ftw = dbox
    st = _settings(4)
    dr = DiffRenderer(renderer, scene, ftw, x, y, spp, st)
    target, alb, em, orig = _target_and_start(dr)
    dr.set_parameters(alb, em)
    img = dr.render().astype(np.float64)
    d_img = (2.0 * (img - target)).astype(np.float32)
    ga, ge = dr.adjoint(d_img)
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, spp)
    ra, re, _, _ = O.adjoint(scene.flat(), ftw, st, spp, x, y, tab, d_img.astype(np.float64))
    g = np.concatenate([ga.ravel(), ge[7]])
    ref = np.concatenate([ra.ravel(), re[7]])
    err = np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    print(f"{x}x{y} {spp} spp: max rel err vs FP64 adjoint", err.max())
    assert err.max() <= 1e-3
    dr.set_parameters(*orig)


def _central_differences(alb, em, loss, h=1e-3):
    """dloss over the 8 albedo colors and the emitter radiance (material 7),
    with the step actually representable in float32."""
    checks = [("a", m, c) for m in range(8) for c in range(3)] + [("e", 7, c) for c in range(3)]
    fds = []
    for kind, m, c in checks:
        base = alb if kind == "a" else em
        vals, xs = [], []
        for sgn in (+1, -1):
            p = base.copy()
            p[m, c] = np.float32(base[m, c] + sgn * h)
            xs.append(float(p[m, c]))
            vals.append(loss(p, em) if kind == "a" else loss(alb, p))
        fds.append((vals[0] - vals[1]) / (xs[0] - xs[1]))
    return checks, np.array(fds)


def _rel_err(g, ref):
    return np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())


def test_adjoint_matches_central_differences(renderer, dbox):
    """h = 1e-3 central differences. (1) Through FP64 renders of the same
    estimator (CPU checker), of the functional sum(d_image * I) that the
    adjoint differentiates (d_image = the GPU's 2 (I - T)): <= 1e-3 relative.
    (2) Of the L2 loss itself through the GPU's own FP32 renders, where the
    rounding of the two renders limits the difference quotient to a few 1e-3:
    <= 1e-2."""
    from oracle import oracle as O
    from path_amd.diff import DiffRenderer
    scene, ftw = dbox
    st = _settings()
    dr = DiffRenderer(renderer, scene, ftw, X, Y, SPP, st)
    target, alb, em, orig = _target_and_start(dr)
    dr.set_parameters(alb, em)
    img = dr.render().astype(np.float64)
    d_img = (2.0 * (img - target)).astype(np.float32)
    ga, ge = dr.adjoint(d_img)
    tab = O.sobol_table(O.joe_kuo_directions(st.max_sample_size), st.max_sample_size, SPP)

    def loss64(a, e):
        dr.set_parameters(a, e)
        im, _ = O.render(scene.flat(), ftw, st, SPP, X, Y, tab, 64, 0)
        return float((d_img.astype(np.float64) * im).sum())

    def loss32(a, e):
        dr.set_parameters(a, e)
        return float(((dr.render().astype(np.float64) - target) ** 2).sum())

    checks, fd64 = _central_differences(alb, em, loss64)
    _, fd32 = _central_differences(alb, em, loss32)
    g = np.array([float((ga if k == "a" else ge)[m, c]) for k, m, c in checks])
    e64, e32 = _rel_err(g, fd64).max(), _rel_err(g, fd32).max()
    print("max rel err vs central differences: FP64 renders", e64, "FP32 renders", e32)
    assert e64 <= 1e-3, list(zip(checks, g, fd64))
    assert e32 <= 1e-2, list(zip(checks, g, fd32))
    dr.set_parameters(*orig)


def test_autograd_function(renderer, dbox):
    import torch
    from path_amd.diff import DiffRenderer
    scene, ftw = dbox
    dr = DiffRenderer(renderer, scene, ftw, 64, 64, 8, _settings())
    alb0, em0 = dr.parameters()
    target = torch.from_numpy(dr.render())
    alb = (alb0 * 0.9).requires_grad_(True)
    em = (em0 * 1.1).requires_grad_(True)
    img = dr(alb, em)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    d_img = (2.0 * (img.detach() - target)).numpy()
    ga, ge = dr.adjoint(d_img)
    assert torch.equal(alb.grad, torch.from_numpy(ga))
    assert torch.equal(em.grad, torch.from_numpy(ge))
    # a few steps of gradient descent on the albedos reduce the loss
    opt = torch.optim.SGD([alb], lr=float(0.05 / alb.grad.abs().max()))
    first = float(loss)
    for _ in range(5):
        opt.zero_grad()
        l2 = ((dr(alb, em.detach()) - target) ** 2).sum()
        l2.backward()
        opt.step()
    assert float(l2) < first
    dr.set_parameters(alb0.numpy(), em0.numpy())


def test_autograd_two_forwards_one_backward(renderer, dbox):
    """Two forwards through one DiffRenderer at different parameters, then one
    backward of the summed loss: each forward's gradient is the adjoint at its
    own parameters (the second forward must not leak into the first's
    backward), and the scene keeps the latest parameters afterwards."""
    import torch
    from path_amd.diff import DiffRenderer
    scene, ftw = dbox
    dr = DiffRenderer(renderer, scene, ftw, 64, 64, 8, _settings())
    alb0, em0 = dr.parameters()
    target = torch.from_numpy(dr.render())
    a1 = (alb0 * 0.9).requires_grad_(True)
    a2 = (alb0 * 0.6).requires_grad_(True)
    e1 = (em0 * 1.1).requires_grad_(True)
    e2 = (em0 * 0.7).requires_grad_(True)
    img1 = dr(a1, e1)
    img2 = dr(a2, e2)
    (((img1 - target) ** 2).sum() + ((img2 - target) ** 2).sum()).backward()
    cur_a, cur_e = dr.parameters()
    assert torch.equal(cur_a, a2.detach()) and torch.equal(cur_e, e2.detach())
    # each against a single forward + adjoint at that forward's parameters
    for a, e, img in ((a1, e1, img1), (a2, e2, img2)):
        dr.set_parameters(a.detach().numpy(), e.detach().numpy())
        ref = dr.render()
        assert np.array_equal(ref.view(np.uint32), img.detach().numpy().view(np.uint32))
        ga, ge = dr.adjoint((2.0 * (ref - target.numpy())).astype(np.float32))
        assert torch.equal(a.grad, torch.from_numpy(ga))
        assert torch.equal(e.grad, torch.from_numpy(ge))
    assert not torch.equal(a1.grad, a2.grad)
    dr.set_parameters(alb0.numpy(), em0.numpy())


def test_adjoint_rejects_unsupported(renderer, dbox):
    from path_amd._lib import PathError
    scene, ftw = dbox
    s = _settings()
    s.term_prob = "MultiplierFunc"
    with pytest.raises(PathError):
        renderer.adjoint(scene, ftw, 16, 16, 4, np.zeros((16, 16, 3), np.float32), s)


def test_gradient_step_in_hbm(renderer, dbox):
    """Render into device memory, dL/dI = 2 (I - T) on the GPU, adjoint from
    the device gradient: the same gradients as the host-array step."""
    from path_amd.diff import DiffRenderer
    from path_amd.renderer import DeviceBuffer
    scene, ftw = dbox
    st = _settings(4)
    dr = DiffRenderer(renderer, scene, ftw, X, Y, SPP, st)
    target, alb, em, _ = _target_and_start(dr)
    dr.set_parameters(alb, em)
    img = dr.render().astype(np.float64)
    d_host = (2.0 * (img - target)).astype(np.float32)
    ga, ge = dr.adjoint(d_host)
    n = X * Y * 3
    dev_img, dev_tgt, dev_d = (DeviceBuffer(renderer, n * 4) for _ in range(3))
    dev_tgt.write(np.asarray(target, np.float32))
    renderer.render(scene, ftw, X, Y, SPP, st, output="DeviceFloatRGB", device_out=dev_img)
    renderer.l2_loss_gradient(dev_img, dev_tgt, n, dev_d)
    assert np.array_equal(dev_d.numpy((Y, X, 3)).view(np.uint32), d_host.view(np.uint32))
    gd, ged, _ = renderer.adjoint(scene, ftw, X, Y, SPP, dev_d, st)
    assert np.array_equal(gd, ga) and np.array_equal(ged, ge)
