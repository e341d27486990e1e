"""The numpy restatement of the radiance MLP (CPU only): encoding closed
forms, BF16 rounding, and its backward pass against finite differences."""
import numpy as np

from oracle import mlp_ref as M


def _weights(seed=0):
    rng = np.random.default_rng(seed)
    ws = []
    for l, (out, inp) in enumerate(M.SHAPES):
        w = np.zeros((out, inp), np.float32)
        iv = 225 if l == 0 else inp
        ov = 3 if l == len(M.SHAPES) - 1 else out
        b = np.sqrt(6.0 / iv) * (0.1 if l == len(M.SHAPES) - 1 else 1.0)
        w[:ov, :iv] = rng.uniform(-b, b, (ov, iv))
        ws.append(w.ravel())
    return np.concatenate(ws)


def test_encoding_closed_form():
    f = np.full((1, 9), 0.5, np.float32)
    e = M.encode(f)[0]
    assert np.allclose(e[:9], 0.5)
    for d in range(9):
        s = e[9 + 24 * d: 9 + 24 * d + 24: 2]
        c = e[10 + 24 * d: 10 + 24 * d + 24: 2]
        # sin(pi 2^k / 2): 1 for k = 0, then 0; cos: 0, -1, then 1
        assert np.allclose(s, [1.0] + [0.0] * 11, atol=1e-9)
        assert np.allclose(c, [0.0, -1.0] + [1.0] * 10, atol=1e-9)
    assert np.all(e[225:] == 0)


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e-3], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                    # tie -> even
    assert r[2] == 1 + 2 ** -6            # tie -> even (odd mantissa rounds up)
    assert r[3] == 1 + 2 ** -7            # representable
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_backward_matches_central_differences():
    w = _weights(1)
    rng = np.random.default_rng(2)
    f = rng.uniform(-1, 1, (64, 9)).astype(np.float32)
    t = rng.uniform(0, 2, (64, 3))
    loss, g = M.loss_and_grad(w, f, t, emulate=False)
    _, den = M.loss_only(w, f, t)

    def surrogate(wv):  # the denominator is held at the base point
        p, _, _ = M.forward(wv, f, emulate=False)
        return ((p - t) ** 2 / den).sum() / (3 * p.shape[0])
    assert abs(surrogate(w) - loss) <= 1e-12 * abs(loss)
    offs = np.cumsum([0] + [o * i for o, i in M.SHAPES])
    picks = []
    for l, (out, inp) in enumerate(M.SHAPES):
        ov = 3 if l == len(M.SHAPES) - 1 else out
        iv = 225 if l == 0 else inp
        for _ in range(3):
            picks.append(offs[l] + rng.integers(ov) * inp + rng.integers(iv))
    w64 = w.astype(np.float64)
    for i in picks:
        h = 1e-6
        wp, wm = w64.copy(), w64.copy()
        wp[i] += h
        wm[i] -= h
        fd = (surrogate(wp) - surrogate(wm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-4 * np.abs(g).max()), (i, fd, g[i])
    # padded input columns and output rows get no gradient
    assert np.all(g[offs[0]:offs[1]].reshape(256, 256)[:, 225:] == 0)
    assert np.all(g[offs[6]:offs[7]].reshape(16, 256)[3:] == 0)
