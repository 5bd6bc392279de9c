"""The FP64 MLP oracle, pinned by central finite differences of its loss (the backward pass),
closed-form Adam steps and the encoding's layout (no GPU)."""
import numpy as np

from oracle import mlp_ref as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = M.he_uniform(3).astype(np.float64)
    q = M.random_queries(40, 1)
    x = M.encode(q, (-1, -1, -1), (1, 1, 1))
    t = rng.uniform(0.0, 2.0, (40, 3))
    # The gradient treats y^2 + 0.01 as a constant: differentiate the loss with that constant.
    y0, _, _ = M.forward(params, x)
    den0 = y0 * y0 + 0.01
    _, g = M.loss_and_grad(params, x, t)
    idx = rng.choice(params.size, 60, replace=False)
    idx = np.concatenate([idx, [params.size - 1, params.size - 64, 0, 4095, 4096]])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (M.loss_and_grad(p1, x, t, den=den0)[0] - M.loss_and_grad(p2, x, t, den=den0)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_adam_closed_form():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m, v = M.adam(p, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # First step: m/(1-b1) = g, v/(1-b2) = g^2 -> update = lr * sign(g) (up to eps).
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-7)


def test_encoding_layout():
    q = M.random_queries(5, 2)
    q["pos"][0] = (-1, -1, -1)
    x = M.encode(q, (-1, -1, -1), (1, 1, 1))
    assert x.shape == (5, 64)
    # At the box corner every position sine is 0 and every cosine 1 (position slots: 0-23 and
    # 24-27, 48-55).
    pos = [s for s, (c, k, cos, ident) in enumerate(M.SLOTS) if c < 3]
    assert len(pos) == 36
    assert all(np.isclose(x[0, s], 1.0 if M.SLOTS[s][2] else 0.0) for s in pos)
    assert np.allclose(x[:, 60], q["roughness"]) and np.allclose(x[:, 61:], q["albedo"])
    assert np.allclose(x[:, 28], np.sin(np.pi * q["dir"][:, 0]))
    assert np.allclose(x[:, 59], np.cos(2 * np.pi * q["nrm"][:, 2]))


def test_bf16_rounding():
    assert M.to_bf16(1.0) == 1.0
    assert M.to_bf16(1.0 + 2 ** -8) == 1.0          # tie -> even
    assert M.to_bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert M.to_bf16(-3.0) == -3.0


def test_encoding_covers_every_feature_once():
    """The 64 input slots hold position octaves 0-5, direction and normal octaves 0-1 (sin and cos
    each), roughness and albedo, each exactly once; a slot's value is the closed form of its
    feature."""
    feats = [(c, k, cos) for c, k, cos, ident in M.SLOTS if not ident]
    want = [(c, k, cos) for c in range(3) for k in range(6) for cos in (0, 1)]
    want += [(c, k, cos) for c in range(3, 9) for k in range(2) for cos in (0, 1)]
    assert sorted(feats) == sorted(want) and len(feats) == len(set(feats)) == 60
    assert sorted(c for c, _, _, ident in M.SLOTS if ident) == [9, 10, 11, 12]
    q = np.zeros(1, M.QUERY_DTYPE)
    q["pos"], q["dir"], q["nrm"], q["roughness"], q["albedo"] = (0.25, -0.5, 0.75), (0.1, 0.2, 0.3), (0.4, 0.5, 0.6), 0.7, (0.8, 0.9, 1.0)
    x = M.encode(q, (-1, -1, -1), (1, 1, 1))[0]
    u = (np.array([0.25, -0.5, 0.75]) + 1) / 2
    assert np.isclose(x[0], np.sin(4 * np.pi * u[0])) and np.isclose(x[7], np.cos(32 * np.pi * u[0]))
    assert np.isclose(x[9], np.cos(4 * np.pi * u[1])) and np.isclose(x[17], np.cos(np.pi * u[0]))
    assert np.isclose(x[28], np.sin(np.pi * 0.1)) and np.isclose(x[55], np.cos(32 * np.pi * u[2]))
    assert np.isclose(x[59], np.cos(2 * np.pi * 0.6)) and np.allclose(x[60:], [0.7, 0.8, 0.9, 1.0])
