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
    # At the box corner every position sine is 0 and every cosine 1.
    assert np.allclose(x[0, 0:36:2], 0) and np.allclose(x[0, 1:36:2], 1)
    assert np.allclose(x[:, 60], q["roughness"]) and np.allclose(x[:, 61:], q["albedo"])
    assert np.allclose(x[:, 36], np.sin(np.pi * q["dir"][:, 0]))
    assert np.allclose(x[:, 48 + 4 * 2 + 3], np.cos(2 * np.pi * q["nrm"][:, 2]))


def test_bf16_rounding():
    assert M.to_bf16(1.0) == 1.0
    assert M.to_bf16(1.0 + 2 ** -8) == 1.0          # tie -> even
    assert M.to_bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert M.to_bf16(-3.0) == -3.0
