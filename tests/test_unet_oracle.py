"""The denoiser oracle (oracle/unet_ref.py) pinned by what the mathematics fixes: central finite
differences of its backward pass, adjoint pairs, closed-form convolutions and the splitmix64
definition of its initialisation."""
import numpy as np
import pytest

from oracle import unet_ref as U


def test_parameter_count_and_layout():
    # 5 levels 32/48/64/96/128, two 3x3 convs per level on each side, 1x1 output.
    assert U.N_PARAMS == 1015651  # hand count: 19 layers, weights + biases
    names = [l[0] for l in U.LAYERS]
    assert names[:2] == ["enc0a", "enc0b"] and names[-3:] == ["dec0a", "dec0b", "out"]
    assert dict((l[0], l[1]) for l in U.LAYERS)["dec3a"] == (128, 96)
    p = np.arange(U.N_PARAMS, dtype=np.float64)
    assert np.array_equal(U.pack(U.unpack(p)), p)


def test_splitmix_stream_matches_definition():
    """Vectorised stream against the scalar splitmix64 definition (Steele et al. 2014)."""
    mask = (1 << 64) - 1
    x = 12345
    ref = []
    for _ in range(50):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        ref.append(np.float32(z >> 40) * np.float32(1.0 / 16777216.0))
    assert np.array_equal(U._splitmix_uniform(12345, 50), np.array(ref, np.float32))
    w = U.he_uniform(7)
    L = U.unpack(w)
    bound = np.sqrt(6.0 / (9 * 10))
    assert np.abs(L["enc0a"][0]).max() <= bound and np.abs(L["enc0a"][0]).max() > 0.9 * bound
    assert not L["enc0a"][1].any()
    assert np.abs(L["out"][0]).max() <= 0.1 * np.sqrt(6.0 / 32) * 1.0001


def test_conv_closed_forms():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 6, 5, 4))
    # A kernel that picks channel 2 of the upper-left neighbour: out[y, x] = x[y - 1, x - 1, 2].
    W = np.zeros((1, 9, 4))
    W[0, 0, 2] = 1.0
    out = U.conv(x, W, np.array([0.5]), 9)
    expect = np.full((2, 6, 5), 0.5)
    expect[:, 1:, 1:] += x[:, :-1, :-1, 2]
    assert np.allclose(out[..., 0], expect)


def test_adjoint_pairs():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 8, 6, 3))
    y = rng.normal(size=(1, 4, 3, 3))
    assert np.isclose((U.pool(x) * y).sum(), (x * U.pool_backward(y)).sum())
    assert np.isclose((U.up(y) * x).sum(), (y * U.up_backward(x)).sum())
    W = rng.normal(size=(5, 9, 3))
    dy = rng.normal(size=(1, 8, 6, 5))
    _, _, dx = U.conv_backward(x, W, dy, 9)
    assert np.isclose((U.conv(x, W, np.zeros(5), 9) * dy).sum(), (x * dx).sum())


@pytest.mark.parametrize("loss", ["l2", "l1"])
def test_backward_matches_central_differences(loss):
    rng = np.random.default_rng(3)
    p = U.he_uniform(11).astype(np.float64)
    L = U.unpack(p)
    # Non-zero biases so that every layer's bias gradient is exercised.
    for name in L:
        L[name] = (L[name][0], rng.normal(scale=0.05, size=L[name][1].shape))
    p = U.pack(L)
    x = rng.uniform(0, 1, size=(1, 16, 16, U.IN_CH))
    t = rng.uniform(0, 2, size=(1, 16, 16, 3))
    val, g, _ = U.loss_and_grad(p, x, t, loss)
    # One weight and one bias from every layer.
    idx, k = [], 0
    for nw, nb in U.layer_sizes():
        idx += [k + int(rng.integers(0, nw)), k + nw + int(rng.integers(0, nb))]
        k += nw + nb
    h = 1e-6
    bad = 0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (U.loss_and_grad(pp, x, t, loss)[0] - U.loss_and_grad(pm, x, t, loss)[0]) / (2 * h)
        if not np.isclose(fd, g[i], rtol=1e-4, atol=1e-9):
            bad += 1
    # l1 has kinks where a residual crosses zero; l2 must agree everywhere.
    assert bad == 0 if loss == "l2" else bad <= 2
