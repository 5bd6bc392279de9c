"""Post-process passes: the oracle pinned by closed forms (CPU), the GPU passes against it."""
import numpy as np
import pytest

import myrt_amd as m
from oracle import post_ref as P


def test_filmic_closed_form():
    assert P.filmic(0.004) == 0.0 and P.filmic(0.0) == 0.0
    x = 1.004  # X = 1: (6.2 + 0.5) / (6.2 + 1.7 + 0.06), to the power 1.3
    assert np.isclose(P.filmic(x), (6.7 / 7.96) ** 1.3)
    assert np.isclose(P.filmic(1e9), 1.0)
    xs = np.linspace(0, 10, 1000)
    assert np.all(np.diff(P.filmic(xs)) >= 0)


def test_bilateral_closed_forms():
    const = np.full((12, 10, 4), 0.3)
    assert np.allclose(P.bilateral(const, 3.0, 0.2)[..., :3], 0.3)
    # A huge range sigma leaves the spatial Gaussian: one impulse spreads as the outer product of
    # normpdf(j, sigma) weights, renormalised by the window (the impulse is far from the border).
    img = np.zeros((21, 21, 4))
    img[10, 10, :3] = 1.0
    out = P.bilateral(img, 2.0, 1e6)
    k = P.normpdf(np.arange(-4, 5, dtype=np.float64), 2.0)
    expect = k[4] * k[4] / (k.sum() ** 2)
    assert np.isclose(out[10, 10, 0], expect)
    assert np.isclose(out[10, 12, 0], k[2] * k[4] / (k.sum() ** 2))


def test_fxaa_closed_forms():
    const = np.full((8, 9, 4), 0.7)
    assert np.allclose(P.fxaa(const)[..., :3], 0.7)
    # A vertical edge: every output lies between the two sides.
    img = np.zeros((16, 16, 4))
    img[:, 8:, :3] = 1.0
    out = P.fxaa(img)
    assert np.all(out[..., :3] >= -1e-12) and np.all(out[..., :3] <= 1 + 1e-12)
    assert np.allclose(out[:, :5, :3], 0) and np.allclose(out[:, 12:, :3], 1)


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["bilateral", "tonemap", "fxaa", "chain"])
def test_post_passes_match_oracle(ensure_built, op):
    rng = np.random.default_rng(4)
    img = np.concatenate([rng.gamma(1.0, 0.5, (37, 53, 3)), np.ones((37, 53, 1))], -1).astype(np.float32)
    img[5:20, 10:30, :3] *= 3.0  # some HDR values above 1
    if op == "bilateral":
        got, ref = m.post_process(img, [m.POST_BILATERAL], 2.5, 0.3), P.bilateral(img, 2.5, 0.3)
    elif op == "tonemap":
        got, ref = m.post_process(img, [m.POST_TONEMAP_FILMIC]), P.tonemap(img)
    elif op == "fxaa":
        got, ref = m.post_process(img, [m.POST_FXAA]), P.fxaa(img)
    else:
        got = m.post_process(img, [m.POST_BILATERAL, m.POST_TONEMAP_FILMIC, m.POST_FXAA], 10.0, 0.05)
        ref = P.fxaa(P.tonemap(P.bilateral(img, 10.0, 0.05)))
    assert np.abs(got - ref).max() < 2e-5
