"""Post-process passes: the oracle pinned by closed forms (CPU), the GPU passes against it."""
import numpy as np
from scipy.ndimage import maximum_filter
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


def test_bilateral_shift_changes_nothing_where_the_reference_is_defined():
    # The reference's unshifted weights 0.39894 exp(-q / 2 s^2) / s * bZ * k_i * k_j give the same
    # normalised sum wherever they do not all underflow.
    rng = np.random.default_rng(3)
    img = np.concatenate([rng.uniform(0, 1.2, (9, 11, 3)), np.ones((9, 11, 1))], -1)
    sig, bs = 2.0, 0.3
    k = P.normpdf(np.arange(-4, 5, dtype=np.float64), sig)
    bz = 1.0 / P.normpdf(0.0, bs)
    lit = np.zeros((9, 11, 3))
    for y in range(9):
        for x in range(11):
            c = img[y, x, :3]
            z, acc = 0.0, np.zeros(3)
            for i in range(-4, 5):
                for j in range(-4, 5):
                    cc = np.clip(img[min(max(y + j, 0), 8), min(max(x + i, 0), 10), :3], 0, 1)
                    f = P.normpdf(np.sqrt(((cc - c) ** 2).sum()), bs) * bz * k[j + 4] * k[i + 4]
                    z += f
                    acc += f * cc
            lit[y, x] = acc / z
    assert np.allclose(P.bilateral(img, sig, bs)[..., :3], lit, rtol=1e-12, atol=0)
    # A centre far above every clamped neighbour (where the reference's weights underflow to 0/0)
    # averages the clamped neighbours: here all 1.
    assert np.array_equal(P.bilateral(np.full((6, 7, 4), 5.0), 10.0, 0.05)[..., :3], np.ones((6, 7, 3)))


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
    img[30:33, 40:44, :3] = 40.0  # a bright patch whose unshifted range weights underflow
    src = img  # the input of the last pass
    if op == "bilateral":
        got, ref = m.post_process(img, [m.POST_BILATERAL], 2.5, 0.3), P.bilateral(img, 2.5, 0.3)
    elif op == "tonemap":
        got, ref = m.post_process(img, [m.POST_TONEMAP_FILMIC]), P.tonemap(img)
    elif op == "fxaa":
        got, ref = m.post_process(img, [m.POST_FXAA]), P.fxaa(img)
    else:
        got = m.post_process(img, [m.POST_BILATERAL, m.POST_TONEMAP_FILMIC, m.POST_FXAA], 10.0, 0.05)
        src = P.tonemap(P.bilateral(img, 10.0, 0.05))
        ref = P.fxaa(src)
    assert np.isfinite(got).all()
    # fp32 against binary64. FXAA samples at fp32 texture coordinates up to 4 texels away, so its
    # error scales with the largest value it can reach: the bound is 2e-5 of the largest input
    # magnitude (at least 1) in the 17 x 17 window around the pixel.
    scale = maximum_filter(np.abs(src[..., :3]).max(-1), size=17, mode="nearest")
    err = np.abs(got - ref)[..., :3].max(-1) / np.maximum(1.0, scale)
    print("%s: max scaled error %.2e at %s" % (op, err.max(), np.unravel_index(err.argmax(), err.shape)))
    assert err.max() <= 2e-5
