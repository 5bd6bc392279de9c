"""The FP64 oracle's material model pinned by the mathematics (no GPU).

The render parity tests compare the GPU with the oracle using the same sampler and the same
BSDF, so a mistake the two share would pass them. These tests pin the oracle's lobes
(oracle/pathtrace_ref.cpp: make_lobes, eval_bsdf, sample_bsdf) against published forms and
sampling theory instead:

* pointwise: the GGX lobe equals Walter et al. 2007's D in its slope form (eq. 33) times Heitz
  2014's Smith G1 in its Lambda form, over 4 n.v n.l, with alpha = roughness^2 (ggx.glsl:9-15,
  29-35; pbr.glsl:72-87); its density is D n.h / (4 o.h) (Walter eq. 14); the Lambert lobe and
  Schlick's Fresnel (f0 = ((ior-1)/(ior+1))^2 mixed to 1 by metallic, pbr.glsl:29-34) weigh them;
* sampling: chi-square tests of sampled directions against the oracle's own density at
  alpha = 0.0025, 0.09 and 1, binned in half-vector space (and in direction space for the
  GGX + Lambert mixture); the GGX half-vector law P(tan^2 theta_h <= x) = x / (alpha^2 + x)
  holds for alpha = roughness^2 and is rejected for alpha = roughness;
* reciprocity of the reflection lobe, and its directional albedo: the sampler's estimator agrees
  with a quadrature of f cos, stays <= 1, and is ~1 for a near-mirror;
* the smooth dielectric: Snell's law, total internal reflection beyond the critical angle, the
  reflect/refract choice with probability Schlick F, and a white furnace: a glass sphere in a
  uniform environment returns the environment.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as orc
from myrt_amd import scenes

NRM = np.array([0.0, 0.0, 1.0])


def material(rgb=(1.0, 1.0, 1.0), ior=1.5, roughness=0.5, metallic=0.0, transmission=0.0):
    m = np.zeros(1, orc.MATERIAL_DTYPE)
    m[0] = (scenes._rgba8(rgb), ior, roughness, metallic, transmission, 0, (0, 0))
    return m


def f32(x):
    """Material parameters are binary32 in the record; the published forms take the same value."""
    return float(np.float32(x))


def direction(theta, phi=0.0):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


def hemisphere(rng, n):
    """Directions uniform in cos(theta) over the upper hemisphere, away from the horizon."""
    mu = rng.uniform(0.02, 1.0, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - mu * mu)
    return np.stack([s * np.cos(phi), s * np.sin(phi), mu], 1)


# Published forms, written independently of the oracle's algebra.
def walter_D(cos_h, alpha):
    """GGX distribution, Walter et al. 2007 eq. 33: alpha^2 / (pi cos^4 (alpha^2 + tan^2)^2)."""
    c2 = cos_h * cos_h
    tan2 = (1.0 - c2) / c2
    return alpha ** 2 / (np.pi * c2 * c2 * (alpha ** 2 + tan2) ** 2)


def heitz_G1(cos_v, alpha):
    """Smith masking for GGX, Heitz 2014: 1 / (1 + Lambda), Lambda = (-1 + sqrt(1 + a^2 tan^2)) / 2."""
    tan2 = (1.0 - cos_v * cos_v) / (cos_v * cos_v)
    lam = (-1.0 + np.sqrt(1.0 + alpha ** 2 * tan2)) / 2.0
    return 1.0 / (1.0 + lam)


def schlick(cos_v, ior, metallic=0.0):
    f0 = ((1.0 - ior) / (1.0 + ior)) ** 2
    f0 = f0 + (1.0 - f0) * metallic
    return f0 + (1.0 - f0) * (1.0 - cos_v) ** 5


def samples(mat, wo, n, seed, entering=True):
    u = np.random.default_rng(seed).random((n, 3))
    return orc.bsdf_sample(mat, NRM, wo, u, entering)


# ---- pointwise against the published forms ---------------------------------------------------

@pytest.mark.parametrize("rough", [0.05, 0.3, 0.7, 1.0])
def test_ggx_lobe_equals_published_forms(ensure_built, rough):
    rng = np.random.default_rng(11)
    alpha = f32(rough) ** 2
    mat = material(metallic=1.0, roughness=rough)  # f0 = 1: F = 1, pure GGX, tint = albedo = 1
    for _ in range(8):
        wo = hemisphere(rng, 1)[0]
        wi = hemisphere(rng, 256)
        f, pdf = orc.bsdf_eval(mat, NRM, wo, wi)
        h = wo + wi
        h /= np.linalg.norm(h, axis=1, keepdims=True)
        nv, nl, nh, vh = wo[2], wi[:, 2], h[:, 2], h @ wo
        want_f = walter_D(nh, alpha) * heitz_G1(nv, alpha) * heitz_G1(nl, alpha) / (4 * nv * nl)
        want_pdf = walter_D(nh, alpha) * nh / (4 * vh)
        np.testing.assert_allclose(f[:, 0], want_f, rtol=1e-9)
        np.testing.assert_allclose(f[:, 1], want_f, rtol=1e-9)
        np.testing.assert_allclose(pdf, want_pdf, rtol=1e-9)


def test_lobe_mixture_is_schlick_weighted_ggx_and_lambert(ensure_built):
    rng = np.random.default_rng(12)
    for rough, ior, metallic, rgb in ((0.4, 1.5, 0.0, (0.8, 0.4, 0.2)), (0.8, 1.33, 0.25, (0.2, 0.6, 0.9))):
        alpha = f32(rough) ** 2
        mat = material(rgb=rgb, ior=ior, roughness=rough, metallic=metallic)
        albedo = np.array([int(round(c * 255)) / 255.0 for c in rgb])
        for _ in range(6):
            wo = hemisphere(rng, 1)[0]
            wi = hemisphere(rng, 128)
            f, pdf = orc.bsdf_eval(mat, NRM, wo, wi)
            h = wo + wi
            h /= np.linalg.norm(h, axis=1, keepdims=True)
            nv, nl, nh, vh = wo[2], wi[:, 2], h[:, 2], h @ wo
            F = schlick(nv, f32(ior), metallic)
            spec = walter_D(nh, alpha) * heitz_G1(nv, alpha) * heitz_G1(nl, alpha) / (4 * nv * nl)
            tint = 1.0 + (albedo - 1.0) * np.sqrt(metallic)
            want = F * spec[:, None] * tint[None, :] + (1 - F) * (1 - metallic) / np.pi * albedo[None, :]
            np.testing.assert_allclose(f, want, rtol=1e-9)
            want_pdf = F * walter_D(nh, alpha) * nh / (4 * vh) + (1 - F) * nl / np.pi
            np.testing.assert_allclose(pdf, want_pdf, rtol=1e-9)


# ---- sampling against the density -------------------------------------------------------------

def gauss_grid(k_u, k_phi, order=8):
    """Gauss-Legendre nodes and weights over [0,1] x [0, 2 pi) split into k_u x k_phi bins:
    returns per-bin arrays (k_u, k_phi, order^2) of u, phi and weights (bin measure included)."""
    x, w = np.polynomial.legendre.leggauss(order)
    x, w = (x + 1) / 2, w / 2
    iu, ip = np.meshgrid(np.arange(k_u), np.arange(k_phi), indexing="ij")
    u = (iu[..., None, None] + x[None, None, :, None]) / k_u
    ph = (ip[..., None, None] + x[None, None, None, :]) / k_phi * 2 * np.pi
    wt = (w[:, None] * w[None, :])[None, None] / k_u * (2 * np.pi / k_phi)
    u, ph = np.broadcast_arrays(u, ph)
    return u.reshape(k_u, k_phi, -1), ph.reshape(k_u, k_phi, -1), np.broadcast_to(wt, u.shape).reshape(k_u, k_phi, -1)


def half_vector_bins(alpha, k_u=20, k_phi=16):
    """Half vectors h(u, phi) with tan^2(theta_h) = alpha^2 u / (1 - u) (the GGX CDF variable u is
    uniform for GGX-distributed h) and the solid-angle measure d omega_h per (u, phi) weight."""
    u, ph, wt = gauss_grid(k_u, k_phi)
    s = np.sqrt(u / (1 - u))
    th = np.arctan(alpha * s)
    dth_du = alpha / (2 * s * (1 - u) ** 2) / (1 + alpha ** 2 * s * s)
    h = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    return h, wt * np.sin(th) * dth_du


def half_vector_coords(h, alpha):
    tan2 = (h[:, 0] ** 2 + h[:, 1] ** 2) / h[:, 2] ** 2
    u = tan2 / (alpha ** 2 + tan2)
    phi = np.mod(np.arctan2(h[:, 1], h[:, 0]), 2 * np.pi)
    return u, phi


def chi2_pvalue(observed, expected, min_expected=5.0):
    observed = np.asarray(observed, np.float64).ravel()
    expected = np.asarray(expected, np.float64).ravel()
    small = expected < min_expected
    o = np.append(observed[~small], observed[small].sum())
    e = np.append(expected[~small], expected[small].sum())
    if e[-1] < min_expected:  # the pooled small bins are still small: merge them into the largest
        big = int(np.argmax(e[:-1]))
        o[big] += o[-1]
        e[big] += e[-1]
        o, e = o[:-1], e[:-1]
    keep = e > 0
    assert o[~keep].sum() == 0, "samples where the density is zero"
    o, e = o[keep], e[keep]
    chi2 = float(((o - e) ** 2 / e).sum())
    return float(stats.chi2.sf(chi2, len(o) - 1))


@pytest.mark.parametrize("rough", [0.05, 0.3, 1.0])  # alpha = 0.0025, 0.09, 1
@pytest.mark.parametrize("theta_o", [0.25, 1.2])
def test_ggx_samples_follow_their_density(ensure_built, rough, theta_o):
    """Directions drawn by the sampler, binned by their half vector, against the expected counts
    from the oracle's density integrated over each bin: p(wi) |d wi / d h| = p(wi) 4 |o.h|
    (the reflection Jacobian, Walter 2007 eq. 14). A density without the 1/(4 o.h) factor, or with
    another alpha than the sampler's, fails."""
    n = 200000
    alpha = f32(rough) ** 2
    mat = material(metallic=1.0, roughness=rough)
    wo = direction(theta_o, 0.4)
    smp = samples(mat, wo, n, seed=21)
    ok = smp["side"] == 1
    h = wo[None, :] + smp["wi"][ok]
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    u, phi = half_vector_coords(h, alpha)
    k_u, k_phi = 20, 16
    obs, _, _ = np.histogram2d(u, phi, bins=[k_u, k_phi], range=[[0, 1], [0, 2 * np.pi]])
    hb, dw = half_vector_bins(alpha, k_u, k_phi)
    vh = hb @ wo
    wi = 2 * vh[..., None] * hb - wo
    _, pdf = orc.bsdf_eval(mat, NRM, wo, wi.reshape(-1, 3))
    pdf = pdf.reshape(vh.shape)
    expected = n * (pdf * 4 * np.abs(vh) * dw).sum(-1)
    p_valid = expected.sum() / n
    assert 0.3 < p_valid <= 1.0 + 1e-6  # alpha = 1 reflects about half of the samples below the horizon
    pval = chi2_pvalue(np.append(obs.ravel(), (~ok).sum()), np.append(expected.ravel(), n - expected.sum()))
    assert pval > 1e-4, (rough, theta_o, pval)


def test_ggx_lambert_mixture_samples_follow_their_density(ensure_built):
    """A rough dielectric-free plastic: lobe choice by Schlick F, GGX and cosine-weighted Lambert
    samples, binned in direction space (cos^2 theta, phi) against the mixture density."""
    n = 200000
    mat = material(rgb=(0.6, 0.5, 0.4), ior=1.5, roughness=0.9)
    wo = direction(0.9, 1.0)
    smp = samples(mat, wo, n, seed=22)
    ok = smp["side"] == 1
    wi = smp["wi"][ok]
    k_u, k_phi = 16, 16
    obs, _, _ = np.histogram2d(wi[:, 2] ** 2, np.mod(np.arctan2(wi[:, 1], wi[:, 0]), 2 * np.pi), bins=[k_u, k_phi],
                               range=[[0, 1], [0, 2 * np.pi]])
    q, ph, wt = gauss_grid(k_u, k_phi)
    mu = np.sqrt(q)  # u = cos^2: d omega = d(cos) d phi = du / (2 cos) d phi
    s = np.sqrt(1 - q)
    d = np.stack([s * np.cos(ph), s * np.sin(ph), mu], -1)
    _, pdf = orc.bsdf_eval(mat, NRM, wo, d.reshape(-1, 3))
    expected = n * (pdf.reshape(q.shape) * wt / (2 * mu)).sum(-1)
    pval = chi2_pvalue(np.append(obs.ravel(), (~ok).sum()), np.append(expected.ravel(), n - expected.sum()))
    assert pval > 1e-4, pval


@pytest.mark.parametrize("rough", [0.3, 0.7])
def test_half_vectors_follow_ggx_with_alpha_roughness_squared(ensure_built, rough):
    """Walter 2007's sampling law: for GGX with parameter alpha, u = tan^2/(alpha^2 + tan^2) of the
    half vector is uniform on [0, 1). It holds for alpha = roughness^2 (ggx.glsl:29-35) and is
    rejected for alpha = roughness, the plausible slip."""
    n = 100000
    mat = material(metallic=1.0, roughness=rough)
    wo = direction(0.7)
    smp = samples(mat, wo, n, seed=23)
    h = wo[None, :] + smp["wi"]
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    h *= np.sign(h[:, 2:3])  # a half vector with o.h < 0 reflects to -h's side
    u_right, _ = half_vector_coords(h, f32(rough) ** 2)
    u_wrong, _ = half_vector_coords(h, f32(rough))
    assert stats.kstest(u_right, "uniform").pvalue > 1e-4
    assert stats.kstest(u_wrong, "uniform").pvalue < 1e-12


# ---- reciprocity and albedo -----------------------------------------------------------------

@pytest.mark.parametrize("rough", [0.05, 0.3, 1.0])
def test_ggx_reflection_is_reciprocal(ensure_built, rough):
    rng = np.random.default_rng(31)
    mat = material(rgb=(0.9, 0.7, 0.5), metallic=1.0, roughness=rough)
    a = hemisphere(rng, 400)
    b = hemisphere(rng, 400)
    fab = np.array([orc.bsdf_eval(mat, NRM, a[k], b[k:k + 1])[0][0] for k in range(len(a))])
    fba = np.array([orc.bsdf_eval(mat, NRM, b[k], a[k:k + 1])[0][0] for k in range(len(a))])
    np.testing.assert_allclose(fab, fba, rtol=1e-12, atol=0)


def albedo_quadrature(mat, wo, alpha):
    """rho(wo) = int f(wo, wi) cos(theta_i) d wi, integrated over half vectors (Jacobian 4 |o.h|)."""
    hb, dw = half_vector_bins(alpha, 48, 32)
    vh = hb @ wo
    wi = 2 * vh[..., None] * hb - wo
    f, _ = orc.bsdf_eval(mat, NRM, wo, wi.reshape(-1, 3))
    fcos = f[:, 0].reshape(vh.shape) * np.maximum(wi[..., 2], 0)
    return float((fcos * 4 * np.abs(vh) * dw).sum())


@pytest.mark.parametrize("rough", [0.05, 0.3, 1.0])
def test_ggx_albedo_quadrature_matches_the_sampling_estimator(ensure_built, rough):
    n = 200000
    alpha = f32(rough) ** 2
    mat = material(metallic=1.0, roughness=rough)
    for theta_o in (0.0, 0.5, 1.0, 1.4):
        wo = direction(theta_o, 0.3)
        smp = samples(mat, wo, n, seed=41)
        w = np.where(smp["side"] == 1, smp["weight"][:, 0], 0.0)
        est, err = w.mean(), w.std() / np.sqrt(n)
        quad = albedo_quadrature(mat, wo, alpha)
        assert abs(est - quad) < 4 * err + 1e-4, (rough, theta_o, est, quad)
        assert quad <= 1.0 + 1e-6  # a microfacet lobe with F = 1 reflects at most what arrives
        if rough == 0.05 and theta_o <= 1.0:
            assert quad > 0.99  # near-mirror: shadowing and masking take almost nothing


# ---- the smooth dielectric --------------------------------------------------------------------

@pytest.mark.parametrize("theta_o", [0.0, 0.3, 0.8, 1.3])
def test_dielectric_refracts_by_snell_with_schlick_selection(ensure_built, theta_o):
    ior = 1.5
    n = 100000
    mat = material(ior=ior, roughness=0.05, transmission=1.0)
    wo = direction(theta_o, 2.0)
    smp = samples(mat, wo, n, seed=51, entering=True)
    refr = smp["side"] == 2
    # Reflect with probability F (Schlick, f0 = ((1 - ior)/(1 + ior))^2), else refract.
    F = schlick(np.cos(theta_o), ior)
    k = refr.sum()
    assert abs(k - n * (1 - F)) < 5 * np.sqrt(n * F * (1 - F)) + 1, (k, n * (1 - F))
    assert np.all(smp["delta"][refr] == 1)
    np.testing.assert_array_equal(smp["weight"][refr], 1.0)  # albedo 1
    wi = smp["wi"][refr]
    np.testing.assert_allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-12)
    assert np.all(wi[:, 2] < 0)  # the other side
    sin_t = np.linalg.norm(np.cross(wi, NRM), axis=1)
    np.testing.assert_allclose(sin_t, np.sin(theta_o) / ior, atol=1e-12)  # Snell, eta = 1 / ior
    # In the plane of incidence, on the far side of the normal.
    np.testing.assert_allclose(wi @ np.cross(wo, NRM), 0.0, atol=1e-12)
    if theta_o > 0:
        assert np.all(wi[:, :2] @ wo[:2] <= 0)


def test_dielectric_leaving_has_total_internal_reflection(ensure_built):
    ior = 1.5
    crit = np.arcsin(1 / ior)
    mat = material(ior=ior, roughness=0.05, transmission=1.0)
    u = np.array([[0.999999, 0.3, 0.6]])  # the transmission lobe (ul above F)
    for theta_o in (0.2, 0.6, crit - 0.01):
        wo = direction(theta_o)
        s = orc.bsdf_sample(mat, NRM, wo, u, entering=False)
        assert s["side"][0] == 2
        sin_t = np.linalg.norm(np.cross(s["wi"][0], NRM))
        assert abs(sin_t - ior * np.sin(theta_o)) < 1e-12  # eta = ior from inside
    for theta_o in (crit + 0.01, 1.0, 1.4):
        wo = direction(theta_o, 0.7)
        s = orc.bsdf_sample(mat, NRM, wo, u, entering=False)
        assert s["side"][0] == 1 and s["delta"][0] == 1
        mirror = 2 * wo[2] * NRM - wo
        np.testing.assert_allclose(s["wi"][0], mirror, atol=1e-12)


def test_glass_sphere_in_a_white_furnace_returns_the_environment(ensure_built):
    """A smooth glass icosphere (flat facets, ior 1.5, white) in a uniform environment of radiance
    1: reflection with probability F and refraction otherwise conserve energy, so every pixel
    converges to 1 (the GGX reflection lobe at the alpha floor loses a fraction of a per mille)."""
    fi, fp, _ = scenes.icosphere(3)
    mat = material(ior=1.5, roughness=0.0, transmission=1.0)
    R = orc.OracleRender([(fi, fp, None)], [(0, np.eye(4, dtype=np.float32), 0)], mat, None, 1.0)
    R.set_cubemap(np.ones((6, 16, 16, 3), np.float32))
    W = H = 12
    view, proj = scenes.camera_matrices((0.0, 0.0, 3.2), (0.0, 0.0, 0.0), 40.0, 1.0)
    img = R.render(view, proj, W, H, 256, 48)
    assert np.isfinite(img).all()
    mean = img.mean()
    assert abs(mean - 1.0) < 0.01, mean
    assert np.all(np.abs(img - 1.0) < 0.08), (img.min(), img.max())
    # A glass that absorbs 10 % per refraction (albedo 0.9) falls below the bar: it separates an
    # energy loss of that size.
    lossy = orc.OracleRender([(fi, fp, None)], [(0, np.eye(4, dtype=np.float32), 0)],
                             material(rgb=(0.9, 0.9, 0.9), ior=1.5, roughness=0.0, transmission=1.0), None, 1.0)
    lossy.set_cubemap(np.ones((6, 16, 16, 3), np.float32))
    assert lossy.render(view, proj, W, H, 64, 48).mean() < 0.97
