"""Denoiser U-Net on the GPU against the FP64 oracle (oracle/unet_ref.py).

Bars: initial parameters equal; fp32 training forward within 1e-4 of the oracle (relative to the
largest output); fp16 inference forward within 3e-2; the loss gradient within 1e-3 (relative
L2 norm) and Adam steps within 1e-5; training lowers the loss; the renderer's primary-hit features
match a closed-form scene; denoising a rendered image runs end to end."""
import sys

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes
from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def raw_batch(B, H, W, seed):
    """Smooth random feature images (radiance 0..3, albedo, unit normals, distances, some misses)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.linspace(0, 1, H), np.linspace(0, 1, W), indexing="ij")
    raw = np.zeros((B, H, W, 10), np.float32)
    for b in range(B):
        f = rng.uniform(1, 4, 6)
        for c in range(3):
            raw[b, ..., c] = 1.5 + 1.5 * np.sin(f[c] * xx * 6 + f[c + 3] * yy * 4 + c)
        raw[b, ..., 3:6] = rng.uniform(0.1, 0.9, 3)
        n = rng.normal(size=(H, W, 3))
        raw[b, ..., 6:9] = n / np.linalg.norm(n, axis=-1, keepdims=True)
        raw[b, ..., 9] = 2 + 8 * yy
        raw[b, :2, :, 9] = np.inf
    raw[..., :3] += rng.normal(scale=0.3, size=raw[..., :3].shape).astype(np.float32)
    return raw


def oracle_x(raw):
    return U.prep(raw[..., 0:3], raw[..., 3:6], raw[..., 6:9], raw[..., 9])


def test_initial_params_match_oracle(ensure_built):
    d = m.Denoiser(seed=21)
    assert d.n_params == U.N_PARAMS
    assert np.array_equal(d.params(), U.he_uniform(21))


def test_forward_fp32_and_fp16_match_oracle(ensure_built):
    d = m.Denoiser(seed=3)
    p = d.params().astype(np.float64)
    L = U.unpack(p)
    rng = np.random.default_rng(4)
    for name in L:  # non-zero biases and a larger output layer, so that every path shows
        W, b = L[name]
        L[name] = (W * (5.0 if name == "out" else 1.0), rng.normal(scale=0.05, size=b.shape))
    p = U.pack(L)
    d.set_params(p.astype(np.float32))
    raw = raw_batch(2, 32, 48, seed=5)
    say("denoiser forward 2x32x48")
    ref, _ = U.forward(p.astype(np.float32).astype(np.float64), oracle_x(raw))
    y32 = d.forward(raw, half=False)
    scale = np.abs(ref).max()
    err32 = np.abs(y32 - ref).max() / scale
    y16 = d.forward(raw, half=True)
    err16 = np.abs(y16 - ref).max() / scale
    say("max relative error fp32 %.2e fp16 %.2e (scale %.3f)" % (err32, err16, scale))
    assert scale > 0.05
    assert err32 < 1e-4
    assert err16 < 3e-2


def test_fp16_tiled_forward_matches_fp32_at_full_size(ensure_built):
    """At 1024x2048 the fp16 inference takes the 8-row tiled convolutions (levels 0 and 1, the
    latter as two 32-channel groups) and the 4-row ones below; it must stay within the fp16 bar of
    the fp32 forward (untiled kernels, pinned to the oracle at small size above)."""
    d = m.Denoiser(seed=3)
    p = d.params()
    rng = np.random.default_rng(6)
    L = U.unpack(p.astype(np.float64))
    for name in L:
        W, b = L[name]
        L[name] = (W * (5.0 if name == "out" else 1.0), rng.normal(scale=0.05, size=b.shape))
    d.set_params(U.pack(L).astype(np.float32))
    raw = raw_batch(1, 1024, 2048, seed=7)
    say("denoiser forward 1x1024x2048 fp32 and fp16")
    y32 = d.forward(raw, half=False)
    y16 = d.forward(raw, half=True)
    scale = np.abs(y32).max()
    err16 = np.abs(y16 - y32).max() / scale
    say("max relative difference fp16 vs fp32 %.2e (scale %.3f)" % (err16, scale))
    assert scale > 0.05 and np.isfinite(y16).all()
    assert err16 < 3e-2


@pytest.mark.parametrize("H,W", [(40, 72), (1000, 520)])
def test_denoise_fused_steps_match_the_separate_passes(ensure_built, H, W):
    """denoise() runs the 1x1 output layer and the finish step inside the last convolution
    (40x72 takes the 4-row tiles, 1000x520 the 8-row ones; neither is a multiple of 16). It must
    give expm1(max(log1p(max(radiance, 0)) + y, 0)) with y the
    separate fp16 forward of the edge-padded features. The only difference is that the separate
    output layer rounds y to fp16, so in log space (log1p of the output) the two differ by at most
    half an fp16 ulp of y, 2^-11 |y|, plus fp32 rounding."""
    d = m.Denoiser(seed=3)
    rng = np.random.default_rng(12)
    L = U.unpack(d.params().astype(np.float64))
    for name in L:
        Wt, b = L[name]
        L[name] = (Wt * (5.0 if name == "out" else 1.0), rng.normal(scale=0.05, size=b.shape))
    d.set_params(U.pack(L).astype(np.float32))
    raw = raw_batch(1, H, W, seed=13)
    say("denoise %dx%d: fused steps against the separate passes" % (W, H))
    out = d.denoise(raw[0])
    Hp, Wp = (H + 15) // 16 * 16, (W + 15) // 16 * 16
    padded = np.pad(raw, ((0, 0), (0, Hp - H), (0, Wp - W), (0, 0)), mode="edge")
    y = d.forward(padded, half=True)[0, :H, :W].astype(np.float64)
    ref_log = np.maximum(np.log1p(np.maximum(raw[0, ..., :3].astype(np.float64), 0)) + y, 0)
    diff = np.abs(np.log1p(out[..., :3].astype(np.float64)) - ref_log)
    bound = 2.0 ** -11 * np.abs(y) + 1e-5 * (1 + ref_log)
    say("largest log-space difference %.2e, largest share of the fp16 bound %.2f" % (diff.max(), (diff / bound).max()))
    assert np.isfinite(out).all() and np.all(out[..., 3] == 1.0)
    assert np.abs(y).max() > 0.05  # the residual is not negligible
    assert np.all(diff <= bound)


def test_gradient_and_adam_match_oracle(ensure_built):
    d = m.Denoiser(seed=8)
    p = d.params().astype(np.float64)
    raw = raw_batch(2, 32, 32, seed=9)
    target = np.clip(raw[..., :3] + np.random.default_rng(1).normal(scale=0.2, size=raw[..., :3].shape), 0, None)
    say("denoiser gradient 2x32x32")
    g, loss = d.gradient(raw, target)
    ref_loss, ref_g, _ = U.loss_and_grad(p, oracle_x(raw), target.astype(np.float64), "l1")
    rel = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    say("loss %.6f oracle %.6f, gradient relative error %.2e" % (loss, ref_loss, rel))
    assert abs(loss - ref_loss) < 1e-5 * max(1.0, ref_loss)
    assert rel < 1e-3
    # Every layer's weight and bias gradients are non-trivial and agree.
    k = 0
    for (nw, nb) in U.layer_sizes():
        for a, b_ in ((k, k + nw), (k + nw, k + nw + nb)):
            r = ref_g[a:b_]
            assert np.linalg.norm(r) > 0
            assert np.linalg.norm(g[a:b_] - r) <= 2e-3 * np.linalg.norm(r) + 1e-9
        k += nw + nb
    # Two Adam steps track the oracle's.
    d.set_adam(lr=1e-3)
    mo = np.zeros_like(p)
    vo = np.zeros_like(p)
    po = p.copy()
    for step in (1, 2):
        d.train(raw, target)
        _, go, _ = U.loss_and_grad(po, oracle_x(raw), target.astype(np.float64), "l1")
        po, mo, vo = U.adam(po, go, mo, vo, step)
    diff = np.abs(d.params() - po)
    say("after 2 Adam steps max |param - oracle| %.2e, above 1e-5: %d" % (diff.max(), (diff > 1e-5).sum()))
    # Adam moves each parameter by about lr * sign(g): components whose gradient is within fp32
    # summation error of zero may move the other way (at most 2 lr per step); all others agree.
    assert (diff > 1e-5).mean() < 1e-3
    assert diff.max() <= 4.0 * 1e-3 + 1e-6


def test_training_lowers_the_loss(ensure_built):
    d = m.Denoiser(seed=2)
    d.set_adam(lr=1e-3)
    rng = np.random.default_rng(0)
    clean = raw_batch(4, 64, 64, seed=11)
    clean[..., :3] = np.clip(clean[..., :3], 0, None)
    losses = []
    for it in range(100):
        raw = clean.copy()
        raw[..., :3] = clean[..., :3] * rng.gamma(2.0, 0.5, size=clean[..., :3].shape)  # multiplicative noise
        losses.append(d.train(raw, clean[..., :3]))
    say("loss first %.4f last %.4f" % (np.mean(losses[:5]), np.mean(losses[-5:])))
    assert np.isfinite(losses).all()
    assert np.mean(losses[-5:]) < 0.85 * np.mean(losses[:5])


def test_renderer_features_closed_form(ensure_built):
    gi, gp, gn = scenes.ground_quad(size=20.0, y=0.0)
    s = m.Scene()
    g = s.push_geometry(gi, gp, gn)
    mat = s.push_material(m.material((0.5, 0.25, 1.0), ior=1.0, roughness=1.0))
    s.set_environment(1.0)
    s.enqueue(g, None, mat)
    view, proj = scenes.camera_matrices((0, 5, 0.001), (0, 0, 0), 30.0, 1.0)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(2)
    r.render(s, 16, 16, 4)
    albedo, normal, depth = r.aux()
    assert np.allclose(albedo, np.array([128, 64, 255]) / 255.0, atol=1e-6)
    assert np.allclose(normal, [0, 1, 0], atol=1e-5)
    assert np.all((depth > 5.0 - 1e-3) & (depth < 5.0 / np.cos(np.radians(15 * 1.5))))
    raw = r.raw_features()
    assert raw.shape == (16, 16, 10)


def test_denoise_rendered_image(ensure_built):
    sc = scenes.sphere_field_scene(30, 2, seed=4)
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    for gg, mt, mi in sc["instances"]:
        s.enqueue(gids[gg], mt, mi)
    W, H = 72, 40  # not multiples of 16: padded inside
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(4)
    r.render(s, W, H, 2)
    d = m.Denoiser(seed=1)
    out = r.denoise(d)
    assert out.shape == (H, W, 4) and np.isfinite(out).all() and (out[..., :3] >= 0).all()
    # The device path equals the host-buffer path on the same features.
    out2 = d.denoise(r.raw_features())
    assert np.allclose(out, out2, rtol=1e-3, atol=1e-4)
    # The untrained network's residual is small: the output stays near its input.
    img = r.image()[..., :3]
    assert np.abs(np.log1p(out[..., :3]) - np.log1p(img)).mean() < 0.2


def test_trained_network_gains_on_a_held_out_scene(ensure_built):
    """The shipped weights (myrt_amd/data/denoiser_unet.npy, trained on training_scene seeds
    4-515 at 512^2, as myrt_amd/data/denoiser_unet.json records) on training_scene(100000), which
    they never saw: denoising a 1-spp render cuts the MSE against a 1024-spp render of the same scene at
    least 4x, a 16-spp render at least 2x."""
    import os
    weights = os.path.join(os.path.dirname(m.__file__), "data", "denoiser_unet.npy")
    sc = scenes.training_scene(100000)
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(sc.get("env", 1.0))
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    if "sdf_tape" in sc:
        s.enqueue_sdf(s.push_sdf(sc["sdf_tape"]), sc["sdf_transform"])
    res = 192
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.0)

    def renderer():
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(8)
        return r

    say("held-out scene: 1024-spp reference at %d^2" % res)
    ref_r = renderer()
    for _ in range(4):
        ref_r.render(s, res, res, 256)
    ref = ref_r.image()[..., :3].astype(np.float64)
    d = m.Denoiser()
    d.load(weights)
    for spp, need in ((1, 4.0), (16, 2.0)):
        r = renderer()
        r.render(s, res, res, spp)
        noisy = ((r.image()[..., :3] - ref) ** 2).mean()
        den = ((r.denoise(d)[..., :3] - ref) ** 2).mean()
        say("%d spp: noisy MSE %.3e, denoised %.3e (x%.1f)" % (spp, noisy, den, noisy / den))
        assert noisy / den >= need
    s.close()
