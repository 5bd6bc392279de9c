"""The radiance cache on its configuration: the 4M-triangle wavefront scene (2000 instanced
icospheres, displaced ground, 64-primitive SDF cluster) at 480x270, depth 8, with the cache
trained online from 1/16 of the paths (2^20-query inference and 2^16-sample training per frame at
this size and 16 samples per pixel).

The cache runs with the renderer's defaults: render paths end in it at vertex 2, training paths
(one in 16, stratified) at vertex 4 with the cache's prediction as their tail (self-training), 8
Adam steps per 2^21 paths.

* Bias: after 16 warm frames, the cache's 256-sample render is compared with a 4096-sample plain
  reference (binary64 over disjoint 256-sample chunks): its error may exceed the plain 256-sample
  render's (the reference's own noise included) by at most 25 %, and its mean radiance stays
  within 1.5 % of the reference's.
* Equal time: at the time the plain tracer takes for 16 and for 64 samples per pixel, the warm
  cache's image has a lower MSE than plain's, cache + denoiser a lower MSE than plain + denoiser,
  and cache + denoiser at least 2.5x lower than plain (the bench's equal-time study, which fits
  each sample count to the budget, measures 3.4-3.5x at 64 samples; the bar leaves room for the
  timing noise of single renders).
"""
import sys
import time

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes

pytestmark = pytest.mark.gpu
W, H, DEPTH = 480, 270, 8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(ensure_built):
    sc = scenes.wavefront_scene()
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(1.0)
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    s.enqueue_sdf(s.push_sdf(sc["sdf_tape"]), sc["sdf_transform"])
    s.prepare()
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)

    def renderer():
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(DEPTH)
        return r

    say("reference: 4096 spp in 16 chunks")
    rr = renderer()
    acc = np.zeros((H, W, 3))
    for k in range(16):
        rr.set_sample_base(1 << 20 | (256 * k))
        rr.render(s, W, H, 256)
        acc += rr.image()[..., :3]
    ref = acc / 16
    cache = m.RadianceCache(seed=1)
    cache.set_bounds((-12.5, -7.0, -9.5), (12.5, 6.0, 9.5))
    warm = renderer()
    warm.set_cache(cache, mode=2)
    say("warming the cache: 16 frames of 16 spp")
    for _ in range(16):
        warm.render(s, W, H, 16)
    yield s, renderer, ref, cache
    s.close()


def mse(img, ref):
    return float(((img[..., :3].astype(np.float64) - ref) ** 2).mean())


def median_time(fn, n=3):
    ts = []
    for _ in range(n):
        m.synchronize()
        t0 = time.perf_counter()
        fn()
        m.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def test_cache_bias_is_bounded(setup):
    s, renderer, ref, cache = setup
    plain = renderer()
    plain.render(s, W, H, 256)
    cr = renderer()
    cr.set_cache(cache, mode=2)
    cr.render(s, W, H, 256)
    e_plain, e_cache = mse(plain.image(), ref), mse(cr.image(), ref)
    mean_ref, mean_cache = ref.mean(), cr.image()[..., :3].mean()
    say("256 spp MSE vs 4096-spp reference: plain %.3e, cache %.3e; mean %.4f vs %.4f" % (
        e_plain, e_cache, mean_cache, mean_ref))
    assert np.isfinite(cr.image()).all()
    assert e_cache <= 1.25 * e_plain
    assert abs(mean_cache - mean_ref) <= 0.015 * mean_ref


@pytest.mark.parametrize("budget", [16, 64])
def test_cache_gains_at_equal_time(setup, budget):
    s, renderer, ref, cache = setup
    import os
    weights = os.path.join(os.path.dirname(m.__file__), "data", "denoiser_unet.npy")
    den = m.Denoiser()
    den.load(weights)
    plain, cr = renderer(), renderer()
    cr.set_cache(cache, mode=2)
    plain.render(s, W, H, 16)
    cr.render(s, W, H, 16)
    T = median_time(lambda: (plain.invalidate(), plain.render(s, W, H, budget)))
    t_dn = median_time(lambda: plain.denoise(den))
    t_c = median_time(lambda: (cr.invalidate(), cr.render(s, W, H, budget)))
    # Sample counts that fit the budget, from the measured per-sample times (denoiser included).
    spp_pd = max(1, int(budget * (T - t_dn) / T))
    spp_c = max(1, int(budget * T / t_c))
    spp_cd = max(1, int(budget * (T - t_dn) / t_c))
    say("budget %d spp = %.1f ms; denoiser %.2f ms; cache %.1f ms per %d spp -> plain+den %d, cache %d, "
        "cache+den %d spp" % (budget, T * 1e3, t_dn * 1e3, t_c * 1e3, budget, spp_pd, spp_c, spp_cd))
    plain.invalidate()
    plain.render(s, W, H, budget)
    e_plain = mse(plain.image(), ref)
    plain.invalidate()
    plain.render(s, W, H, spp_pd)
    e_plain_den = mse(plain.denoise(den), ref)
    cr.invalidate()
    cr.render(s, W, H, spp_c)
    e_cache = mse(cr.image(), ref)
    cr.invalidate()
    cr.render(s, W, H, spp_cd)
    e_cache_den = mse(cr.denoise(den), ref)
    say("MSE: plain %.3e, cache %.3e (x%.2f); plain+den %.3e, cache+den %.3e (x%.2f)" % (
        e_plain, e_cache, e_plain / e_cache, e_plain_den, e_cache_den, e_plain_den / e_cache_den))
    say("plain / (cache+den) = %.2f" % (e_plain / e_cache_den))
    assert e_cache < e_plain
    assert e_cache_den < e_plain_den
    assert e_plain / e_cache_den >= 2.5
