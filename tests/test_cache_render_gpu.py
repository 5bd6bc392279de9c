"""Radiance cache inside the wavefront renderer.

Training paths are ordinary unbiased paths that also record their vertices, so a render that only
trains (mode 1) must equal the plain render; a render that ends paths in the cache (mode 2) traces
fewer rays and stays finite."""
import sys

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_scene():
    sc = scenes.sphere_field_scene(60, 2, seed=3, glass_every=7, metal_every=5)
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(1.0)
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    return sc, s


def renderer(sc, W, H, depth=6):
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    return r


def cache_for_scene():
    c = m.RadianceCache(seed=5)
    c.set_bounds((-8.5, -5.5, -8.5), (8.5, 7.5, 8.5))
    return c


def test_train_only_mode_equals_plain(ensure_built):
    sc, s = small_scene()
    W, H, spp = 64, 48, 8
    say("plain vs train-only %dx%d %d spp" % (W, H, spp))
    a = renderer(sc, W, H)
    a.render(s, W, H, spp)
    b = renderer(sc, W, H)
    c = cache_for_scene()
    b.set_cache(c, mode=1, train_fraction=0.25, train_steps=2)
    b.render(s, W, H, spp)
    ia, ib = a.image(), b.image()
    assert np.allclose(ia, ib, rtol=1e-4, atol=1e-6)
    assert a.stats()["rays"] == b.stats()["rays"]
    # The training steps changed the weights.
    c0 = cache_for_scene()
    assert not np.array_equal(c.params(), c0.params())


def test_cache_mode_ends_paths_early(ensure_built):
    sc, s = small_scene()
    W, H, spp = 64, 48, 8
    say("cache mode %dx%d %d spp" % (W, H, spp))
    a = renderer(sc, W, H)
    a.render(s, W, H, spp)
    c = cache_for_scene()
    t = renderer(sc, W, H)
    t.set_cache(c, mode=1)
    t.render(s, W, H, 4)  # trains the cache on unbiased paths
    b = renderer(sc, W, H)
    b.set_cache(c, mode=2)
    b.render(s, W, H, spp)
    img = b.image()
    assert np.isfinite(img).all()
    assert (img[..., :3] >= 0).all()
    ra, rb = a.stats()["rays"], b.stats()["rays"]
    say("rays plain %d, cache %d" % (ra, rb))
    assert rb < ra
    # Same exposure: the cache estimate is of the same image (loose: an early, partly trained cache).
    assert abs(img[..., :3].mean() / a.image()[..., :3].mean() - 1) < 0.25
