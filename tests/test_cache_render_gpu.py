"""Radiance cache inside the wavefront renderer.

Training paths are ordinary unbiased paths that also record their vertices, so a render that only
trains (mode 1) must equal the plain render; a render that ends paths in the cache (mode 2) traces
fewer rays and stays finite. Self-training (training paths end in the cache at train_end and their
earlier vertices learn the cache's prediction as their tail) is checked in a closed room, where
most radiance arrives after several bounces: a cache whose targets lost the tail would render the
room about half as bright as the plain tracer."""
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


def closed_room(albedo=0.5):
    """The inside of the cube [-2, 2]^3 (12 triangles, normals inward), Lambert walls of the given
    albedo, a 1 x 1 quad light under the ceiling emitting downward, no environment."""
    idx, pts, nrm = [], [], []
    for a in range(3):
        u, v = (a + 1) % 3, (a + 2) % 3
        for sign in (-1.0, 1.0):
            base = len(pts)
            for cu, cv in ((-2, -2), (2, -2), (2, 2), (-2, 2)):
                p = [0.0, 0.0, 0.0]
                p[a], p[u], p[v] = 2.0 * sign, float(cu), float(cv)
                n = [0.0, 0.0, 0.0]
                n[a] = -sign
                pts.append(p)
                nrm.append(n)
            # (0, 1, 2) winds around +e_a; the wall at +2 faces -e_a.
            tri = [(0, 2, 1), (0, 3, 2)] if sign > 0 else [(0, 1, 2), (0, 2, 3)]
            idx += [[base + i for i in t] for t in tri]
    s = m.Scene()
    g = s.push_geometry(np.array(idx, np.uint32), np.array(pts, np.float32), np.array(nrm, np.float32))
    mat = np.zeros(1, m.MATERIAL_DTYPE)
    mat[0] = (scenes._rgba8((albedo, albedo, albedo)), 1.0, 1.0, 0.0, 0.0, 0, (0, 0))
    s.push_material(mat[0])
    lights = np.zeros(1, m.LIGHT_DTYPE)
    lights[0] = ((-0.5, 1.95, -0.5), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (10.0, 10.0, 10.0))
    s.set_lights(lights)
    s.set_environment(0.0)
    s.enqueue(g, np.eye(4, dtype=np.float32), 0)
    return dict(eye=(0.0, 0.0, -1.9), target=(0.0, 0.0, 0.0), fovy=70.0), s


def test_self_training_learns_the_tail(ensure_built):
    sc, s = closed_room()
    W, H, depth = 64, 64, 16
    ref = renderer(sc, W, H, depth)
    ref.render(s, W, H, 1024)
    mean_ref = float(ref.image()[..., :3].mean())
    c = m.RadianceCache(seed=3)
    c.set_bounds((-2.2, -2.2, -2.2), (2.2, 2.2, 2.2))
    r = renderer(sc, W, H, depth)
    r.set_cache(c, mode=2, min_bounce=1)
    r.set_cache_training(2, 0.0)  # training paths end at vertex 2: vertex 1 learns direct + cache
    for _ in range(40):
        r.render(s, W, H, 16)
    r.invalidate()
    r.render(s, W, H, 256)
    img = r.image()
    mean_c = float(img[..., :3].mean())
    say("closed room: plain depth %d mean %.4f, self-trained cache %.4f; rays %d vs %d" % (
        depth, mean_ref, mean_c, ref.stats()["rays"], r.stats()["rays"]))
    assert np.isfinite(img).all()
    assert abs(mean_c / mean_ref - 1.0) < 0.05


def test_unbiased_training_paths_run_to_full_depth(ensure_built):
    sc, s = small_scene()
    W, H, spp = 64, 48, 8
    c = cache_for_scene()
    rays = {}
    for frac in (0.0, 1.0):
        r = renderer(sc, W, H)
        r.set_cache(c, mode=2, train_fraction=0.25)
        r.set_cache_training(2, frac)
        r.render(s, W, H, spp)
        assert np.isfinite(r.image()).all()
        rays[frac] = r.stats()["rays"]
    r = renderer(sc, W, H)
    r.set_cache(c, mode=2, train_fraction=0.25)
    r.set_cache_training(0, 0.0)
    r.render(s, W, H, spp)
    say("rays: train_end 2 %d, train_end 2 all unbiased %d, train_end 0 %d" % (rays[0.0], rays[1.0], r.stats()["rays"]))
    assert rays[0.0] < rays[1.0] <= r.stats()["rays"]
    with pytest.raises(m.MrtError):
        r.set_cache_training(-1, 0.0)
    with pytest.raises(m.MrtError):
        r.set_cache_training(2, 1.5)
