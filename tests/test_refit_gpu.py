"""GPU refit of the instance BVH for frames that only move instances.

The reference rebuilds its instance BVH whenever the drawable hash changes, a transform included
(src/pathtracer/scene.cpp:255-271, 336-344). With refits enabled the product keeps the TLAS
topology and refits it on the GPU. Bars:
* a refit with unchanged matrices reproduces the host-built TLAS nodes and instance records bit
  for bit (the device repeats the host's inverse, world-box and quantisation arithmetic);
* after moving every instance, the refitted scene's closest hits equal, bit for bit, those of a
  scene rebuilt from scratch with the same matrices, and its inverses equal the host's;
* every decoded child box of the refitted TLAS contains the world boxes of its instances;
* device-resident matrices (mrt_scene_update_transforms_device) give the same result as the
  enqueue path; a singular matrix is reported and leaves its instance unhittable.
"""
import sys
import time

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes

pytestmark = pytest.mark.gpu
N_SPHERES = 1500


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rot_y(a):
    c, s = np.cos(a), np.sin(a)
    r = np.eye(4, dtype=np.float32)
    r[0, 0], r[0, 2], r[2, 0], r[2, 2] = c, s, -s, c
    return r


def moved(instances, seed):
    rng = np.random.default_rng(seed)
    out = []
    for g, mt, mi in instances:
        t = np.eye(4, dtype=np.float32)
        t[:3, 3] = rng.normal(size=3) * 0.7
        out.append((g, (t @ rot_y(rng.uniform(0, 2 * np.pi)) @ np.asarray(mt, np.float32)).astype(np.float32), mi))
    return out


def make_scene(sc, instances, refit=0):
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    s.set_tlas_refit(refit)
    for g, mt, mi in instances:
        s.enqueue(gids[g], mt, mi)
    return s, gids


def probe_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3)) * 6.0
    d = rng.normal(size=(n, 3))
    return m.make_rays(o, d / np.linalg.norm(d, axis=1, keepdims=True))


def colmajor(instances):
    return np.stack([np.ascontiguousarray(np.asarray(mt, np.float32).T).ravel() for _, mt, _ in instances])


@pytest.fixture(scope="module", params=["1500 instances", "30000 instances"])
def field(ensure_built, request):
    """1,500 icospheres (a TLAS small enough for the one-workgroup refit) or 30,000 (the
    per-level launches)."""
    if request.param.startswith("1500"):
        return scenes.sphere_field_scene(N_SPHERES, 2, seed=4)
    return scenes.sphere_field_scene(30000, 0, seed=5)


def test_refit_with_unchanged_matrices_reproduces_the_build(field):
    s, _ = make_scene(field, field["instances"], refit=4)
    p = s.prepare()
    _, _, tlas0, items0, inst0 = s.wide_arrays()
    say("refit of %d instances / %d TLAS nodes with unchanged matrices" % (p["num_instances"], p["tlas_wide_nodes"]))
    assert (p["tlas_wide_nodes"] <= 4096) == (p["num_instances"] < 10000)  # both refit forms are covered
    d = m.DeviceBuffer.from_array(colmajor(field["instances"]))
    s.update_transforms_device(d, len(field["instances"]))
    assert s.update_status() == 0
    _, _, tlas1, items1, inst1 = s.wide_arrays()
    assert np.array_equal(items0, items1)
    assert np.array_equal(tlas0.view(np.uint8), tlas1.view(np.uint8))
    assert np.array_equal(inst0.view(np.uint8), inst1.view(np.uint8))
    s.close()


def decoded_boxes(node, c):
    sc = np.ldexp(np.float32(1.0), node["exp"].astype(np.int32)).astype(np.float32)
    lo = np.float32(node["qlo"][:, c]).astype(np.float32) * sc + node["origin"]
    hi = np.float32(node["qhi"][:, c]).astype(np.float32) * sc + node["origin"]
    return lo, hi


def test_refit_after_motion_matches_a_rebuild(field):
    s, gids = make_scene(field, field["instances"], refit=4)
    s.prepare()
    frames = [moved(field["instances"], seed) for seed in (11, 12)]
    rays = probe_rays(1 << 18, 5)
    for k, inst in enumerate(frames):
        for g, mt, mi in inst:
            s.enqueue(gids[g], mt, mi)
        t0 = time.perf_counter()
        p = s.prepare()
        m.synchronize()
        t_refit = time.perf_counter() - t0
        assert p["tlas_refitted"] == 1 and p["drawables_changed"] == 1
        ref, _ = make_scene(field, inst)
        t0 = time.perf_counter()
        ref.prepare()
        m.synchronize()
        t_build = time.perf_counter() - t0
        say("frame %d: prepare with refit %.2f ms, fresh scene build %.2f ms" % (k, t_refit * 1e3, t_build * 1e3))
        a, b = s.trace_closest(rays), ref.trace_closest(rays)
        assert (a["inst"] != m.MISS).mean() > 0.2
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), "frame %d" % k
        n = len(inst)
        assert np.array_equal(s.instance_inverses(n), ref.instance_inverses(n))
        ref.close()
    # Containment: every decoded TLAS child box holds its instances' world boxes.
    _, _, tlas, items, _ = s.wide_arrays()
    geo_lo = [np.min(np.asarray(p_, np.float64).reshape(-1, 3), 0) for _, p_, _ in field["geometries"]]
    geo_hi = [np.max(np.asarray(p_, np.float64).reshape(-1, 3), 0) for _, p_, _ in field["geometries"]]
    world = []
    for g, mt, _ in frames[-1]:
        corners = np.array([[geo_hi[g][0] if i & 1 else geo_lo[g][0], geo_hi[g][1] if i & 2 else geo_lo[g][1],
                             geo_hi[g][2] if i & 4 else geo_lo[g][2], 1.0] for i in range(8)])
        w = corners @ np.asarray(mt, np.float64).T
        world.append((w[:, :3].min(0), w[:, :3].max(0)))
    checked = 0
    for node in tlas:
        for c in range(node["nchildren"]):
            e = int(node["entry"][c])
            if not e & 0x80000000:
                continue
            lo, hi = decoded_boxes(node, c)
            first, count = e & 0xFFFFFF, (e >> 24) & 0x7F
            for it in items[first:first + count]:
                wl, wh = world[it]
                assert np.all(lo <= wl + 1e-6 * (1 + np.abs(wl))) and np.all(hi >= wh - 1e-6 * (1 + np.abs(wh)))
                checked += 1
    assert checked == len(frames[-1])
    # The refit budget: after 4 refits in a row the next change rebuilds.
    for k in range(3):
        for g, mt, mi in moved(field["instances"], 20 + k):
            s.enqueue(gids[g], mt, mi)
        assert s.prepare()["tlas_refitted"] == (1 if k < 2 else 0)
    s.close()


def test_device_matrices_and_singular_update(field):
    inst = moved(field["instances"], 31)
    s, gids = make_scene(field, field["instances"], refit=0)
    s.prepare()
    d = m.DeviceBuffer.from_array(colmajor(inst))
    s.update_transforms_device(d, len(inst))
    assert s.update_status() == 0
    ref, _ = make_scene(field, inst)
    ref.prepare()
    rays = probe_rays(1 << 17, 6)
    a, b = s.trace_closest(rays), ref.trace_closest(rays)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    n = len(inst)
    assert np.array_equal(s.instance_inverses(n), ref.instance_inverses(n))
    # Enqueueing the first frame's matrices again is a change (the device moved on).
    for g, mt, mi in field["instances"]:
        s.enqueue(gids[g], mt, mi)
    assert s.prepare()["drawables_changed"] == 1
    # A singular matrix: reported, and its instance is no longer hit.
    mats = colmajor(inst)
    n_spheres = len(field["instances"]) - 1
    hits_per = np.bincount(b["inst"][b["inst"] < n_spheres].astype(np.int64), minlength=n_spheres)
    victim = int(np.argmax(hits_per))
    assert hits_per[victim] > 0
    mats[victim] = 0.0
    s.update_transforms_device(m.DeviceBuffer.from_array(mats), n)
    assert s.update_status() & 1
    h = s.trace_closest(rays)
    assert not np.any(h["inst"] == victim)
    others = b["inst"] != victim  # rays whose closest hit was elsewhere keep it exactly
    assert np.array_equal(h[others].view(np.uint8), b[others].view(np.uint8))
    with pytest.raises(m.MrtError, match="one matrix per prepared instance"):
        s.update_transforms_device(d, n - 1)
    ref.close()
    s.close()
