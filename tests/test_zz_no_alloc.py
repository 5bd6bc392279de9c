"""mrt_trace_closest_device_noalloc against mrt_trace_closest_device through the C ABI.

The no-allocation entry point runs the same kernels with their scratch in the library's static
per-GPU array, so its hit records must equal the allocating entry point's bit for bit: on a
triangle-only scene (spill stack only) and on a triangle + SDF scene (SDF hit points and hit list
as well), on two calls in a row, the second with a smaller input. Sizes beyond its limit are
rejected before anything runs."""
import sys

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def both(scene, rays):
    n = rays.size
    d_rays = m.DeviceBuffer.from_array(rays)
    d_a = m.DeviceBuffer(n * 16)
    d_b = m.DeviceBuffer(n * 16)
    scene.trace_closest_device(d_rays, n, d_a)
    scene.trace_closest_device_noalloc(d_rays, n, d_b)
    m.synchronize()
    a = d_a.to_array(m.HIT_DTYPE, n)
    b = d_b.to_array(m.HIT_DTYPE, n)
    return a, b


def random_rays(n, seed, center=(0.0, 0.0, 0.0), spread=6.0):
    rng = np.random.default_rng(seed)
    o = np.asarray(center) + rng.normal(size=(n, 3)) * spread
    d = rng.normal(size=(n, 3))
    return m.make_rays(o, d / np.linalg.norm(d, axis=1, keepdims=True))


def test_noalloc_matches_the_allocating_trace_triangles(ensure_built):
    (idx, pts, nrm, uv), coh, inc = scenes.traversal_workload(n_coherent=1 << 18, n_incoherent=1 << 20)
    s = m.Scene()
    s.enqueue(s.push_geometry(idx, pts, nrm, uv))
    s.prepare()
    for label, rays in (("incoherent 2^20", inc), ("coherent 2^18", coh)):
        say("noalloc vs allocating, heightfield, %s" % label)
        a, b = both(s, rays)
        assert (a["inst"] != m.MISS).mean() > 0.05  # the first 2^18 coherent rays are mostly sky
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), label


def test_noalloc_matches_the_allocating_trace_with_sdf(ensure_built):
    sc = scenes.sphere_field_scene(40, 2, seed=2)
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    s.enqueue_sdf(s.push_sdf(scenes.sdf_cluster(32, seed=5)), scenes.translate_scale((0.0, 0.0, 0.0), 1.5))
    s.prepare()
    for label, n, seed in (("2^19 rays", 1 << 19, 1), ("2^15 rays", 1 << 15, 2)):
        rays = random_rays(n, seed)
        say("noalloc vs allocating, spheres + SDF, %s" % label)
        a, b = both(s, rays)
        sdf_hits = ((a["inst"] != m.MISS) & ((a["inst"] & m.SDF_HIT) != 0)).sum()
        assert sdf_hits > 100
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), label
    big = (1 << 20) + 1
    d_rays = m.DeviceBuffer(big * 32)
    d_hits = m.DeviceBuffer(big * 16)
    with pytest.raises(m.MrtError, match="at most 1048576 rays"):
        s.trace_closest_device_noalloc(d_rays, big, d_hits)
