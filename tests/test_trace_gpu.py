"""GPU parity of the closest-hit / any-hit traversal against the CPU oracle.

Bar: triangle and instance ids bit-exact (ties -> lowest (instance, triangle)), hit distance
bit-exact as well (the kernel and the oracle run the same binary32 operations; the stated
tolerance of 1e-6 relative is checked on top), occlusion flags identical.
"""
import sys

import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes

pytestmark = pytest.mark.gpu
REL_TOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_scene(geoms, instances):
    s = m.Scene()
    gids = [s.push_geometry(i, p) for i, p in geoms]
    for g, mat in instances:
        s.enqueue(gids[g], mat)
    s.prepare()
    return s, orc.OracleScene(geoms, instances)


def assert_closest_equal(gpu, ref, label):
    ref_t = ref["t"].astype(np.float32)
    found_g = gpu["inst"] != m.MISS
    found_r = ref["inst"] != m.MISS
    bad = np.flatnonzero(found_g != found_r)
    assert bad.size == 0, "%s: hit/miss differs on %d rays, first %s" % (label, bad.size, bad[:5])
    assert np.array_equal(gpu["inst"], ref["inst"]), label + ": instance ids differ"
    bad = np.flatnonzero(gpu["prim"] != ref["prim"])
    assert bad.size == 0, "%s: triangle ids differ on %d rays, first %s" % (label, bad.size, bad[:5])
    f = found_g
    rel = np.abs(gpu["t"][f].astype(np.float64) - ref_t[f]) / np.maximum(np.abs(ref_t[f]), 1e-30)
    assert rel.size == 0 or rel.max() <= REL_TOL, label + ": t beyond 1e-6 relative"
    assert np.array_equal(gpu["t"][f].view(np.uint32), ref_t[f].view(np.uint32)), label + ": t not bit-exact"
    assert np.all(np.isinf(gpu["t"][~f])), label + ": misses must report t = inf"


@pytest.fixture(scope="module")
def sphere_scene(ensure_built):
    (fi, fp, _), mats = scenes.sphere_field(200, 2, seed=1)
    gi, gp, _ = scenes.ground_quad()
    geoms = [(fi, fp), (gi, gp)]
    inst = [(0, mt) for mt in mats] + [(1, np.eye(4, dtype=np.float32))]
    return build_scene(geoms, inst)


def random_rays(n, seed, box=6.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-box, box, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return m.make_rays(o, d)


def test_closest_spheres_vs_brute_force(sphere_scene):
    s, o = sphere_scene
    rays = random_rays(20000, 11)
    say("spheres: 20000 random rays, 200 instances x 320 triangles + quad")
    assert_closest_equal(s.trace_closest(rays), o.brute(rays), "spheres")


def test_any_spheres_vs_brute_force(sphere_scene):
    s, o = sphere_scene
    rays = random_rays(20000, 12)
    rays["tmax"] = np.random.default_rng(3).uniform(0.5, 8.0, rays.size).astype(np.float32)
    occ = s.trace_any(rays)
    ref = np.isfinite(o.brute(rays, any_hit=True)["t"]).astype(np.uint8)
    assert np.array_equal(occ, ref)


def test_any_agrees_with_closest(sphere_scene):
    s, _ = sphere_scene
    rays = random_rays(20000, 13)
    hits = s.trace_closest(rays)
    occ = s.trace_any(rays)
    assert np.array_equal(occ.astype(bool), hits["inst"] != m.MISS)


@pytest.mark.parametrize("start", [(0.0, 3.0), (-3.0, 0.0)], ids=["ahead", "behind"])
def test_tmin_tmax_window(sphere_scene, start):
    """Hit windows [tmin, tmax] ahead of the origin and reaching behind it (a negative tmin: the
    far-plane widening works on the distance's magnitude, so boxes behind the origin stay in)."""
    s, o = sphere_scene
    rays = random_rays(5000, 14)
    rng = np.random.default_rng(15)
    rays["tmin"] = rng.uniform(start[0], start[1], rays.size).astype(np.float32)
    rays["tmax"] = rays["tmin"] + rng.uniform(0.0, 5.0, rays.size).astype(np.float32)
    ref = o.brute(rays)
    if start[1] <= 0.0:
        assert (ref["t"][np.isfinite(ref["t"])] < 0).sum() > 100  # hits behind the origin are in the test
    assert_closest_equal(s.trace_closest(rays), ref, "window")


def test_degenerate_and_empty(ensure_built):
    s = m.Scene()
    g = s.push_geometry(np.array([0, 1, 2], np.uint32), np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))
    s.prepare()  # no drawables: every ray misses
    rays = m.make_rays([[0.2, 0.2, -1]], [[0, 0, 1]])
    h = s.trace_closest(rays)
    assert h["inst"][0] == m.MISS and np.isinf(h["t"][0])
    s.enqueue(g)
    s.prepare()
    h = s.trace_closest(m.make_rays([[0.2, 0.2, -1], [0.2, 0.2, -1]], [[0, 0, 1], [0, 0, 0]]))
    assert h["inst"][0] == 0 and h["prim"][0] == 0 and abs(float(h["t"][0]) - 1.0) < 1e-6
    assert h["inst"][1] == m.MISS  # zero direction
    p = s.pick([0.2, 0.2, -1], [0, 0, 1])
    assert p is not None and p["prim"] == 0


def test_shared_edges_are_watertight(ensure_built):
    """Rays aimed exactly at grid vertices and cell diagonals of a flat grid never fall through."""
    idx, pts, _, _ = scenes.heightfield(40, 20, seed=3, height=0.0)
    s, o = build_scene([(idx, pts)], [(0, np.eye(4, dtype=np.float32))])
    xs = pts[:, 0].reshape(21, 41)
    zs = pts[:, 2].reshape(21, 41)
    inner = (slice(1, -1), slice(1, -1))
    targets = np.stack([xs[inner].ravel(), np.zeros(xs[inner].size), zs[inner].ravel()], 1)
    mids = np.stack([(xs[1:, 1:] + xs[:-1, :-1]).ravel() / 2, np.zeros(40 * 20), (zs[1:, 1:] + zs[:-1, :-1]).ravel() / 2], 1)
    tg = np.concatenate([targets, mids]).astype(np.float32)
    orig = tg + np.array([0.0, 3.0, 0.0], np.float32)
    rays = m.make_rays(orig, tg - orig)
    hits = s.trace_closest(rays)
    assert np.all(hits["inst"] == 0), "a ray through a shared edge or vertex fell through"
    assert_closest_equal(hits, o.brute(rays), "edges")


def test_heightfield_full_size(ensure_built):
    """The 1M-triangle heightfield with 8,388,608 coherent + incoherent rays against the oracle BVH,
    and a subset against brute force."""
    (idx, pts, nrm, uv), coh, inc = scenes.traversal_workload()
    s = m.Scene()
    g = s.push_geometry(idx, pts, nrm, uv)
    s.enqueue(g)
    s.prepare()
    o = orc.OracleScene([(idx, pts)], [(0, np.eye(4, dtype=np.float32))])
    for name, rays in (("coherent", coh), ("incoherent", inc)):
        say("heightfield: %s %d rays" % (name, rays.size))
        gpu = s.trace_closest(rays)
        ref = o.bvh(rays)
        assert_closest_equal(gpu, ref, name)
        sub = rays[:: rays.size // 1024][:1024]
        assert_closest_equal(s.trace_closest(sub), o.brute(sub), name + " brute subset")
        sh = scenes.shadow_rays(rays, gpu)
        occ = s.trace_any(sh)
        ref_occ = np.isfinite(o.bvh(sh, any_hit=True)["t"]).astype(np.uint8)
        assert np.array_equal(occ, ref_occ), name + ": shadow occlusion differs"


@pytest.mark.parametrize("grab", ["1", "4096"])
def test_ray_reserve_size_gives_identical_hits(ensure_built, monkeypatch, grab):
    """The number of rays a traversal wave reserves per atomic on the ray counter (MRT_TRACE_GRAB,
    128 by default) changes which wave takes which ray, never a result: hits and occlusion flags
    equal the default's bit for bit."""
    (idx, pts, nrm, uv), coh, inc = scenes.traversal_workload(n_coherent=1 << 16, n_incoherent=1 << 18)
    plain = m.Scene()
    plain.enqueue(plain.push_geometry(idx, pts, nrm, uv))
    plain.prepare()
    monkeypatch.setenv("MRT_TRACE_GRAB", grab)
    other = m.Scene()
    other.enqueue(other.push_geometry(idx, pts, nrm, uv))
    other.prepare()
    for name, rays in (("coherent", coh), ("incoherent", inc)):
        say("reserve %s: %s %d rays" % (grab, name, rays.size))
        a, b = plain.trace_closest(rays), other.trace_closest(rays)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
        sh = scenes.shadow_rays(rays, a)
        assert np.array_equal(plain.trace_any(sh), other.trace_any(sh)), name + " shadow"


@pytest.mark.parametrize("case", ["heightfield 1M", "wavefront 4M"])
def test_kernel_fetch_counts_match_the_cpu_replica(ensure_built, case):
    """The roofline's fetch counts per ray come from the kernel's counting variant; the CPU
    replica (oracle/fetch_count.cpp) walks the same device arrays the same way and reproduces
    them. The kernel's 1/d is the hardware reciprocal (1 ulp) and the replica's the correctly
    rounded one, so box tests on the edge of a slab may differ: totals within 0.1 %."""
    if case == "heightfield 1M":
        (idx, pts, nrm, uv), coh, inc = scenes.traversal_workload()
        s = m.Scene()
        s.enqueue(s.push_geometry(idx, pts, nrm, uv))
        sets = (("incoherent", inc), ("coherent", coh))
    else:
        sc = scenes.wavefront_scene()
        s = m.Scene()
        gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
        for g, mt, mi in sc["instances"]:
            s.enqueue(gids[g], mt, mi)
        tris, refs = scenes.world_triangles(sc["geometries"], sc["instances"])
        sets = (("incoherent", scenes.scene_bounce_rays(tris, refs, 1 << 20, seed=4)),)
        del tris, refs
    s.prepare()
    wide = s.wide_arrays()
    for name, rays in sets:
        d_r = m.DeviceBuffer.from_array(rays)
        for any_hit in (False, True):
            say("fetch counts, %s, %s %d rays, any=%d" % (case, name, rays.size, any_hit))
            gpu = s.trace_stats(d_r, rays.size, any_hit=any_hit)
            cpu, _ = orc.fetch_counts(wide, rays, any_hit=any_hit)
            tot = cpu.sum(0, dtype=np.int64)
            say("  kernel nodes %d tris %d inst %d | replica %d %d %d" % (
                gpu["node_visits"], gpu["tri_tests"], gpu["instance_visits"], tot[0], tot[1], tot[2]))
            assert gpu["rays"] == rays.size
            for key, k in (("node_visits", 0), ("tri_tests", 1), ("instance_visits", 2)):
                assert abs(gpu[key] - int(tot[k])) <= 1e-3 * max(int(tot[k]), 1), (name, key)
