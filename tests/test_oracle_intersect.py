"""The oracle's ray queries, pinned without the reference's code (it has no tests and cannot be
built here):

* the watertight test (Woop, Benthin and Wald, JCGT 2013) against hand-computed hits, exact
  shared-edge/vertex rays and a double-precision exact test away from edges;
* Moeller-Trumbore as the reference writes it (src/pt/intersect.glsl:3-28): hand cases for its
  epsilon and border rules;
* the oracle BVH query against brute force; the reference's stackless traversal against brute
  force on the closest distance (its ties go to the first triangle found, not the lowest id).
"""
import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes

TRI = [np.float32([0, 0, 0]), np.float32([1, 0, 0]), np.float32([0, 1, 0])]


def test_watertight_hand_case():
    # o = (0.25, 0.25, -1), d = +z: U = -0.5, V = W = -0.25, det = -1, T = -1 -> t = 1,
    # weights of v1 and v2 are 0.25 each (all exact in binary32).
    hit, t, b1, b2 = orc.watertight([0.25, 0.25, -1], [0, 0, 1], *TRI)
    assert hit and t == 1.0 and b1 == 0.25 and b2 == 0.25
    # Both faces are hit (no culling), from the other side too.
    hit, t, _, _ = orc.watertight([0.25, 0.25, 2], [0, 0, -1], *TRI)
    assert hit and t == 2.0
    # Scaling the direction scales t inversely (t is in units of |d|).
    hit, t, _, _ = orc.watertight([0.25, 0.25, -1], [0, 0, 4], *TRI)
    assert hit and t == 0.25


def test_watertight_misses():
    assert not orc.watertight([2, 2, -1], [0, 0, 1], *TRI)[0]          # outside
    assert not orc.watertight([0.25, 0.25, -1], [1, 0, 0], *TRI)[0]    # parallel to the plane
    hit, t, _, _ = orc.watertight([0.25, 0.25, 1], [0, 0, 1], *TRI)     # plane behind: t < 0
    assert hit and t == -1.0


def test_watertight_edges_and_vertices_never_leak():
    """Rays passing exactly through the shared diagonal of a split quad, and exactly through the
    shared centre vertex of a six-triangle fan, hit at least one triangle, for many directions.
    Directions use multiples of 1/8 so that o = p - 2 d is exact and the ray contains p."""
    a, b, c, d = (np.float32(v) for v in ([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]))
    rng = np.random.default_rng(1)

    def dirs(k):
        out = rng.integers(-16, 17, size=(k, 3)).astype(np.float32) / np.float32(8)
        out[:, 2] = -np.abs(out[:, 2]) - np.float32(0.125)
        return out

    for k in range(1, 16):
        p = np.float32([k / 16, k / 16, 0])
        for dv in dirs(40):
            o = p - dv * np.float32(2.0)
            assert orc.watertight(o, dv, a, b, c)[0] or orc.watertight(o, dv, a, c, d)[0], \
                "ray through %s fell through the shared edge" % p
    centre = np.float32([0, 0, 0])
    ring = [np.float32([np.cos(t), np.sin(t), 0]) for t in np.linspace(0, 2 * np.pi, 7)[:-1]]
    fan = [(centre, ring[i], ring[(i + 1) % 6]) for i in range(6)]
    for dv in dirs(200):
        o = centre - dv * np.float32(2.0)
        assert any(orc.watertight(o, dv, *tri)[0] for tri in fan), "ray through the fan centre fell through"


def exact_hit(o, d, v0, v1, v2):
    """Exact-ish reference in double precision: barycentrics by Cramer's rule."""
    o, d, v0, v1, v2 = (np.asarray(x, np.float64) for x in (o, d, v0, v1, v2))
    e1, e2 = v1 - v0, v2 - v0
    mat = np.stack([-d, e1, e2], axis=1)
    det = np.linalg.det(mat)
    if abs(det) < 1e-12:
        return None
    t, u, v = np.linalg.solve(mat, o - v0)
    return t, u, v


def test_watertight_matches_exact_away_from_edges():
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(3000):
        v = rng.uniform(-2, 2, (3, 3)).astype(np.float32)
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        ex = exact_hit(o, d, *v)
        if ex is None:
            continue
        t, u, w = ex
        margin = min(u, w, 1 - u - w)
        hit, tt, b1, b2 = orc.watertight(o, d, *v)
        if abs(margin) < 1e-4:
            continue
        checked += 1
        assert hit == (margin > 0)
        if hit:
            assert abs(tt - t) <= 1e-5 * max(1.0, abs(t))
            assert abs(b1 - u) < 1e-4 and abs(b2 - w) < 1e-4
    assert checked > 2000


def test_moller_trumbore_reference_rules():
    v1, v2, v3 = TRI
    hit, t, bx, by = orc.moller_trumbore([0.25, 0.25, -1], [0, 0, 1], v1, v2, v3)
    assert hit and t == 1.0 and bx == 0.25 and by == 0.25
    # Border epsilon 1e-6: a point 5e-7 outside the edge bx >= 0 is still a hit, 1e-5 is not.
    assert orc.moller_trumbore([-5e-7, 0.25, -1], [0, 0, 1], v1, v2, v3)[0]
    assert not orc.moller_trumbore([-1e-5, 0.25, -1], [0, 0, 1], v1, v2, v3)[0]
    # |det| < 1e-23: a ray in the triangle's plane misses.
    assert not orc.moller_trumbore([0.25, 0.25, 0], [1, 0, 0], v1, v2, v3)[0]
    # t must exceed 1e-23: a triangle behind the origin is not hit.
    assert not orc.moller_trumbore([0.25, 0.25, 1], [0, 0, 1], v1, v2, v3)[0]


def sphere_scene(n=30, level=1, seed=4):
    (fi, fp, _), mats = scenes.sphere_field(n, level, seed=seed)
    gi, gp, _ = scenes.ground_quad()
    geoms = [(fi, fp), (gi, gp)]
    return geoms, [(0, mt) for mt in mats] + [(1, np.eye(4, dtype=np.float32))]


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-6, 6, (n, 3))
    d = rng.normal(size=(n, 3))
    return m.make_rays(o, d / np.linalg.norm(d, axis=1, keepdims=True))


def test_oracle_bvh_equals_brute_force(ensure_built):
    geoms, inst = sphere_scene()
    s = orc.OracleScene(geoms, inst)
    rays = random_rays(3000, 5)
    a, b = s.brute(rays), s.bvh(rays)
    assert np.array_equal(a["inst"], b["inst"]) and np.array_equal(a["prim"], b["prim"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    rays["tmax"] = np.float32(3.0)
    assert np.array_equal(np.isfinite(s.brute(rays, True)["t"]), np.isfinite(s.bvh(rays, True)["t"]))


def test_brute_force_ties_go_to_lowest_ids(ensure_built):
    """Two coincident copies of one triangle (as two instances and as two triangles of one
    geometry): the closest hit reports the lowest (instance, triangle) pair."""
    idx = np.array([0, 1, 2, 0, 1, 2], np.uint32)
    pts = np.array(TRI, np.float32)
    s = orc.OracleScene([(idx, pts)], [(0, np.eye(4, dtype=np.float32)), (0, np.eye(4, dtype=np.float32))])
    h = s.brute(m.make_rays([[0.25, 0.25, -1]], [[0, 0, 1]]))
    assert h["inst"][0] == 0 and h["prim"][0] == 0 and h["t"][0] == 1.0


def test_reference_traversal_finds_closest_distance(ensure_built):
    """The restated stackless traversal (with Moeller-Trumbore) against brute-force
    Moeller-Trumbore on a shallow scene: the closest distances agree."""
    geoms, inst = sphere_scene(12, 1, seed=6)
    s = orc.OracleScene(geoms, inst)
    rays = random_rays(400, 7)
    ref = s.reference(rays)
    for i in range(rays.size):
        o, d = rays["o"][i], rays["d"][i]
        best = np.inf
        for k, (g, mt) in enumerate(inst):
            inv = orc.invert(mt).reshape(3, 4)
            oo = (inv[:, :3] @ o.astype(np.float32) + inv[:, 3]).astype(np.float32)
            od = (inv[:, :3] @ d.astype(np.float32)).astype(np.float32)
            gi, gp = geoms[g]
            gi = np.asarray(gi).reshape(-1, 3)
            for tri in gi:
                hit, t, _, _ = orc.moller_trumbore(oo, od, gp[tri[0]], gp[tri[1]], gp[tri[2]])
                if hit and t < best:
                    best = t
        if np.isinf(best):
            assert np.isinf(ref["t"][i])
        else:
            assert abs(ref["t"][i] - best) <= 1e-5 * max(1.0, best)


def test_inverse_matches_product(ensure_built):
    rng = np.random.default_rng(8)
    mats = []
    for _ in range(20):
        a = rng.normal(size=(3, 3))
        q, _ = np.linalg.qr(a)
        mt = np.eye(4, dtype=np.float32)
        mt[:3, :3] = (q * rng.uniform(0.2, 3.0, 3)).astype(np.float32)
        mt[:3, 3] = rng.uniform(-5, 5, 3)
        mats.append(mt)
    s = m.Scene()
    g = s.push_geometry(np.array([0, 1, 2], np.uint32), np.array(TRI, np.float32))
    for mt in mats:
        s.enqueue(g, mt)
    s.build_host()
    prod = s.instance_inverses(len(mats))
    for k, mt in enumerate(mats):
        ref = orc.invert(mt)
        assert prod[k].tobytes() == ref.tobytes()
        full = np.vstack([ref.reshape(3, 4), [0, 0, 0, 1]]).astype(np.float64)
        assert np.allclose(full @ mt.astype(np.float64), np.eye(4), atol=1e-5)
