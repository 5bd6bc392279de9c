"""SDF programs on the CPU: tape validation (product), closed-form distances and sphere tracing
against analytic intersections (oracle), and conservative drawable bounds (product)."""
import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes


def test_map_closed_forms(ensure_built):
    t = scenes.sdf_tape([("SPHERE", 3, 2.0)])
    d, mat = orc.sdf_map(t, [3.0, 4.0, 0.0])
    assert d == 3.0 and mat == 3
    t = scenes.sdf_tape([("BOX", 1, 1.0, 2.0, 3.0)])
    assert orc.sdf_map(t, [2.0, 0.0, 0.0])[0] == 1.0          # outside a face
    assert orc.sdf_map(t, [0.0, 0.0, 0.0])[0] == -1.0         # inside: -(distance to nearest face)
    assert orc.sdf_map(t, [4.0, 6.0, 3.0])[0] == 5.0          # corner region: |(3, 4, 0)|
    t = scenes.sdf_tape([("TORUS", 0, 2.0, 0.5)])
    assert orc.sdf_map(t, [2.0, 0.0, 0.0])[0] == -0.5
    t = scenes.sdf_tape([("CAPSULE", 0, 2.0, 0.5)])
    assert orc.sdf_map(t, [0.0, 3.0, 0.0])[0] == 0.5
    assert orc.sdf_map(t, [1.0, 1.0, 0.0])[0] == 0.5
    # smooth union of two spheres at the midpoint: d1 = d2 = 0.5 -> h = 0.5, d = 0.5 - k/4 (the
    # material there is either operand's, drawn per ray: tests/test_sdf_blend.py).
    t = scenes.sdf_tape([("PUSH_TRANSLATE", 0, -1, 0, 0), ("SPHERE", 1, 0.5), ("POP", 0),
                         ("PUSH_TRANSLATE", 0, 1, 0, 0), ("SPHERE", 2, 0.5), ("POP", 0), ("SMOOTH_UNION", 0, 0.2)])
    d, mat = orc.sdf_map(t, [0.0, 0.0, 0.0])
    assert abs(d - (0.5 - 0.05)) < 1e-7 and mat in (1, 2)
    # hard union picks the closer operand's material; subtraction max(-a, b).
    t = scenes.sdf_tape([("PUSH_TRANSLATE", 0, -1, 0, 0), ("SPHERE", 1, 0.5), ("POP", 0),
                         ("PUSH_TRANSLATE", 0, 1, 0, 0), ("SPHERE", 2, 0.5), ("POP", 0), ("UNION", 0)])
    assert orc.sdf_map(t, [0.9, 0.0, 0.0]) == (pytest.approx(-0.4, abs=1e-6), 2)
    t = scenes.sdf_tape([("SPHERE", 1, 0.5), ("SPHERE", 2, 1.0), ("SUBTRACT", 0)])
    assert orc.sdf_map(t, [0.0, 0.0, 0.0])[0] == 0.5
    t = scenes.sdf_tape([("PUSH_MIRROR", 0, 0), ("PUSH_TRANSLATE", 0, 2, 0, 0), ("SPHERE", 0, 1.0), ("POP", 0),
                         ("POP", 0)])
    assert orc.sdf_map(t, [-2.0, 0.0, 0.0])[0] == -1.0


def march_sphere_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-4, 4, (n, 3))
    o[:, 2] = -6.0
    tgt = rng.uniform(-0.9, 0.9, (n, 3)) * np.array([1, 1, 0])
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return m.make_rays(o, d)


def test_march_sphere_matches_analytic(ensure_built):
    tape = scenes.sdf_tape([("SPHERE", 0, 1.0)])
    dr = np.zeros(1, orc.SDF_DRAWABLE_DTYPE)
    dr["inv"] = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    dr["box_lo"], dr["box_hi"] = (-1.1, -1.1, -1.1), (1.1, 1.1, 1.1)
    dr["tape_len"] = 1
    rays = march_sphere_rays(2000, 1)
    hits, normals = orc.sdf_trace(tape, dr, rays)
    o = rays["o"].astype(np.float64)
    d = rays["d"].astype(np.float64)
    b = (o * d).sum(1)
    disc = b * b - ((o * o).sum(1) - 1.0)
    hit = disc > 0
    t_exact = -b - np.sqrt(np.maximum(disc, 0))
    assert np.array_equal(np.isfinite(hits["t"]) & (hits["inst"] != 0xFFFFFFFF), hit)
    err = np.abs(hits["t"][hit] - t_exact[hit])
    p = o[hit] + d[hit] * t_exact[hit][:, None]
    # Sphere tracing stops when |d| < eps: the distance error is at most ~eps / cos(incidence).
    cos_i = np.abs((d[hit] * p).sum(1))
    assert np.all(err <= 2e-5 / np.maximum(cos_i, 1e-2) + 1e-5)
    assert np.abs((normals[hit] * p).sum(1) - 1).max() < 1e-3   # normal = position on a unit sphere


def test_invalid_programs_rejected(ensure_built):
    s = m.Scene()
    bad = [
        [("SPHERE", 0, 1.0), ("UNION", 0)],
        [("PUSH_TRANSLATE", 0, 1, 0, 0), ("SPHERE", 0, 1.0)],
        [("POP", 0)],
        [("SPHERE", 0, 1.0), ("SPHERE", 0, 1.0)],
        [("SPHERE", 0, -1.0)],
        [("SPHERE", 0, 1.0), ("SPHERE", 0, 1.0), ("SMOOTH_UNION", 0, 0.0)],
    ]
    for prog in bad:
        with pytest.raises(m.MrtError):
            s.push_sdf(scenes.sdf_tape(prog))
    deep = [("SPHERE", 0, 1.0)] * 9 + [("UNION", 0)] * 8
    with pytest.raises(m.MrtError, match="deeper than 8"):
        s.push_sdf(scenes.sdf_tape(deep))
    assert s.push_sdf(scenes.sdf_cluster(8)) == 0


def test_drawable_box_contains_surface(ensure_built):
    """Points where the oracle's march stops lie inside the product's world box."""
    tape = scenes.sdf_cluster(64, seed=12)
    s = m.Scene()
    sid = s.push_sdf(tape)
    xf = scenes.translate_scale((0.5, 1.0, -2.0), 1.7)
    s.enqueue_sdf(sid, xf)
    s.build_host()
    dr = s.sdf_drawables(1)
    rng = np.random.default_rng(2)
    o = rng.uniform(-8, 8, (3000, 3))
    d = np.array([0.5, 1.0, -2.0]) + rng.uniform(-2, 2, (3000, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = m.make_rays(o, d)
    # A huge box for the oracle so that its march is not clipped; then check the product box.
    big = dr.copy()
    big["box_lo"], big["box_hi"] = (-1e3,) * 3, (1e3,) * 3
    hits, _ = orc.sdf_trace(tape, big, rays)
    f = hits["inst"] != 0xFFFFFFFF
    assert f.sum() > 200
    p = rays["o"][f] + rays["d"][f] * hits["t"][f][:, None]
    assert np.all(p >= dr["box_lo"][0]) and np.all(p <= dr["box_hi"][0])


@pytest.mark.parametrize("op", ["SMOOTH_UNION", "SMOOTH_SUBTRACT", "SMOOTH_INTERSECT", "UNION", "SUBTRACT",
                                "INTERSECT"])
def test_drawable_box_contains_negative_region(ensure_built, op):
    """Every grid point where the oracle's distance is negative lies inside the product's box, for
    each operator under a mirror and translations (a chain deep enough that loose bounds add up)."""
    spread = 0.1 if "INTERSECT" in op else 0.8  # keep the intersection non-empty
    nodes = [("PUSH_MIRROR", 0, 0), ("PUSH_TRANSLATE", 0, 6 * spread / 8, 0.1, 0.0), ("SPHERE", 1, 0.5), ("POP",),
             ("POP",)]
    rng = np.random.default_rng(3)
    for i in range(5):
        c = rng.uniform(-spread, spread, 3)
        if "INTERSECT" in op:
            prim = ("BOX", 2, 0.6, 0.5, 0.55) if i % 2 else ("SPHERE", 3, 0.7)
        else:
            prim = ("BOX", 2, 0.4, 0.3, 0.35) if i % 2 else ("TORUS", 3, 0.4, 0.12)
        nodes += [("PUSH_TRANSLATE", 0, *c), prim, ("POP",), (op, 0, 0.35)]
    tape = scenes.sdf_tape(nodes)
    s = m.Scene()
    s.enqueue_sdf(s.push_sdf(tape), np.eye(4, dtype=np.float32))
    s.build_host()
    dr = s.sdf_drawables(1)
    xs = np.linspace(-2.5, 2.5, 41)
    pts = np.stack(np.meshgrid(xs, xs, xs, indexing="ij"), -1).reshape(-1, 3)
    d = np.array([orc.sdf_map(tape, p)[0] for p in pts])
    inside = pts[d < 0]
    assert inside.shape[0] > 20
    assert np.all(inside >= dr["box_lo"][0]) and np.all(inside <= dr["box_hi"][0])
