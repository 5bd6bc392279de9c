"""Stochastic material choice of smooth SDF operators (src/pt/raymarch.glsl:43-56 `_mxm`, called
with t = 1 - h by src/pathtracer/sdf.hpp:462-522): at a hit point where a smooth operator's blend
weight is h, the right operand's material comes with probability 1 - h and the left one's with
probability h; hard operators stay deterministic. The oracle draws one uniform number per ray and
operator from a hash of the world ray's bits, so the choice is reproducible and independent of a
ray's position in a batch. The GPU march uses the same draw (tests/test_sdf_gpu.py compares its
materials with the oracle's bit for bit)."""
import numpy as np

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes

K = 0.6


def two_spheres(op):
    return scenes.sdf_tape([("PUSH_TRANSLATE", 0, -0.35, 0, 0), ("SPHERE", 1, 0.5), ("POP", 0),
                            ("PUSH_TRANSLATE", 0, 0.35, 0, 0), ("SPHERE", 2, 0.5), ("POP", 0), (op, 0, K)])


def drawable(tape):
    dr = np.zeros(1, orc.SDF_DRAWABLE_DTYPE)
    dr["inv"] = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    dr["box_lo"], dr["box_hi"] = (-1.2, -1.0, -1.0), (1.2, 1.0, 1.0)
    dr["tape_len"] = len(tape)
    return dr


def seam_rays(n, seed):
    """Rays from z = -5 towards the seam between the spheres (|x| < 0.3)."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n), np.full(n, -5.0)], 1)
    tgt = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n), np.zeros(n)], 1)
    d = tgt - o
    return m.make_rays(o, d / np.linalg.norm(d, axis=1, keepdims=True))


def blend_weight(p):
    """h of the smooth union at points p (binary64): left operand at x = -0.35, right at +0.35."""
    d1 = np.linalg.norm(p - [-0.35, 0, 0], axis=1) - 0.5
    d2 = np.linalg.norm(p - [0.35, 0, 0], axis=1) - 0.5
    return np.clip(0.5 + 0.5 * (d2 - d1) / K, 0.0, 1.0)


def test_smooth_union_material_frequency_is_one_minus_h(ensure_built):
    tape = two_spheres("SMOOTH_UNION")
    rays = seam_rays(40000, 1)
    hits, _ = orc.sdf_trace(tape, drawable(tape), rays)
    ok = hits["inst"] != 0xFFFFFFFF
    assert ok.mean() > 0.9
    p = rays["o"][ok].astype(np.float64) + rays["d"][ok].astype(np.float64) * hits["t"][ok, None]
    h = blend_weight(p)
    right = hits["prim"][ok] == 2
    assert set(np.unique(hits["prim"][ok])) == {1, 2}
    # Overall and per band of h: the count of right-operand materials is binomial with mean sum(1 - h).
    for lo, hi in ((0.0, 1.01), (0.0, 0.3), (0.3, 0.7), (0.7, 1.01)):
        sel = (h >= lo) & (h < hi)
        assert sel.sum() > 500
        expect, var = (1 - h[sel]).sum(), (h[sel] * (1 - h[sel])).sum()
        assert abs(right[sel].sum() - expect) <= 5 * np.sqrt(var) + 1, (lo, hi, right[sel].sum(), expect)
    # Where one operand dominates (h = 0 or 1 exactly) the choice is certain.
    assert np.all(right[h == 1.0] == False) and np.all(right[h == 0.0] == True)  # noqa: E712


def test_choice_is_per_ray_and_reproducible(ensure_built):
    tape = two_spheres("SMOOTH_UNION")
    rays = seam_rays(4000, 2)
    a, _ = orc.sdf_trace(tape, drawable(tape), rays)
    perm = np.random.default_rng(0).permutation(rays.size)
    b, _ = orc.sdf_trace(tape, drawable(tape), rays[perm])
    assert np.array_equal(a[perm].view(np.uint8), b.view(np.uint8))
    # The hit record's material is the seeded map at the hit point with the ray's seed.
    for i in range(0, 4000, 97):
        if a["inst"][i] == 0xFFFFFFFF:
            continue
        p = rays["o"][i] + rays["d"][i] * a["t"][i]
        seed = orc.sdf_blend_seed(rays["o"][i], rays["d"][i])
        assert orc.sdf_map_seeded(tape, p.astype(np.float32), seed)[1] == a["prim"][i]


def test_hard_union_stays_deterministic(ensure_built):
    tape = two_spheres("UNION")
    rays = seam_rays(4000, 3)
    hits, _ = orc.sdf_trace(tape, drawable(tape), rays)
    ok = hits["inst"] != 0xFFFFFFFF
    p = rays["o"][ok].astype(np.float64) + rays["d"][ok].astype(np.float64) * hits["t"][ok, None]
    nearer_right = np.linalg.norm(p - [0.35, 0, 0], axis=1) < np.linalg.norm(p - [-0.35, 0, 0], axis=1) - 1e-4
    nearer_left = np.linalg.norm(p - [-0.35, 0, 0], axis=1) < np.linalg.norm(p - [0.35, 0, 0], axis=1) - 1e-4
    assert np.all(hits["prim"][ok][nearer_right] == 2) and np.all(hits["prim"][ok][nearer_left] == 1)
