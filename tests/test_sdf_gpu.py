"""GPU SDF sphere tracing against the oracle's binary32 restatement: same hit/miss, bit-exact
distance and material, normals within the 16-bit octahedral encoding, for SDF-only and mixed
triangle + SDF scenes, closest and any hit."""
import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes

pytestmark = pytest.mark.gpu


def oct_decode(w):
    u = (w & 0xFFFF).astype(np.float64) * (2.0 / 65535.0) - 1.0
    v = (w >> 16).astype(np.float64) * (2.0 / 65535.0) - 1.0
    z = 1.0 - np.abs(u) - np.abs(v)
    uu = np.where(z < 0, (1.0 - np.abs(v)) * np.where(u >= 0, 1, -1), u)
    vv = np.where(z < 0, (1.0 - np.abs(u)) * np.where(v >= 0, 1, -1), v)
    n = np.stack([uu, vv, z], 1)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def rays_at(center, n, seed, spread=2.5, dist=8.0):
    rng = np.random.default_rng(seed)
    o = np.asarray(center) + rng.normal(size=(n, 3)) * dist
    tgt = np.asarray(center) + rng.uniform(-spread, spread, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return m.make_rays(o, d)


def check(gpu, ref, ref_n, label):
    assert np.array_equal(gpu["inst"], ref["inst"]), label + ": instance ids differ"
    assert np.array_equal(gpu["prim"], ref["prim"]), label + ": triangle ids / materials differ"
    f = gpu["inst"] != m.MISS
    assert np.array_equal(gpu["t"][f].view(np.uint32), ref["t"][f].astype(np.float32).view(np.uint32)), label
    s = (gpu["inst"] != m.MISS) & ((gpu["inst"] & m.SDF_HIT) != 0)
    if s.any():
        n = oct_decode(gpu["bary"][s])
        assert np.abs(n - ref_n[s]).max() < 2e-3, label + ": normals"


def test_sdf_cluster_closest_and_any(ensure_built):
    tape = scenes.sdf_cluster(64, seed=12)
    s = m.Scene()
    sid = s.push_sdf(tape)
    xf = scenes.translate_scale((0.0, 0.5, 0.0), 1.3)
    s.enqueue_sdf(sid, xf)
    s.prepare()
    dr = s.sdf_drawables(1)
    rays = rays_at((0, 0.5, 0), 20000, 1)
    gpu = s.trace_closest(rays)
    ref, ref_n = orc.sdf_trace(tape, dr, rays)
    assert (gpu["inst"] != m.MISS).mean() > 0.3
    check(gpu, ref, ref_n, "sdf")
    rays["tmax"] = np.random.default_rng(2).uniform(1, 10, rays.size).astype(np.float32)
    occ = s.trace_any(rays)
    ref_any, _ = orc.sdf_trace(tape, dr, rays, any_hit=True)
    assert np.array_equal(occ.astype(bool), np.isfinite(ref_any["t"]))


def test_sdf_with_triangles(ensure_built):
    idx, pts, nrm, _ = scenes.heightfield(80, 40, seed=3, extent=(8.0, 4.0), height=0.5)
    tape = scenes.sdf_cluster(16, seed=5, box=1.0)
    s = m.Scene()
    g = s.push_geometry(idx, pts, nrm)
    sid = s.push_sdf(tape)
    s.enqueue(g)
    s.enqueue_sdf(sid, scenes.translate_scale((0.0, 1.0, 0.0), 1.0))
    s.enqueue_sdf(sid, scenes.translate_scale((2.5, 0.8, 0.5), 0.6))
    s.prepare()
    dr = s.sdf_drawables(2)
    rays = rays_at((0.5, 0.5, 0), 20000, 3, spread=3.0)
    gpu = s.trace_closest(rays)
    o = orc.OracleScene([(idx, pts)], [(0, np.eye(4, dtype=np.float32))])
    ref, ref_n = orc.sdf_trace(tape, dr, rays, hits=o.brute(rays))
    assert ((gpu["inst"] != m.MISS) & ((gpu["inst"] & m.SDF_HIT) != 0)).sum() > 500 and (gpu["inst"] == 0).sum() > 500
    check(gpu, ref, ref_n, "mixed")


def mixed_chain(n, seed, op):
    """A union chain of translated spheres, boxes, tori and capsules (the culled program form)."""
    rng = np.random.default_rng(seed)
    kinds = [("SPHERE", 0.3), ("BOX", 0.25, 0.2, 0.3), ("TORUS", 0.3, 0.08), ("CAPSULE", 0.4, 0.12)]
    nodes = []
    for i in range(n):
        k = kinds[i % 4]
        nodes += [("PUSH_TRANSLATE", 0, *rng.uniform(-1.2, 1.2, 3)), (k[0], i % 5, *k[1:]), ("POP",)]
        if i:
            nodes.append((op, 0, 0.15))
    return scenes.sdf_tape(nodes)


@pytest.mark.parametrize("op", ["SMOOTH_UNION", "UNION"])
def test_sdf_culled_chain_is_exact(ensure_built, op):
    """Chains of mixed primitives evaluate through the culling grid: distances, materials and hit
    sets stay bit-identical to the full program of the oracle."""
    tape = mixed_chain(48, 31, op)
    s = m.Scene()
    s.enqueue_sdf(s.push_sdf(tape), scenes.translate_scale((0.2, -0.1, 0.3), 1.7))
    s.prepare()
    dr = s.sdf_drawables(1)
    rays = rays_at((0.2, -0.1, 0.3), 30000, 7, spread=2.0)
    gpu = s.trace_closest(rays)
    ref, ref_n = orc.sdf_trace(tape, dr, rays)
    assert (gpu["inst"] != m.MISS).mean() > 0.2
    check(gpu, ref, ref_n, "chain " + op)


def general_program():
    """A program the culling grid does not cover: mirror and translate modifiers, smooth and plain
    subtraction, a torus, a Menger sponge, unions. It runs through the general tape interpreter."""
    return scenes.sdf_tape([
        ("CAPSULE", 2, 0.6, 0.15),
        ("PUSH_MIRROR", 0, 0), ("PUSH_TRANSLATE", 0, 0.6, 0.0, 0.0), ("SPHERE", 1, 0.45), ("POP", 0), ("POP", 0),
        ("BOX", 2, 0.5, 0.3, 0.4), ("SMOOTH_SUBTRACT", 0, 0.05), ("TORUS", 3, 0.7, 0.1), ("UNION", 0),
        ("PUSH_TRANSLATE", 0, 0.0, 0.8, 0.0), ("MENGER", 1), ("POP", 0), ("SMOOTH_UNION", 0, 0.1),
        ("SUBTRACT", 0)])


@pytest.mark.parametrize("with_chain", [False, True])
def test_sdf_general_programs(ensure_built, with_chain):
    """Programs without a culling grid, alone (the general kernel) and next to a culled chain (both
    drawables then march in the general kernel): hits, materials and t bit-exact, normals close."""
    tape = general_program()
    s = m.Scene()
    s.enqueue_sdf(s.push_sdf(tape), scenes.translate_scale((0.0, 0.2, 0.0), 1.2))
    tapes = [tape]
    if with_chain:
        chain = scenes.sdf_cluster(16, seed=9, box=0.6)
        s.enqueue_sdf(s.push_sdf(chain), scenes.translate_scale((1.8, 0.0, 0.3), 0.8))
        tapes.append(chain)
    s.prepare()
    dr = s.sdf_drawables(len(tapes))
    rays = rays_at((0.5, 0.2, 0.0), 20000, 21, spread=3.0)
    gpu = s.trace_closest(rays)
    # The drawable records address the scene's tapes back to back, in push order.
    joint = np.concatenate(tapes)
    ref, ref_n = orc.sdf_trace(joint, dr, rays)
    assert (gpu["inst"] != m.MISS).mean() > 0.15
    check(gpu, ref, ref_n, "general" + (" + chain" if with_chain else ""))
    rays["tmax"] = np.random.default_rng(22).uniform(1, 10, rays.size).astype(np.float32)
    occ = s.trace_any(rays)
    ref_any, _ = orc.sdf_trace(joint, dr, rays, any_hit=True)
    assert np.array_equal(occ.astype(bool), np.isfinite(ref_any["t"]))
