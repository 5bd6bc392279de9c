"""The 64-byte wide BVH the kernels read (host collapse and quantisation).

Checks that (a) every triangle is reachable: on every root-to-leaf path to a leaf slot holding a
triangle referenced once (not box-split), each decoded child box contains the triangle's vertices
strictly; (b) leaf slots hold exactly the reordered reference array of the binary BVH; (c) the
records match the mesh."""
import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes


def decode(node, c):
    # code * 2^e is exact in binary32, so float32 multiply-then-add rounds once, like fmaf.
    scale = np.ldexp(np.float32(1.0), node["exp"].astype(np.int32)).astype(np.float32)
    lo = np.array([np.float32(node["qlo"][a][c]) * scale[a] + node["origin"][a] for a in range(3)], np.float64)
    hi = np.array([np.float32(node["qhi"][a][c]) * scale[a] + node["origin"][a] for a in range(3)], np.float64)
    return lo, hi


def walk(nodes):
    """Yields (leaf item range, list of decoded boxes on the path) for every leaf child."""
    stack = [(0, [])]
    while stack:
        k, path = stack.pop()
        n = nodes[k]
        for c in range(n["nchildren"]):
            box = decode(n, c)
            e = int(n["entry"][c])
            if not e & 0x80000000:
                stack.append((e, path + [box]))
            else:
                yield e & 0xFFFFFF, (e >> 24) & 0x7F, path + [box]


@pytest.mark.parametrize("builder", ["sah", "reference"])
def test_heightfield_wide_bvh_contains_every_triangle(ensure_built, builder, monkeypatch):
    # The scene's traversal trees come from the SAH builder by default; MRT_BVH_BUILDER=0 (read
    # when a scene is created) selects the reference builder's trees.
    monkeypatch.setenv("MRT_BVH_BUILDER", "1" if builder == "sah" else "0")
    idx, pts, _, _ = scenes.heightfield(120, 60, seed=3)
    s = m.Scene()
    g = s.push_geometry(idx, pts)
    s.enqueue(g)
    r = s.build_host()
    blas, tris, tlas, items, inst = s.wide_arrays()
    assert blas.size == r["blas_wide_nodes"] and tris.size >= idx.shape[0]
    refs = np.bincount(tris["prim"], minlength=idx.shape[0])
    assert (refs == 1).mean() > 0.9
    seen = np.zeros(idx.shape[0], bool)
    for start, count, path in walk(blas):
        for k in range(start, start + count):
            t = tris[k]
            v = np.stack([t["v0"], t["v1"], t["v2"]]).astype(np.float64)
            for lo, hi in path if refs[t["prim"]] == 1 else []:
                assert np.all(v > lo) and np.all(v < hi), "triangle %d escapes a decoded box" % t["prim"]
            seen[t["prim"]] = True
            assert np.array_equal(t["v0"], pts[idx[t["prim"], 0]])
            assert np.array_equal(t["v2"], pts[idx[t["prim"], 2]])
    assert seen.all()
    if builder == "sah":
        # No reference duplication: every triangle sits in exactly one leaf.
        assert np.array_equal(np.sort(tris["prim"]), np.arange(idx.shape[0]))
    else:
        # The leaf slots are the reference builder's reordered references, grouped per leaf.
        _, ref_idx, _ = orc.bvh_build(orc.triangle_bounds(idx, pts), True)
        assert sorted(tris["prim"].tolist()) == sorted(ref_idx.tolist())
    # TLAS: one instance, identity.
    assert items.tolist() == [0] and inst["blas_root"][0] == 0


def test_sphere_field_records(ensure_built):
    (fi, fp, _), mats = scenes.sphere_field(40, 2, seed=3)
    s = m.Scene()
    g = s.push_geometry(fi, fp)
    for mt in mats:
        s.enqueue(g, mt)
    r = s.build_host()
    blas, tris, tlas, items, inst = s.wide_arrays()
    assert sorted(items.tolist()) == list(range(40))
    # Every instance appears in exactly one TLAS leaf slot and each decoded instance box on the
    # path contains the instance's transformed vertices.
    world = {}
    for k, mt in enumerate(mats):
        world[k] = (np.c_[fp, np.ones(len(fp))] @ mt.T.astype(np.float64))[:, :3]
    for start, count, path in walk(tlas):
        for slot in range(start, start + count):
            w = world[int(items[slot])]
            for lo, hi in path:
                assert np.all(w >= lo) and np.all(w <= hi)
    assert set(np.unique(tris["prim"]).tolist()) == set(range(fi.shape[0]))
    assert r["max_stack_entries"] >= 8
