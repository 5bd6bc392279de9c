"""The BVH builder (src/pathtracer/bvh.cpp:43-161, 356-512).

* The oracle's literal restatement is pinned by hand-derived cases (closed forms from the SAH
  constants) and by structural invariants that a wrong index, a lost or duplicated reference, a
  wrong box or a broken queue order would violate.
* The product builder (node-local reference lists, parallel subtrees, breadth-first renumbering)
  must reproduce the oracle bit for bit: nodes, boxes, parent links and reordered indices.
"""
import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes


def product_build(boxes, split):
    return m.bvh_build(boxes.view(m.AABB_DTYPE), split)


def assert_same_build(boxes, split):
    n1, i1 = product_build(boxes, split)
    n2, i2, _ = orc.bvh_build(boxes, split)
    assert n1.size == n2.size, "node count %d vs %d" % (n1.size, n2.size)
    assert n1.tobytes() == n2.view(m.NODE_DTYPE).tobytes(), "node arrays differ"
    assert np.array_equal(i1, i2), "reordered indices differ"
    return n2, i2


def check_invariants(nodes, indices, final_boxes, input_boxes):
    n = nodes.size
    leaf = (nodes["type_and_parent"] & 0x80000000) != 0
    parent = nodes["type_and_parent"] & 0x7FFFFFFF
    # Root box = union of the input boxes.
    assert np.array_equal(nodes["min"][0], input_boxes["min"].min(axis=0))
    assert np.array_equal(nodes["max"][0], input_boxes["max"].max(axis=0))
    # Queue (breadth-first) order: inner node k's children are the next unused pair, in order.
    nxt = 1
    for k in range(n):
        if not leaf[k]:
            assert nodes["first"][k] == nxt and nodes["second"][k] == nxt + 1
            assert parent[nxt] == k and parent[nxt + 1] == k
            nxt += 2
    assert nxt == n
    # Inner boxes enclose their children.
    for k in np.flatnonzero(~leaf):
        for c in (nodes["first"][k], nodes["second"][k]):
            assert np.all(nodes["min"][k] <= nodes["min"][c]) and np.all(nodes["max"][k] >= nodes["max"][c])
    # Leaf ranges tile the index array in in-order (left-to-right) order.
    order = []
    stack = [0]
    while stack:
        k = stack.pop()
        if leaf[k]:
            order.append(k)
        else:
            stack.append(nodes["second"][k])
            stack.append(nodes["first"][k])
    pos = 0
    for k in order:
        f, s = int(nodes["first"][k]), int(nodes["second"][k])
        assert f == pos and s >= f
        # Leaf box = union of its (possibly clipped) reference boxes.
        assert np.array_equal(nodes["min"][k], final_boxes["min"][f:s + 1].min(axis=0))
        assert np.array_equal(nodes["max"][k], final_boxes["max"][f:s + 1].max(axis=0))
        pos = s + 1
    assert pos == indices.size
    # Every primitive is referenced; a clipped piece lies inside the primitive's original box.
    assert set(indices.tolist()) == set(range(input_boxes.size))
    src = input_boxes[indices]
    assert np.all(final_boxes["min"] >= src["min"]) and np.all(final_boxes["max"] <= src["max"])


def tiny_tri(center, size=1e-3):
    c = np.asarray(center, np.float32)
    return np.array([c, c + [size, 0, 0], c + [0, size, 0]], np.float32)


def mesh_from_tris(tris):
    pts = np.concatenate(tris).astype(np.float32)
    idx = np.arange(pts.shape[0], dtype=np.uint32)
    return idx, pts


# ---- hand-derived cases -------------------------------------------------------------------


@pytest.mark.parametrize("count", [1, 2, 3])
def test_three_or_fewer_never_split(ensure_built, count):
    # Any split costs at least 7 (bvh.cpp:84) while not splitting costs 2 * count <= 6.
    idx, pts = mesh_from_tris([tiny_tri([3.0 * i, 0, 0]) for i in range(count)])
    boxes = orc.triangle_bounds(idx, pts)
    nodes, ind = assert_same_build(boxes, 1)
    assert nodes.size == 1 and nodes["first"][0] == 0 and nodes["second"][0] == count - 1
    assert nodes["type_and_parent"][0] == 0x80000000
    assert ind.tolist() == list(range(count))


def test_triangle_bounds_closed_form(ensure_built):
    pts = np.array([[0, 0, 0], [1, 2, 0], [0, 0, 3]], np.float32)
    b = orc.triangle_bounds(np.array([0, 1, 2], np.uint32), pts)[0]
    assert np.array_equal(b["min"], np.float32([0, 0, 0]) - np.float32(1e-5))
    assert np.array_equal(b["max"], np.float32([1, 2, 3]) + np.float32(1e-5))
    assert np.array_equal(b["centroid"], (pts[0] + pts[1] + pts[2]) / np.float32(3.0))
    assert b["has_centroid"] == 1
    p = m.triangle_bounds(np.array([0, 1, 2], np.uint32), pts)[0]
    assert p.tobytes() == b.tobytes()


def test_two_clusters(ensure_built):
    """Two clusters of 8 triangles with one shared centroid each, 10 apart on x.

    Root: x bins put cluster A in bin 0 and B in bin 15; every partition i >= 1 has the same cost,
    so the last one (i = 14) wins and separates the clusters; y and z centroid extents are 0, all
    references fall in bin 0 there and cost 7 + 2*(8 + 8) + ... > x's. Each cluster (one centroid)
    costs 7 + 2*(8 + 8) = 39 > 16 to split, so it stays a leaf of 8 in input order."""
    a = [tiny_tri([0, 0, 0]) for _ in range(8)]
    b = [tiny_tri([10, 0, 0]) for _ in range(8)]
    idx, pts = mesh_from_tris(a + b)
    boxes = orc.triangle_bounds(idx, pts)
    nodes, ind = assert_same_build(boxes, 1)
    assert nodes.size == 3
    assert nodes["first"][0] == 1 and nodes["second"][0] == 2
    assert (nodes["first"][1], nodes["second"][1]) == (0, 7)
    assert (nodes["first"][2], nodes["second"][2]) == (8, 15)
    assert ind.tolist() == list(range(16))


def test_wide_reference_is_box_split(ensure_built):
    """A triangle spanning the whole x range is wider than 5 % of the root extent: with
    split_aabbs it is clipped at the plane and referenced from both sides; with dont_split it is
    referenced once."""
    small = [tiny_tri([x, 0, 0]) for x in (0, 1, 2, 3, 7, 8, 9, 10)] * 2
    long = np.array([[0, 1, 0], [10, 1, 0], [0, 1.5, 0]], np.float32)
    idx, pts = mesh_from_tris(small + [long])
    boxes = orc.triangle_bounds(idx, pts)
    nodes, ind, fb = orc.bvh_build(boxes, True)
    assert (ind == 16).sum() >= 2
    check_invariants(nodes, ind, fb, boxes)
    assert_same_build(boxes, 1)
    nodes0, ind0, fb0 = orc.bvh_build(boxes, False)
    assert (ind0 == 16).sum() == 1
    check_invariants(nodes0, ind0, fb0, boxes)
    assert_same_build(boxes, 0)


# ---- seeded meshes: invariants and product == oracle --------------------------------------


def soup(n, seed, skinny=False):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    e = rng.normal(size=(n, 3, 3)) * (0.05 if not skinny else 0.02)
    if skinny:
        e[:, 1, 0] += rng.uniform(-4, 4, n)
    pts = (c + e).reshape(-1, 3).astype(np.float32)
    return np.arange(pts.shape[0], dtype=np.uint32), pts


CASES = [
    ("heightfield 40x20", lambda: scenes.heightfield(40, 20, seed=3)[:2]),
    ("heightfield 100x50 seed 9", lambda: scenes.heightfield(100, 50, seed=9)[:2]),
    ("icosphere 0", lambda: scenes.icosphere(0)[:2]),
    ("icosphere 2", lambda: scenes.icosphere(2)[:2]),
    ("icosphere 3", lambda: scenes.icosphere(3)[:2]),
    ("soup 3000", lambda: soup(3000, 1)),
    ("skinny soup 2000", lambda: soup(2000, 2, skinny=True)),
    ("flat grid", lambda: scenes.heightfield(30, 30, seed=1, height=0.0)[:2]),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("split", [0, 1])
def test_builder_matches_oracle(ensure_built, name, make, split):
    idx, pts = make()
    boxes = orc.triangle_bounds(idx, pts)
    assert m.triangle_bounds(idx, pts).tobytes() == boxes.tobytes()
    nodes, ind, fb = orc.bvh_build(boxes, split)
    check_invariants(nodes, ind, fb, boxes)
    assert_same_build(boxes, split)


def test_instance_boxes_without_centroid(ensure_built):
    """The instance BVH uses dont_split over boxes without a weighted centroid (scene.cpp:338)."""
    rng = np.random.default_rng(5)
    lo = rng.uniform(-5, 5, (500, 3)).astype(np.float32)
    hi = lo + rng.uniform(0.01, 1.0, (500, 3)).astype(np.float32)
    boxes = np.zeros(500, orc.AABB_DTYPE)
    boxes["min"], boxes["max"] = lo, hi
    nodes, ind, fb = orc.bvh_build(boxes, False)
    check_invariants(nodes, ind, fb, boxes)
    assert_same_build(boxes, 0)


def test_degenerate_inputs(ensure_built):
    # Identical triangles (zero centroid extent on every axis): NaN bins clamp to bin 0.
    idx, pts = mesh_from_tris([tiny_tri([1, 1, 1])] * 50)
    boxes = orc.triangle_bounds(idx, pts)
    nodes, ind = assert_same_build(boxes, 1)
    assert nodes.size == 1
    # Zero-area and NaN triangles mixed into a soup.
    idx, pts = soup(300, 7)
    pts[5] = pts[4]
    pts[6] = pts[4]
    pts[30] = np.nan
    boxes = orc.triangle_bounds(idx, pts)
    assert_same_build(boxes, 1)
    assert_same_build(boxes, 0)


def test_large_heightfield_matches(ensure_built):
    idx, pts, _, _ = scenes.heightfield(300, 150, seed=3)
    boxes = orc.triangle_bounds(idx, pts)
    assert_same_build(boxes, 1)
