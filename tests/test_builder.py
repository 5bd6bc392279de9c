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


# ---- the reference's SAH quirks, each pinned by a case the textbook SAH splits elsewhere ------

LEAF = 0x80000000


def boxes_of(spec):
    """Boxes without a weighted centroid (centroid = centre) from (centre, half extents) pairs."""
    b = np.zeros(len(spec), orc.AABB_DTYPE)
    for k, (c, h) in enumerate(spec):
        b["min"][k] = np.asarray(c, np.float32) - np.asarray(h, np.float32)
        b["max"][k] = np.asarray(c, np.float32) + np.asarray(h, np.float32)
    return b


def nodes_of(rows):
    """Expected node records from (min, leaf, parent, max, first, second) rows."""
    out = np.zeros(len(rows), orc.NODE_DTYPE)
    for k, (mn, leaf, parent, mx, first, second) in enumerate(rows):
        out[k] = (mn, (LEAF if leaf else 0) | parent, mx, first, second, (0, 0, 0))
    return out


def textbook_split_plane(boxes):
    """The textbook binned SAH on x (16 bins over the centroid bounds, left = bins 0..i, right =
    bins i+1..15 with its own count, cost 7 + 2 (n_l A_l + n_r A_r) / A, last minimum wins, plane
    at the upper edge of bin i). For the cases below the centroids vary in x only."""
    c = (boxes["min"][:, 0].astype(np.float64) + boxes["max"][:, 0]) / 2
    lo, step = c.min(), (c.max() - c.min()) / 16
    b = np.clip(((c - lo) / step).astype(int), 0, 15)

    def area(sel):
        mn, mx = boxes["min"][sel].min(0).astype(np.float64), boxes["max"][sel].max(0).astype(np.float64)
        d = mx - mn
        return ((d[0] + d[1]) * d[2] + d[0] * d[1]) * 2

    total = area(np.ones(len(boxes), bool))
    best, plane = np.inf, None
    for i in range(15):
        left, right = b <= i, b > i
        if not left.any() or not right.any():
            continue
        cost = 7 + 2 * (left.sum() * area(left) + right.sum() * area(right)) / total
        if cost <= best:
            best, plane = cost, lo + step * (i + 1)
    return plane


def test_sah_second_count_is_the_prefix_count(ensure_built):
    """bvh.cpp:96 counts the second side as accum_counts[15 - i], the references in bins
    0..15-i, not those in bins i+1..15.

    Boxes (centre x; full size x*y*z; surface area): L at 0 (2*1*2; 16), M at 10 (2*1*2; 16),
    three R at 16 (2*2*1; 16). Root box 18*2*2, A = 152; A(M u R) = 8*2*2 -> 72, A(L u M) =
    12*1*2 -> 76. Centroid bins (step 1): L 0, M 10, R 15; y and z extents are 0 (all in bin 0,
    every i >= 1 there has an empty right box of area +inf).
    Cost 7 + 2 S / 152 with S = n_1 A_1 + accum_counts[15-i] A(bins i..15):
      i in 1..5:  1*16 + 2*72 = 160 (bins 0..10-14 hold L and M);
      i in 6..9:  1*16 + 1*72 = 88  (bins 0..6-9 hold L only)  <- minimum, last one wins: i = 9;
      i = 10:     2*76 + 1*72 = 224;  i in 11..14: 2*76 + 1*16 = 168;  i = 0: 1*16 + 5*152.
    7 + 176/152 = 8.16 < 2 * 5: split at x = 10, so M (centroid 10, not < 10) goes right:
    {L} | {M, R, R, R}. The right child (centroids 10 and 16, step 0.375; area 72) costs
    7 + 2 (1*16 + 1*16) / 72 = 7.89 < 8 at every i >= 1, splits at 10 + 0.375*15 = 15.625 into
    {M} | {R, R, R}; the three R share a centroid and stay a leaf.
    The textbook SAH (n - P(i) on the right) ends at i = 14 instead: 7 + 2 (2*76 + 3*16) / 152 =
    9.63 against 11 for i <= 9, so its root split is {L, M} | {R, R, R} at x = 15."""
    boxes = boxes_of([((0, 0, 0), (1, 0.5, 1)), ((10, 0, 0), (1, 0.5, 1))] + [((16, 0, 0), (1, 1, 0.5))] * 3)
    want = nodes_of([
        ((-1, -1, -1), False, 0, (17, 1, 1), 1, 2),
        ((-1, -0.5, -1), True, 0, (1, 0.5, 1), 0, 0),
        ((9, -1, -1), False, 0, (17, 1, 1), 3, 4),
        ((9, -0.5, -1), True, 2, (11, 0.5, 1), 1, 1),
        ((15, -1, -0.5), True, 2, (17, 1, 0.5), 2, 4),
    ])
    nodes, ind, _ = orc.bvh_build(boxes, False)
    assert nodes.tobytes() == want.tobytes()
    assert ind.tolist() == [0, 1, 2, 3, 4]
    assert_same_build(boxes, 0)
    assert textbook_split_plane(boxes) == 15.0  # the textbook root split differs (x = 10 above)


def test_sah_second_box_includes_bin_i(ensure_built):
    """bvh.cpp:98 takes the second box as accum_aabbsr[i], bins i..15, so bin i is on both sides.

    Boxes (centre x; size; area): B0 at 0 (1*1*2; 10), B1 at 7 (1*4*1; 18), B2 at 8, B3 at 9,
    B4 at 12 (1*1*1; 6 each), B5 at 16 (1*2*2; 16). Root 17*4*2, A = 220. Bins (step 1): 0, 7, 8,
    9, 12, 15; prefix counts P: 1 (i < 7), 2, 3, 4 (9..11), 5 (12..14), 6.
    Boxes of bins 0..i: 10 (i < 7), 112 (7), 124 (8), 136 (9..11), 172 (12..14). Boxes of bins
    i..15: 136 (1..7), 80 (8), 72 (9), 48 (10..12), 16 (13, 14).
    S = P(i) A(0..i) + P(15-i) A(i..15):  i = 4..6: 10 + 4*136 = 554;  i = 7: 2*112 + 3*136 = 632;
    i = 8: 3*124 + 2*80 = 532 (minimum);  i = 9: 4*136 + 1*72 = 616;  i = 10, 11: 592;  i >= 12:
    876 or more. 7 + 2*532/220 = 11.84 < 12: split at x = 9 into {B0, B1, B2} | {B3, B4, B5};
    three references never split (7 > 6). With the second box over bins i+1..15 instead, i = 7
    would cost 2*112 + 3*80 = 464 < 3*124 + 2*72 = 516 and split at x = 8; the textbook SAH also
    splits at x = 8 (2*112 + 4*80 = 544 against 588 at i = 8)."""
    boxes = boxes_of([((0, 0, 0), (0.5, 0.5, 1)), ((7, 0, 0), (0.5, 2, 0.5)), ((8, 0, 0), (0.5, 0.5, 0.5)),
                      ((9, 0, 0), (0.5, 0.5, 0.5)), ((12, 0, 0), (0.5, 0.5, 0.5)), ((16, 0, 0), (0.5, 1, 1))])
    want = nodes_of([
        ((-0.5, -2, -1), False, 0, (16.5, 2, 1), 1, 2),
        ((-0.5, -2, -1), True, 0, (8.5, 2, 1), 0, 2),
        ((8.5, -1, -1), True, 0, (16.5, 1, 1), 3, 5),
    ])
    nodes, ind, _ = orc.bvh_build(boxes, False)
    assert nodes.tobytes() == want.tobytes()
    assert ind.tolist() == list(range(6))
    assert_same_build(boxes, 0)
    assert textbook_split_plane(boxes) == 8.0


def test_box_split_pieces_go_in_front_and_the_partition_swaps(ensure_built):
    """split_aabbs (bvh.cpp:29-41, 408-512): a reference wider than 5 % of the root extent on the
    split axis is clipped at the plane; the lower piece is inserted at the front of the node's
    range (:460-465), the upper piece stays in place, and the partition scans forward and swaps
    (:477-486), which is not stable on the right side.

    Boxes on x (y, z in [-0.5, 0.5]): S0 [-0.25, 0.25], W1 [-2, 6], S2 [15.75, 16.25], S3 = S0,
    W4 [10, 18], S5 = S2, S6 = S0, S7 = S2. Root 20*1*1 (A = 82), threshold 0.05 * 20 = 1.
    Bins (centroids 0..16, step 1): S at 0 and 15, W1 (centre 2) at 2, W4 (centre 14) at 14; areas
    S 4, W 34, bins 0..i: 34 for 2 <= i <= 13, bins i..15: 34 for 3 <= i <= 13, P(i) = 4 there:
    S = 4*34 + 4*34 = 272, the minimum; i = 13 wins (last), cost 7 + 544/82 = 13.6 < 16: plane 14.
    W1 lies below 14 and stays whole; W4 becomes [10, 14] (inserted at the front, index 4) and
    [14, 18] (in place). Range [L4, S0, W1, S2, S3, W4hi, S5, S6, S7]: the scan passes L4 (centre
    12), S0, W1 and stops at S2; the swap loop then moves S3 to position 3 and S6 to position 4.
    Indices 4 0 1 3 6 | 4 5 2 7. The left child (centroids 0..12, A = 66) costs at best
    7 + 2 (4*34 + 3*18) / 66 = 12.8 > 10, the right one shares one centroid: both are leaves."""
    xs = [(-0.25, 0.25), (-2, 6), (15.75, 16.25), (-0.25, 0.25), (10, 18), (15.75, 16.25), (-0.25, 0.25),
          (15.75, 16.25)]
    boxes = boxes_of([(((a + b) / 2, 0, 0), ((b - a) / 2, 0.5, 0.5)) for a, b in xs])
    want = nodes_of([
        ((-2, -0.5, -0.5), False, 0, (18, 0.5, 0.5), 1, 2),
        ((-2, -0.5, -0.5), True, 0, (14, 0.5, 0.5), 0, 4),
        ((14, -0.5, -0.5), True, 0, (18, 0.5, 0.5), 5, 8),
    ])
    nodes, ind, fb = orc.bvh_build(boxes, True)
    assert nodes.tobytes() == want.tobytes()
    assert ind.tolist() == [4, 0, 1, 3, 6, 4, 5, 2, 7]
    assert fb["min"][:, 0].tolist() == [10, -0.25, -2, -0.25, -0.25, 14, 15.75, 15.75, 15.75]
    assert fb["max"][:, 0].tolist() == [14, 0.25, 6, 0.25, 0.25, 18, 16.25, 16.25, 16.25]
    assert_same_build(boxes, 1)


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
