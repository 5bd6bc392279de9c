"""The CPU replica of the traversal kernel's fetches (oracle/fetch_count.cpp), the source of the
roofline's node / triangle / instance counts per ray, walks the product's own device-format
arrays (no GPU needed to build them). It must be a correct traversal: its closest hits equal the
brute-force hits bit for bit (ids and t), its any-hit flags equal the closest-hit existence, on
an instanced scene and on a single-mesh heightfield. Its totals against the kernel's own
counters are a GPU test (test_trace_gpu.py)."""
import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes


def host_scene(geometries, instances):
    s = m.Scene()
    gids = [s.push_geometry(i, p) for i, p in geometries]
    for g, mt in instances:
        s.enqueue(gids[g], mt)
    s.build_host()
    return s


def random_rays(n, seed, box=6.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-box, box, (n, 3))
    d = rng.normal(size=(n, 3))
    return m.make_rays(o, d / np.linalg.norm(d, axis=1, keepdims=True))


@pytest.mark.parametrize("case", ["instanced spheres", "heightfield"])
def test_replica_is_a_correct_traversal(ensure_built, case):
    if case == "instanced spheres":
        (fi, fp, _), mats = scenes.sphere_field(300, 2, seed=3)
        geoms, insts = [(fi, fp)], [(0, mt) for mt in mats]
        rays = random_rays(20000, 5)
    else:
        idx, pts, _, _ = scenes.heightfield(200, 100, seed=3)
        geoms, insts = [(idx, pts)], [(0, np.eye(4, dtype=np.float32))]
        rays = random_rays(20000, 6, box=4.0)
    s = host_scene(geoms, insts)
    wide = s.wide_arrays()
    counts, hits = orc.fetch_counts(wide, rays)
    ref = orc.OracleScene(geoms, insts).brute(rays)
    assert np.array_equal(hits["inst"], ref["inst"]) and np.array_equal(hits["prim"], ref["prim"])
    f = hits["inst"] != 0xFFFFFFFF
    assert f.mean() > 0.1
    assert np.array_equal(hits["t"][f].view(np.uint32), ref["t"][f].view(np.uint32))
    _, occ = orc.fetch_counts(wide, rays, any_hit=True)
    assert np.array_equal(np.isfinite(occ["t"]), f)
    # Every ray enters the scene (the TLAS root, or the single instance's BLAS root).
    assert (counts[:, 0] >= 1).all()
    if case == "heightfield":
        assert (counts[:, 2] == 1).all()  # the single instance, entered once per ray
