"""Runtime SDF parameter updates on the GPU: after mrt_scene_set_sdf_param and a prepare, the
culling grid rebuilt on the GPU equals, mask for mask, the host-built grid of the edited program
pushed fresh, and closest hits equal that scene's bit for bit. Covers a 64-instruction chain (two
grid passes, 128 cells per axis) and a 6-instruction chain (one pass, 32 cells), and times an
update against a re-push."""
import sys
import time

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def translate_node(i):
    return 0 if i == 0 else 3 + 4 * (i - 1)


def scene_with(tape):
    s = m.Scene()
    sid = s.push_sdf(tape)
    s.enqueue_sdf(sid, scenes.translate_scale((0.0, 0.0, 0.0), 1.5))
    return s, sid


def rays_at_cluster(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3)) * 5.0
    tgt = rng.uniform(-2.0, 2.0, size=(n, 3))
    d = tgt - o
    return m.make_rays(o, d / np.linalg.norm(d, axis=1, keepdims=True))


@pytest.mark.parametrize("n_prims", [64, 6])
def test_gpu_grid_rebuild_matches_a_fresh_push(ensure_built, n_prims):
    tape = scenes.sdf_cluster(n_prims, seed=12)
    s, sid = scene_with(tape)
    s.prepare()
    edits = [(translate_node(1), 0, [0.9, -0.4, 0.2]), (translate_node(2) + 1, 0, [0.35]),
             (translate_node(n_prims - 1) + 3, 0, [0.2])]
    edited = tape.copy()
    t0 = time.perf_counter()
    for node, slot, vals in edits:
        s.set_sdf_param(sid, node, slot, vals)
        edited[node]["p"][slot:slot + len(vals)] = vals
    s.prepare()
    m.synchronize()
    t_update = time.perf_counter() - t0
    t0 = time.perf_counter()
    f, fid = scene_with(edited)
    f.prepare()
    m.synchronize()
    t_push = time.perf_counter() - t0
    dim, lo, inv, masks = s.sdf_grid(sid)
    dim_f, lo_f, inv_f, masks_f = f.sdf_grid(fid)
    say("%d primitives: grid %d^3, update + prepare %.2f ms, fresh push + prepare %.2f ms" % (
        n_prims, dim, t_update * 1e3, t_push * 1e3))
    assert dim == dim_f == (128 if n_prims == 64 else 32)
    assert np.array_equal(lo, lo_f) and np.array_equal(inv, inv_f)
    assert np.array_equal(masks, masks_f)
    rays = rays_at_cluster(1 << 16, 3)
    a, b = s.trace_closest(rays), f.trace_closest(rays)
    assert ((a["inst"] != m.MISS) & ((a["inst"] & m.SDF_HIT) != 0)).mean() > 0.2
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    s.close()
    f.close()
