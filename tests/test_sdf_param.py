"""Runtime SDF parameter updates, host side (scene::set_parameter(sdf, parameter, value),
src/pathtracer/scene.hpp:150-153; sdf.cpp:346-407; parameter.cpp:57-109).

mrt_scene_set_sdf_param writes a pushed program's constants in place: its bounds, compiled form
and drawable boxes follow at once, and its culling grid keeps its frame and is rebuilt on the GPU at the
next prepare. Frames, bounds and drawables after updates equal those of the edited program pushed
fresh; invalid values are rejected and change nothing."""
import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes


def translate_node(i):
    return 0 if i == 0 else 3 + 4 * (i - 1)


def scene_with(tape):
    s = m.Scene()
    sid = s.push_sdf(tape)
    s.enqueue_sdf(sid, np.eye(4, dtype=np.float32))
    s.build_host()
    return s, sid


def test_errors_leave_the_program_unchanged(ensure_built):
    tape = scenes.sdf_cluster(16, seed=3)
    s, sid = scene_with(tape)
    before = s.sdf_drawables(1).copy()
    dim0, lo0, inv0, masks0 = s.sdf_grid(sid)
    for args, msg in (((sid + 1, 0, 0, [1.0]), "unknown sdf"), ((sid, len(tape), 0, [1.0]), "node index"),
                      ((sid, 1, 6, [1.0]), "slots out of range"), ((sid, 1, 4, [1.0, 2.0, 3.0]), "slots out of range"),
                      ((sid, 1, 0, [float("nan")]), "finite"), ((sid, 1, 0, [float("inf")]), "finite"),
                      ((sid, translate_node(3) + 3, 0, [0.0]), "smooth operator needs k > 0")):
        with pytest.raises(m.MrtError, match=msg):
            s.set_sdf_param(*args)
    assert np.array_equal(s.sdf_drawables(1).view(np.uint8), before.view(np.uint8))
    dim1, lo1, inv1, masks1 = s.sdf_grid(sid)
    assert dim1 == dim0 and np.array_equal(lo0, lo1) and np.array_equal(masks0, masks1)
    s.close()


def test_bounds_and_grid_frame_follow_an_update(ensure_built):
    tape = scenes.sdf_cluster(16, seed=3)
    s, sid = scene_with(tape)
    hi0 = s.sdf_drawables(1)[0]["box_hi"].copy()
    dim, lo, inv, _ = s.sdf_grid(sid)
    assert dim == 64  # 16 chain instructions: 64 cells per axis
    # Move primitive 0 to x = 10: the drawable's box and the grid's frame grow to include it.
    s.set_sdf_param(sid, translate_node(0), 0, [10.0, 0.0, 0.0])
    hi1 = s.sdf_drawables(1)[0]["box_hi"]
    assert hi1[0] > 10.0 > hi0[0]
    dim1, lo1, inv1, _ = s.sdf_grid(sid, with_masks=False)
    assert dim1 == 64 and inv1[0] < inv[0]
    # The masks are rebuilt on the GPU at the next prepare; until then they cannot be read.
    with pytest.raises(m.MrtError, match="next prepare"):
        s.sdf_grid(sid)
    s.close()


def test_updated_frame_matches_a_fresh_push(ensure_built):
    tape = scenes.sdf_cluster(16, seed=3)
    s, sid = scene_with(tape)
    sphere = translate_node(4) + 1  # primitive 4 is a sphere
    s.set_sdf_param(sid, sphere, 0, [0.3])
    s.set_sdf_param(sid, translate_node(7), 1, [0.5])
    with pytest.raises(m.MrtError, match="sphere radius must be positive"):
        s.set_sdf_param(sid, sphere, 0, [0.0])
    edited = tape.copy()
    edited[sphere]["p"][0] = 0.3
    edited[translate_node(7)]["p"][1] = 0.5
    f, fid = scene_with(edited)
    d0, lo0, inv0, _ = s.sdf_grid(sid, with_masks=False)
    d1, lo1, inv1, masks = f.sdf_grid(fid)
    assert d0 == d1 == 64 and np.array_equal(lo0, lo1) and np.array_equal(inv0, inv1)
    assert masks is not None and masks.any()
    assert np.array_equal(s.sdf_drawables(1).view(np.uint8), f.sdf_drawables(1).view(np.uint8))
    s.close()
    f.close()
