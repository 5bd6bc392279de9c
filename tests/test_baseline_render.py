"""The CPU baseline's restated reference renderer (oracle/baseline_render.cpp, a timing baseline,
not a checker): it runs the reference's passes on the reference's small configuration, produces a
finite, positive image sum, and gives the same sum for any thread count and on every call (each
pixel has one path, and the pass seeds come from a fixed generator), so its timings measure the
same work."""
import numpy as np

import oracle as orc
from myrt_amd import scenes


def test_reference_render_baseline_is_deterministic(ensure_built):
    sc = scenes.sphere_field_scene(200, 2, seed=1)
    o = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"])
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.0)
    runs = [o.reference_render_time(view, proj, 128, 128, 4, threads=t, reps=1) for t in (1, 3, 1)]
    sums = [r[3] for r in runs]
    assert all(np.isfinite(sums)) and sums[0] > 0
    assert sums[0] == sums[1] == sums[2]
    for sps, rps, used, _ in runs:
        assert sps > 0 and rps >= sps  # at least one ray per sample
