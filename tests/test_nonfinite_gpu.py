"""Non-finite values in the render path (pt_color.comp:84-98, 156-166; pt_trace.comp:186-189).

The reference discards NaN/Inf colours before and after accumulating, ends paths whose BSDF
sample is not finite, and treats a hit with a non-finite normal as a miss. The product does the
same and counts what it rejected (mrt_render_stats.nonfinite): a non-finite geometric normal
makes the hit a miss, a non-finite shading normal falls back to the geometric one, non-finite
contributions are dropped and a non-finite throughput ends the path. The FP64 oracle applies the
same rules, so a scene with NaN vertex normals and zero-area triangles still renders within the
0.5 % relative-MSE bar."""
import sys

import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import scenes

pytestmark = pytest.mark.gpu
REL_MSE_TOL = 0.005


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def degenerate_triangles():
    """A strip of zero-area triangles (repeated vertices, collinear vertices, a point) in the
    camera's view, next to two ordinary triangles that keep the geometry's box non-empty."""
    pts = np.array([
        [-2.0, -1.0, 0.0], [2.0, -1.0, 0.0], [0.0, 1.0, 0.0],      # an ordinary triangle
        [-1.0, -1.0, 0.0], [0.0, 0.0, 0.0], [1.0, 1.0, 0.0],       # collinear
        [0.5, 0.5, 0.2], [0.5, 0.5, 0.2], [1.5, 0.0, 0.2],        # two equal vertices
        [0.3, 0.3, 0.3], [0.3, 0.3, 0.3], [0.3, 0.3, 0.3],        # one point
        [-3.0, 0.0, -0.5], [3.0, 0.0, -0.5], [0.0, 0.000001, -0.5],  # a sliver along x
    ], np.float32)
    idx = np.arange(pts.shape[0], dtype=np.uint32)
    nrm = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (pts.shape[0], 1))
    return idx, pts, nrm


def poisoned_scene():
    """The 40-sphere field with NaN normals on every 7th vertex of the sphere mesh and a
    degenerate-triangle strip in front of the camera."""
    sc = scenes.sphere_field_scene(40, 2, seed=3, glass_every=6, metal_every=5)
    fi, fp, fn = sc["geometries"][0]
    fn = np.array(fn, np.float32, copy=True)
    fn[::7] = np.nan
    sc["geometries"][0] = (fi, fp, fn)
    di, dp, dn = degenerate_triangles()
    sc["geometries"].append((di, dp, dn))
    sc["instances"].append((len(sc["geometries"]) - 1, scenes.translate_scale((0.0, -1.0, -4.0), 1.0),
                            len(sc["materials"]) - 1))
    return sc


def gpu_scene(sc):
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(1.0)
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    return s


def test_nan_normals_and_degenerate_triangles_match_fp64_oracle(ensure_built):
    sc = poisoned_scene()
    W, H, spp, depth = 96, 64, 32, 6
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    say("render with NaN vertex normals and zero-area triangles %dx%d %d spp" % (W, H, spp))
    s = gpu_scene(sc)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    st = r.stats()
    o = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0)
    ref = o.render(view, proj, W, H, spp, depth)
    err = float(((img - ref) ** 2).sum() / (ref ** 2).sum())
    say("relative MSE %.3e, rejected: product %d, oracle %d" % (err, st["nonfinite"], o.last_nonfinite))
    assert np.isfinite(img).all()
    assert np.isfinite(ref).all()
    assert st["nonfinite"] > 0 and o.last_nonfinite > 0
    # Both sides reject at the same events up to the hit set's bit-exact agreement; the counts
    # differ only where fp32 and fp64 shading diverge on the way (a few per mille).
    assert abs(int(st["nonfinite"]) - o.last_nonfinite) <= 0.01 * o.last_nonfinite + 16
    assert err <= REL_MSE_TOL


def test_nonfinite_material_light_and_camera_give_a_finite_image(ensure_built):
    """A NaN roughness, an infinite light and a camera placed at 3e38 (its rays overflow) leave the
    accumulated image finite; every rejected value is counted."""
    sc = scenes.sphere_field_scene(20, 2, seed=5)
    s = gpu_scene(sc)
    s.set_parameter(3, "roughness", [float("nan")])
    s.set_parameter(4, "ior", [float("inf")])
    lights = np.array(sc["lights"], copy=True)
    lights["emission"][0] = (np.inf, 1.0, 1.0)
    s.set_lights(lights)
    W, H = 64, 48
    r = m.Renderer()
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    r.set_camera(view, proj)
    r.set_num_bounces(5)
    say("render with a NaN material and an infinite light")
    r.render(s, W, H, 16)
    img = r.image()
    n1 = r.stats()["nonfinite"]
    say("rejected %d" % n1)
    assert np.isfinite(img).all()
    assert n1 > 0
    far = np.array(view, np.float64, copy=True)
    far[:3, 3] = 3e38  # the eye's translation: generated rays overflow to inf/NaN
    r.set_camera(far.astype(np.float32), proj)
    say("render from a camera at 3e38")
    r.render(s, W, H, 4)
    img = r.image()
    n2 = r.stats()["nonfinite"]
    say("rejected %d in total" % n2)
    assert np.isfinite(img).all()
    assert n2 > n1
