"""Materials of caller-registered layouts render like the same values in the built-in record.

The shaders read each material through its type's bindings (src/pt/material_new.glsl:17-25
`open()` loads the same six inputs from the reference's generated loader), so a scene whose
materials come from a layout with other parameters around the bound ones, in another order and
nested, must give the image of the scene built from mrt_material records: the resolved records
are compared bit for bit, and the two renders agree to a few ulp (per-pixel radiance is summed with
float atomics, whose order varies between runs)."""
import sys

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build(sc, generic):
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    if generic:
        s.register_type("color_t", [("value", "vec4unorm8")])
        lay = s.register_type("layered", [("frame", "mat3"), ("transmission", "float"), ("rough", "float"),
                                          ("base", "color_t"), ("spare", "vec2"), ("metal", "float"),
                                          ("index", "float")])
        mt = s.push_material_type(lay, {"albedo": "base", "roughness": "rough", "metallic": "metal",
                                        "ior": "index"})
        for rec in sc["materials"]:
            r = np.asarray(rec).reshape(-1)[0]
            x = s.push_material(mt)
            s.set_parameter(x, "frame", np.eye(3, dtype=np.float32))
            s.set_parameter(x, "base", np.uint32(r["albedo_rgba8"]))
            s.set_parameter(x, "rough", np.float32(r["roughness"]))
            s.set_parameter(x, "metal", np.float32(r["metallic"]))
            s.set_parameter(x, "index", np.float32(r["ior"]))
            s.set_parameter(x, "transmission", np.float32(r["transmission"]))
            s.set_parameter(x, "spare", np.float32([7.0, -7.0]))
    else:
        for rec in sc["materials"]:
            s.push_material(rec)
    s.set_lights(sc["lights"])
    s.set_environment(1.0)
    for g, mtx, mi in sc["instances"]:
        s.enqueue(gids[g], mtx, mi)
    return s


def test_generic_layout_renders_like_the_record(ensure_built):
    sc = scenes.sphere_field_scene(40, 2, seed=3, glass_every=6, metal_every=5)
    W, H = 96, 64
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    imgs = []
    for generic in (False, True):
        say("render %dx%d, 8 spp, materials from %s" % (W, H, "a registered layout" if generic else "records"))
        s = build(sc, generic)
        for i in range(len(sc["materials"])):
            a = np.asarray(sc["materials"][i]).reshape(-1)[0]
            b = s.get_material(i)[0]
            for k in ("albedo_rgba8", "ior", "roughness", "metallic", "transmission", "texture"):
                assert a[k] == b[k], (i, k)
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(6)
        r.render(s, W, H, 8)
        imgs.append(r.image().copy())
        r.close()
        s.close()
    assert np.isfinite(imgs[0]).all() and imgs[0][..., :3].mean() > 0.01
    diff = np.abs(imgs[0].astype(np.float64) - imgs[1])
    say("largest pixel difference %.3g" % diff.max())
    assert diff.max() <= 1e-5 * max(1.0, float(np.abs(imgs[0]).max()))
