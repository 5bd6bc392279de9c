"""An OBJ scene imported through the library (load_object_file) renders like the FP64 oracle
fed by the oracle's own OBJ parse and material mapping (oracle/obj_ref.py).

Bar: relative MSE <= 0.5 % at 32 spp with the same sampler, as for every render parity test."""
import os
import sys

import numpy as np
import pytest

import myrt_amd as m
import oracle as orc
from myrt_amd import obj as mobj
from myrt_amd import scenes
from oracle import obj_ref

pytestmark = pytest.mark.gpu
PODIUM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "obj", "podium.obj")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_imported_podium_matches_fp64_oracle(ensure_built):
    lights = np.zeros(1, m.LIGHT_DTYPE)
    lights[0] = ((-10.0, 50.0, -10.0), (20.0, 0.0, 0.0), (0.0, 0.0, 20.0), (20.0, 19.0, 18.0))
    s = m.Scene()
    objs = mobj.load_object_file(s, PODIUM, import_scale=6.0)
    mobj.enqueue_objects(s, objs)
    s.set_lights(lights)
    s.set_environment(0.5)

    groups = obj_ref.load_groups(PODIUM)
    geoms = [(g["indices"], g["positions"], g["normals"], g["texcoords"]) for g in groups]
    mats = np.zeros(len(groups), m.MATERIAL_DTYPE)
    for i, g in enumerate(groups):
        rgba, ior, rough, trans = obj_ref.material_record(g["material"])
        mats[i]["albedo_rgba8"], mats[i]["ior"], mats[i]["roughness"], mats[i]["transmission"] = rgba, ior, rough, trans
    scale = np.diag([6.0, 6.0, 6.0, 1.0]).astype(np.float32)
    insts = [(i, scale, i) for i in range(len(groups))]

    W, H, spp, depth = 96, 64, 32, 4
    view, proj = scenes.camera_matrices((0.0, 40.0, 70.0), (0.0, 0.0, 0.0), 40.0, W / H)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    say("imported podium %dx%d %d spp depth %d" % (W, H, spp, depth))
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    ref = orc.OracleRender(geoms, insts, mats, lights, 0.5).render(view, proj, W, H, spp, depth)
    err = float(((img - ref) ** 2).sum() / (ref ** 2).sum())
    say("relative MSE %.3e" % err)
    assert np.isfinite(img).all() and img.mean() > 0.01
    assert err <= 0.005
    s.close()
