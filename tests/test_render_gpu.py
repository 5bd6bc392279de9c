"""GPU wavefront path tracer against the FP64 oracle with the same sampler sequence.

Bar: relative MSE sum((gpu - oracle)^2) / sum(oracle^2) <= 0.5 % on the converged images, for
scenes mixing GGX, metal, glass and Lambert surfaces under a quad light and the sky."""
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


def gpu_scene(sc, env=1.0):
    s = m.Scene()
    gids = [s.push_geometry(i, p, n) for i, p, n in sc["geometries"]]
    for mat in sc["materials"]:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(env)
    for g, mt, mi in sc["instances"]:
        s.enqueue(gids[g], mt, mi)
    return s


def rel_mse(a, b):
    return float(((a - b) ** 2).sum() / (b ** 2).sum())


@pytest.mark.parametrize("depth", [4, 8])
def test_render_matches_fp64_oracle(ensure_built, depth):
    sc = scenes.sphere_field_scene(40, 2, seed=1, glass_every=6, metal_every=5)
    W, H, spp = 96, 64, 32
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    say("render %dx%d %d spp depth %d" % (W, H, spp, depth))
    s = gpu_scene(sc)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    ref = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0).render(
        view, proj, W, H, spp, depth)
    err = rel_mse(img, ref)
    say("relative MSE %.3e" % err)
    assert np.isfinite(img).all()
    assert err <= REL_MSE_TOL


def test_progressive_accumulation(ensure_built):
    """Two calls of 8 spp equal one call of 16 spp (the sampler continues the sample index)."""
    sc = scenes.sphere_field_scene(10, 1, seed=2)
    W, H = 32, 24
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    s = gpu_scene(sc)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(4)
    r.render(s, W, H, 8)
    r.render(s, W, H, 8)
    assert r.sample_count() == 16
    a = r.image()
    r2 = m.Renderer()
    r2.set_camera(view, proj)
    r2.set_num_bounces(4)
    r2.render(s, W, H, 16)
    b = r2.image()
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5)


def test_direct_light_closed_form_on_gpu(ensure_built):
    gi, gp, gn = scenes.ground_quad(size=20.0, y=0.0)
    s = m.Scene()
    g = s.push_geometry(gi, gp, gn)
    mat = s.push_material(m.material((0.5, 0.5, 0.5), ior=1.0, roughness=1.0))
    s.set_lights(m.quad_light((-0.5, 2.0, -0.5), (1, 0, 0), (0, 0, 1), (10, 10, 10)))
    s.set_environment(0.0)
    s.enqueue(g, None, mat)
    view, proj = scenes.camera_matrices((0, 5, 0.001), (0, 0, 0), 1.0, 1.0)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(2)
    r.render(s, 3, 3, 8192)
    n = 400
    xs = (np.arange(n) + 0.5) / n - 0.5
    X, Z = np.meshgrid(xs, xs)
    r2 = X ** 2 + Z ** 2 + 4.0
    E = float((10 * (2.0 / np.sqrt(r2)) ** 2 / r2).sum() / n ** 2)
    expect = (128 / 255) / np.pi * E
    got = r.image()[1, 1, 0]
    assert abs(got - expect) / expect < 0.01


def test_render_with_sdf_matches_fp64_oracle(ensure_built):
    sc = scenes.sphere_field_scene(20, 2, seed=4, glass_every=0, metal_every=3)
    tape = scenes.sdf_cluster(64, seed=12, n_materials=4)
    W, H, spp, depth = 80, 60, 32, 4
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    s = gpu_scene(sc)
    sid = s.push_sdf(tape)
    s.enqueue_sdf(sid, scenes.translate_scale((0.0, -1.0, 0.0), 1.5))
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    dr = s.sdf_drawables(1)
    ref = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0, sdf_tape=tape,
                           sdf_drawables=dr).render(view, proj, W, H, spp, depth)
    err = rel_mse(img, ref)
    say("sdf render relative MSE %.3e" % err)
    assert err <= REL_MSE_TOL


def hexagon_mask(n=32):
    """A coloured hexagonal aperture: inside rgb (1, 0.8, 0.6), outside black."""
    ys, xs = np.mgrid[0:n, 0:n]
    u, v = (xs + 0.5) / n - 0.5, (ys + 0.5) / n - 0.5
    inside = np.ones((n, n), bool)
    for k in range(6):
        a = np.pi / 3 * k
        inside &= u * np.cos(a) + v * np.sin(a) <= 0.4
    mask = np.zeros((n, n, 3), np.float32)
    mask[inside] = (1.0, 0.8, 0.6)
    return mask


@pytest.mark.parametrize("masked", [False, True])
def test_depth_of_field_matches_fp64_oracle(ensure_built, masked):
    sc = scenes.sphere_field_scene(40, 2, seed=1, glass_every=6, metal_every=5)
    W, H, spp, depth = 80, 60, 32, 4
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    mask = hexagon_mask() if masked else None
    say("dof render %dx%d %d spp masked=%s" % (W, H, spp, masked))
    s = gpu_scene(sc)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    r.set_dof(True, (1.2, 1.2), 12.0, bokeh_mask=mask)
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    o = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0)
    o.set_dof(True, (1.2, 1.2), 12.0, bokeh_mask=mask)
    ref = o.render(view, proj, W, H, spp, depth)
    err = rel_mse(img, ref)
    say("dof relative MSE %.3e" % err)
    assert err <= REL_MSE_TOL


def test_textures_and_cubemap_match_fp64_oracle(ensure_built):
    """A checkerboard albedo texture on the ground and on some spheres (icosphere uvs are zero,
    so those read one texel), lit by the quad light and a six-face cube map."""
    sc = scenes.sphere_field_scene(30, 2, seed=5, metal_every=4, ground=False)
    gi, gp, gn, guv = scenes.textured_ground(size=8.0, y=-5.0, tiles=3.0)
    geoms = list(sc["geometries"]) + [(gi, gp, gn, guv)]
    mats = sc["materials"].copy()
    mats[-1]["texture"] = 1
    mats[3]["texture"] = 2
    insts = list(sc["instances"]) + [(1, np.eye(4, dtype=np.float32), len(mats) - 1)]
    tex = [scenes.checker_texture(64, 8), scenes.checker_texture(16, 2, a=(0.9, 0.2, 0.2), b=(0.2, 0.9, 0.2))]
    cube = scenes.gradient_cubemap(32)
    W, H, spp, depth = 80, 60, 32, 4
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    s = m.Scene()
    gids = [s.push_geometry(*g) for g in geoms]
    for t in tex:
        s.push_texture(t)
    for mat in mats:
        s.push_material(mat)
    s.set_lights(sc["lights"])
    s.set_environment(0.8)
    for g, mt, mi in insts:
        s.enqueue(gids[g], mt, mi)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    r.set_cubemap(cube)
    say("textured render %dx%d %d spp" % (W, H, spp))
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    o = orc.OracleRender(geoms, insts, mats, sc["lights"], 0.8, textures=tex)
    o.set_cubemap(cube)
    ref = o.render(view, proj, W, H, spp, depth)
    err = rel_mse(img, ref)
    say("textured relative MSE %.3e" % err)
    assert err <= REL_MSE_TOL


def test_material_sorted_shading_gives_the_same_image(ensure_built):
    """Shading each bounce in material order reorders only the atomic adds into the accumulator:
    the image equals the unsorted one up to fp32 summation order."""
    sc = scenes.sphere_field_scene(40, 2, seed=3, glass_every=4, metal_every=3)
    W, H, spp, depth = 96, 64, 16, 6
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    imgs = []
    for sort in (False, True):
        s = gpu_scene(sc)
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(depth)
        r.set_material_sort(sort)
        say("render %dx%d %d spp, material sort %s" % (W, H, spp, sort))
        r.render(s, W, H, spp)
        imgs.append(r.image()[..., :3].astype(np.float64))
        s.close()
    diff = np.abs(imgs[0] - imgs[1])
    say("max difference %.3e, image mean %.3f" % (diff.max(), imgs[0].mean()))
    assert imgs[0].mean() > 0.01
    assert np.all(diff <= 1e-5 * np.maximum(1.0, np.abs(imgs[0])))


def test_shadow_stream_gives_the_same_image(ensure_built, monkeypatch):
    """Tracing each bounce's shadow rays on a second stream, concurrently with the next bounce's
    closest hits, keeps every data dependency: the image equals the one-stream render up to fp32
    summation order of the atomic adds. The scene has an SDF cluster, so both the triangle and the
    SDF shadow passes run in the second trace context."""
    sc = scenes.sphere_field_scene(40, 2, seed=5, glass_every=4, metal_every=3)
    tape = scenes.sdf_cluster(16, seed=2, k=0.1, box=0.8, n_materials=2)
    W, H, spp, depth = 96, 64, 16, 6
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    imgs = []
    for overlap in ("0", "1"):
        monkeypatch.setenv("MRT_SHADOW_STREAM", overlap)
        s = gpu_scene(sc)
        s.enqueue_sdf(s.push_sdf(tape), scenes.translate_scale(tuple(sc["target"]), 1.0))
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(depth)
        say("render %dx%d %d spp, shadow stream %s" % (W, H, spp, overlap))
        r.render(s, W, H, spp)
        r.render(s, W, H, spp)  # a second call reuses the stream, events and trace contexts
        imgs.append(r.image()[..., :3].astype(np.float64))
        s.close()
    diff = np.abs(imgs[0] - imgs[1])
    say("max difference %.3e, image mean %.3f" % (diff.max(), imgs[0].mean()))
    assert imgs[0].mean() > 0.01
    assert np.all(diff <= 1e-5 * np.maximum(1.0, np.abs(imgs[0])))


def test_batched_samples_give_the_same_image(ensure_built, monkeypatch):
    """A render whose samples run in several batches (MRT_BATCH_PATHS below one call's paths)
    gives the one-batch image up to fp32 summation order, over several progressive calls with
    odd sample counts."""
    sc = scenes.sphere_field_scene(30, 2, seed=7, glass_every=5, metal_every=3)
    tape = scenes.sdf_cluster(12, seed=4, k=0.1, box=0.8, n_materials=2)
    W, H, depth = 80, 48, 6
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    imgs = []
    for batch in (None, str(W * H * 3)):
        if batch:
            monkeypatch.setenv("MRT_BATCH_PATHS", batch)  # three samples per batch
        s = gpu_scene(sc)
        s.enqueue_sdf(s.push_sdf(tape), scenes.translate_scale(tuple(sc["target"]), 1.0))
        r = m.Renderer()
        r.set_camera(view, proj)
        r.set_num_bounces(depth)
        for spp in (5, 1, 10):
            say("render %dx%d %d spp, batch %s" % (W, H, spp, batch))
            r.render(s, W, H, spp)
        assert r.sample_count() == 16
        imgs.append(r.image()[..., :3].astype(np.float64))
        s.close()
    diff = np.abs(imgs[0] - imgs[1])
    say("max difference %.3e, image mean %.3f" % (diff.max(), imgs[0].mean()))
    assert imgs[0].mean() > 0.01
    assert np.all(diff <= 1e-5 * np.maximum(1.0, np.abs(imgs[0])))


def test_reference_small_config_matches_fp64_oracle(ensure_built):
    """The reference's CPU-runnable configuration: ground quad + 200 icospheres (64K triangles,
    radius U[0.1, 0.5], centres U[-5, 5]^3), GGX roughness U[0.05, 1], albedo U[0.1, 0.9], one
    quad light, 512x512 pinhole camera, 1 spp, depth 4. The oracle draws the same sampler
    sequence, so even one sample per pixel agrees within the 0.5 % bar."""
    sc = scenes.sphere_field_scene(200, 2, seed=1)
    n_tris = sum(sc["geometries"][g][0].size // 3 for g, _, _ in sc["instances"])
    assert n_tris == 200 * 320 + 2
    W, H, spp, depth = 512, 512, 1, 4
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    s = gpu_scene(sc)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    say("render %dx%d %d spp depth %d, 200 icospheres" % (W, H, spp, depth))
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    ref = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0).render(
        view, proj, W, H, spp, depth)
    err = rel_mse(img, ref)
    say("relative MSE %.3e" % err)
    assert np.isfinite(img).all() and img.mean() > 0.01
    assert err <= REL_MSE_TOL


def test_wavefront_scene_matches_fp64_oracle(ensure_built):
    """The 4M-triangle wavefront scene itself (2000 instanced icospheres, the displaced
    1000x750 ground, the 64-primitive SDF cluster; GGX, dielectric and Lambert materials), at the
    bench's camera but 96x54 pixels, 8 spp, depth 8, against the FP64 oracle with the same
    sampler sequence. Runs through the default render path: one batch, material-sorted shading,
    shadow passes on the second stream, the chain-only SDF kernel."""
    sc = scenes.wavefront_scene()
    W, H, spp, depth = 96, 54, 8, 8
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    s = gpu_scene(sc)
    s.enqueue_sdf(s.push_sdf(sc["sdf_tape"]), sc["sdf_transform"])
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(depth)
    say("render %dx%d %d spp depth %d, 4M-triangle scene + SDF" % (W, H, spp, depth))
    r.render(s, W, H, spp)
    img = r.image()[..., :3].astype(np.float64)
    dr = s.sdf_drawables(1)
    say("oracle render")
    ref = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0,
                           sdf_tape=sc["sdf_tape"], sdf_drawables=dr).render(view, proj, W, H, spp, depth)
    err = rel_mse(img, ref)
    say("relative MSE %.3e" % err)
    assert np.isfinite(img).all() and img.mean() > 0.01
    assert err <= REL_MSE_TOL


def test_device_image_and_sample_ranges(ensure_built):
    """mrt_renderer_read_device writes the host read's image into device memory bit for bit, and
    disjoint sample ranges (mrt_renderer_set_sample_base) combine: the mean of samples 0-7 and of
    samples 8-15 equals one 16-sample render up to fp32 summation order."""
    sc = scenes.sphere_field_scene(30, 2, seed=9, glass_every=4, metal_every=3)
    W, H = 64, 40
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], W / H)
    s = gpu_scene(sc)
    r = m.Renderer()
    r.set_camera(view, proj)
    r.set_num_bounces(5)
    say("sample ranges %dx%d" % (W, H))
    r.render(s, W, H, 16)
    full = r.image().astype(np.float64)
    d = m.DeviceBuffer(W * H * 16)
    r.image_device(d)
    m.synchronize()
    assert np.array_equal(d.to_array(np.float32, W * H * 4).reshape(H, W, 4), r.image())
    halves = []
    for base in (0, 8):
        r.set_sample_base(base)
        r.render(s, W, H, 8)
        assert r.sample_count() == 8
        halves.append(r.image().astype(np.float64))
    both = (halves[0] + halves[1]) / 2
    assert np.abs(both - full).max() <= 1e-5 * max(1.0, np.abs(full).max())
    assert np.abs(halves[0] - halves[1]).max() > 1e-3  # different samples
