"""The FP64 path-tracing oracle, pinned by closed forms (no GPU):

* direct lighting of a Lambert plane by a quad light equals rho/pi times the irradiance
  integral E = int Le cos(theta) cos(theta') / r^2 dA, evaluated here by a 400x400 midpoint
  rule (this exercises light sampling, the MIS split between light and BSDF samples, the Lambert
  lobe and the camera);
* a white-furnace check: a Lambert sphere under a uniform-looking environment reflects at most
  what it receives and a black sphere reflects nothing;
* determinism: the same sampler sequence gives the same image.
"""
import numpy as np

import oracle as orc
from myrt_amd import scenes


def lambert_plane(rho=0.5, light_h=2.0, emission=10.0):
    gi, gp, gn = scenes.ground_quad(size=20.0, y=0.0)
    mat = np.zeros(1, orc.MATERIAL_DTYPE)
    mat[0] = (scenes._rgba8((rho, rho, rho)), 1.0, 1.0, 0.0, 0.0, 0, (0, 0))
    light = np.zeros(1, orc.LIGHT_DTYPE)
    light[0] = ((-0.5, light_h, -0.5), (1, 0, 0), (0, 0, 1), (emission,) * 3)
    return orc.OracleRender([(gi, gp, gn)], [(0, np.eye(4, dtype=np.float32), 0)], mat, light, 0.0)


def irradiance_below_square(h, emission, n=400):
    xs = (np.arange(n) + 0.5) / n - 0.5
    X, Z = np.meshgrid(xs, xs)
    r2 = X ** 2 + Z ** 2 + h * h
    return float((emission * (h / np.sqrt(r2)) ** 2 / r2).sum() / n ** 2)


def test_direct_light_closed_form(ensure_built):
    for h in (1.0, 2.0, 4.0):
        R = lambert_plane(light_h=h)
        view, proj = scenes.camera_matrices((0, h + 3, 0.001), (0, 0, 0), 1.0, 1.0)
        img = R.render(view, proj, 3, 3, 8192, 2)
        expect = (128 / 255) / np.pi * irradiance_below_square(h, 10.0)
        assert abs(img[1, 1, 0] - expect) / expect < 0.01, (h, img[1, 1, 0], expect)


def test_light_seen_directly(ensure_built):
    """A camera ray hitting the emitting side of the light returns its radiance exactly."""
    R = lambert_plane(light_h=2.0, emission=7.0)
    view, proj = scenes.camera_matrices((0.0, 0.5, 0.0), (0.0, 3.0, 0.001), 1.0, 1.0)
    img = R.render(view, proj, 3, 3, 4, 2)
    assert np.allclose(img[1, 1], 7.0)


def test_furnace_sphere(ensure_built):
    (fi, fp, fn) = scenes.icosphere(3)
    white = np.zeros(2, orc.MATERIAL_DTYPE)
    white[0] = (scenes._rgba8((1.0, 1.0, 1.0)), 1.0, 1.0, 0.0, 0.0, 0, (0, 0))
    # A black metal: Fresnel weight 1 and a black specular tint, so nothing is reflected.
    white[1] = (scenes._rgba8((0.0, 0.0, 0.0)), 1.0, 1.0, 1.0, 0.0, 0, (0, 0))
    view, proj = scenes.camera_matrices((0, 0, -4), (0, 0, 0), 20.0, 1.0)
    for mi, bound in ((0, 1.0), (1, 0.0)):
        R = orc.OracleRender([(fi, fp, fn)], [(0, np.eye(4, dtype=np.float32), mi)], white, None, 1.0)
        img = R.render(view, proj, 8, 8, 256, 6)
        env_max = 1.0  # the sky gradient never exceeds 1 with scale 1
        centre = img[3:5, 3:5].mean()
        if bound == 0.0:
            assert centre == 0.0
        else:
            assert 0.2 < centre <= env_max * 1.001


def test_deterministic(ensure_built):
    sc = scenes.sphere_field_scene(20, 1, seed=3, glass_every=5, metal_every=4)
    R = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0)
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.5)
    a = R.render(view, proj, 24, 16, 8, 4, threads=1)
    b = R.render(view, proj, 24, 16, 8, 4, threads=4)
    assert np.array_equal(a, b)
    c = R.render(view, proj, 24, 16, 8, 4, sample_base=8)
    assert not np.array_equal(a, c)


def small_dof_scene():
    sc = scenes.sphere_field_scene(12, 1, seed=9)
    r = orc.OracleRender(sc["geometries"], sc["instances"], sc["materials"], sc["lights"], 1.0)
    view, proj = scenes.camera_matrices(sc["eye"], sc["target"], sc["fovy"], 1.0)
    return r, view, proj


def test_dof_closed_forms(ensure_built):
    """Thin lens of size 0 is the pinhole camera; an all-black mask yields black; a constant
    mask c scales the lens image by c (the first lens sample is accepted, so paths coincide)."""
    r, view, proj = small_dof_scene()
    pin = r.render(view, proj, 12, 10, 4, 3)
    r.set_dof(True, (0.0, 0.0), 9.0)
    zero_lens = r.render(view, proj, 12, 10, 4, 3)
    assert np.allclose(zero_lens, pin, rtol=1e-6, atol=1e-9)
    r.set_dof(True, (1.5, 1.5), 9.0)
    lens = r.render(view, proj, 12, 10, 4, 3)
    assert not np.allclose(lens, pin, rtol=1e-3)
    r.set_dof(True, (1.5, 1.5), 9.0, bokeh_mask=np.full((16, 16, 3), 0.5))
    half = r.render(view, proj, 12, 10, 4, 3)
    assert np.allclose(half, 0.5 * lens, rtol=1e-9, atol=1e-12)
    r.set_dof(True, (1.5, 1.5), 9.0, bokeh_mask=np.zeros((16, 16, 3)))
    assert not r.render(view, proj, 12, 10, 4, 3).any()


def test_texture_and_cubemap_closed_forms(ensure_built):
    """A white albedo texture equals a white material; a texture of value c gives albedo c^2.2; a
    camera ray that leaves the scene along +x reads face 0 of the cube map at the texel row the
    OpenGL face table gives."""
    gi, gp, gn, guv = scenes.textured_ground(size=3.0, y=0.0, tiles=2.0)
    geo = [(gi, gp, gn, guv)]
    inst = [(0, np.eye(4, dtype=np.float32), 0)]
    lights = scenes.LIGHT_DTYPE and np.zeros(1, scenes.LIGHT_DTYPE)
    lights[0] = ((-0.5, 2.0, -0.5), (1, 0, 0), (0, 0, 1), (10, 10, 10))
    view, proj = scenes.camera_matrices((0, 4, 0.01), (0, 0, 0), 40.0, 1.0)

    def mat(rgb, tex):
        m = np.zeros(1, scenes.MATERIAL_DTYPE)
        m[0] = (scenes._rgba8(rgb), 1.0, 1.0, 0.0, 0.0, tex, (0, 0))
        return m

    white = orc.OracleRender(geo, inst, mat((1, 1, 1), 0), lights, 0.5).render(view, proj, 8, 8, 4, 3)
    tex_white = orc.OracleRender(geo, inst, mat((0.2, 0.3, 0.4), 1), lights, 0.5,
                                 textures=[np.ones((4, 4, 3))]).render(view, proj, 8, 8, 4, 3)
    assert np.allclose(tex_white, white, rtol=1e-12, atol=1e-15)
    k = np.array([51, 102, 204])
    c = ((k / 255.0) ** (1 / 2.2)).astype(np.float32)
    plain = orc.OracleRender(geo, inst, mat(k / 255.0, 0), lights, 0.5).render(view, proj, 8, 8, 4, 3)
    tex_c = orc.OracleRender(geo, inst, mat((1, 1, 1), 1), lights, 0.5,
                             textures=[np.broadcast_to(c, (4, 4, 3))]).render(view, proj, 8, 8, 4, 3)
    assert np.allclose(tex_c, plain, rtol=1e-5)
    # Cube map: one pixel looking along +x, above the ground plane's extent.
    far = [(gi, gp - np.array([0, 100, 0], np.float32), gn, guv)]
    o = orc.OracleRender(far, inst, mat((1, 1, 1), 0), None, 2.0)
    cube = scenes.gradient_cubemap(32)
    o.set_cubemap(cube)
    v2, p2 = scenes.camera_matrices((0, 0, 0), (1, 0.3, 0), 1e-3, 1.0)
    got = o.render(v2, p2, 1, 1, 2, 1)[0, 0]
    d = np.array([1.0, 0.3, 0.0]) / np.linalg.norm([1.0, 0.3, 0.0])
    t = 0.5 * (-d[1] / d[0] + 1) * 32 - 0.5  # face +x: tc = -y / |x|
    row = (np.arange(32) + 0.5) / 32
    expect = 2.0 * np.array([1.0, 0.3, 0.2]) * (0.5 + np.interp(t, np.arange(32), row))
    assert np.allclose(got, expect, rtol=1e-3)


def test_nonfinite_rules(ensure_built):
    """The oracle's rejection rules: NaN vertex normals fall back to the geometric normal (the
    image equals the one rendered without vertex normals, where every shading normal is the
    geometric one), and each fallback is counted; a scene without non-finite input counts none."""
    gi, gp, gn = scenes.ground_quad(size=20.0, y=0.0)
    mat = np.zeros(1, orc.MATERIAL_DTYPE)
    mat[0] = (scenes._rgba8((0.5, 0.5, 0.5)), 1.5, 0.4, 0.0, 0.0, 0, (0, 0))
    light = np.zeros(1, orc.LIGHT_DTYPE)
    light[0] = ((-0.5, 2.0, -0.5), (1, 0, 0), (0, 0, 1), (10.0,) * 3)
    view, proj = scenes.camera_matrices((0, 5, 0.001), (0, 0, 0), 1.0, 1.0)
    inst = [(0, np.eye(4, dtype=np.float32), 0)]
    clean = orc.OracleRender([(gi, gp, None)], inst, mat, light, 0.5)
    a = clean.render(view, proj, 8, 8, 16, 3)
    assert clean.last_nonfinite == 0
    nan_n = np.full_like(np.asarray(gn, np.float32), np.nan)
    poisoned = orc.OracleRender([(gi, gp, nan_n)], inst, mat, light, 0.5)
    b = poisoned.render(view, proj, 8, 8, 16, 3)
    assert np.isfinite(b).all()
    assert poisoned.last_nonfinite >= 8 * 8 * 16  # every primary hit falls back
    assert np.array_equal(a, b)
