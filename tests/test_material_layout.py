"""Named-parameter material layouts (src/types.hpp:216-310, src/material.hpp:10-118,
src/pathtracer/scene.hpp:131-140, 156, 249-252), host side.

The expected block arrays below are written out by hand from the reference's rules: parameters
packed as consecutive 4-byte blocks in declaration order (object_type::add_parameter,
types.hpp:250-258), nested layouts inline, a vec4unorm8 as one block holding its four bytes
(types.hpp:66-68, parameter.cpp:130-133), a material's blocks appended to one buffer and zeroed
(material.hpp:81-91), set_parameter writing the value's bytes at the parameter's offset and
ignoring names the layout lacks (material.hpp:93-102). The application's own calls
(main.cpp:447-464, 526-533) are replayed with the same C++ value types.
"""
import struct

import numpy as np
import pytest

import myrt_amd as m


def f32(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


@pytest.fixture
def scene(ensure_built):
    s = m.Scene()
    yield s
    s.close()


def app_types(s):
    # color_object_type and material_object_type of main.cpp:447-464.
    color = s.register_type("color_t", [("value", "vec4unorm8")])
    mat = s.register_type("material", [("albedo", "color_t"), ("ior", "float"), ("roughness", "float"),
                                       ("metallic", "float"), ("transmission", "float"),
                                       ("albedo_texture", "uvec2")])
    return color, mat


def test_application_layout(scene):
    color, mat = app_types(scene)
    assert scene.layout(color) == (1, [("value", 0, 1, 6, None)])
    assert scene.layout(mat) == (7, [("albedo", 0, 1, -1, color), ("ior", 1, 1, 0, None),
                                     ("roughness", 2, 1, 0, None), ("metallic", 3, 1, 0, None),
                                     ("transmission", 4, 1, 0, None), ("albedo_texture", 5, 2, 10, None)])


def test_set_parameter_every_name_as_the_application_does(scene):
    _, mat = app_types(scene)
    tex = scene.push_texture(np.ones((2, 2, 3), np.float32))
    mt = scene.push_material_type(mat)
    a = scene.push_material(mt)
    assert scene.material_slot(a) == (mt, 0)
    assert list(scene.material_blocks()) == [0] * 7  # material.hpp:85-86: zeroed
    # main.cpp:527-533: rnu::vec4ui8(255 * vec4(0.1, 0.4, 0, 1)) truncates to (25, 102, 0, 255).
    assert scene.set_parameter(a, "albedo", np.array([25, 102, 0, 255], np.uint8))
    assert scene.set_parameter(a, "ior", np.float32(1.4))
    assert scene.set_parameter(a, "roughness", np.float32(0.8))
    assert scene.set_parameter(a, "transmission", np.float32(0.0))
    assert scene.set_parameter(a, "metallic", np.float32(0.25))
    assert scene.set_parameter(a, "albedo_texture", np.array([tex, 0], np.uint32))
    assert list(scene.material_blocks()) == [0xFF006619, f32(1.4), f32(0.8), f32(0.25), 0, tex, 0]
    r = scene.get_material(a)[0]
    assert int(r["albedo_rgba8"]) == 0xFF006619 and int(r["texture"]) == tex
    assert (np.float32(r["ior"]), np.float32(r["roughness"]), np.float32(r["metallic"])) == (
        np.float32(1.4), np.float32(0.8), np.float32(0.25))
    # A second material starts after the first one's 7 blocks.
    b = scene.push_material(mt)
    assert scene.material_slot(b) == (mt, 7)
    assert scene.set_parameter(b, "roughness", np.float32(0.5))
    assert list(scene.material_blocks()[7:]) == [0, 0, f32(0.5), 0, 0, 0, 0]
    # A material with all-zero blocks resolves to zeros: black albedo, ior 0 (material.hpp:86).
    r = scene.get_material(b)[0]
    assert int(r["albedo_rgba8"]) == 0 and float(r["ior"]) == 0.0 and float(r["roughness"]) == 0.5


def test_set_parameter_sizes_and_unknown_names(scene):
    _, mat = app_types(scene)
    a = scene.push_material(scene.push_material_type(mat))
    before = scene.material_blocks().copy()
    # The reference asserts sizeof(value) / sizeof(float) == blocks (material.hpp:97).
    with pytest.raises(m.MrtError, match="Data size not suitable"):
        scene.set_parameter(a, "ior", np.float64(1.4))
    with pytest.raises(m.MrtError, match="Data size not suitable"):
        scene.set_parameter(a, "albedo", np.float32([0.1, 0.4, 0.0, 1.0]))
    with pytest.raises(m.MrtError, match="Data size not suitable"):
        scene.set_parameter(a, "albedo_texture", np.uint32(3))
    with pytest.raises(m.MrtError, match="Data size not suitable"):
        scene.set_parameter(a, "albedo", b"\x01\x02\x03")
    # A name the layout lacks is ignored (material.hpp:96 `if (iter != end(ps))`).
    assert scene.set_parameter(a, "emission", np.float32(3.0)) is False
    assert np.array_equal(scene.material_blocks(), before)
    # Any 4-byte value fits a one-block parameter; its bytes land unchanged.
    assert scene.set_parameter(a, "albedo", np.int32(-2))
    assert scene.material_blocks()[0] == 0xFFFFFFFE


def test_nested_layout_offsets_and_bindings(scene):
    color = scene.register_type("color_t", [("value", "vec4unorm8")])
    fancy = scene.register_type("fancy", [("tint", "vec3"), ("frame", "mat3"), ("rough", "float"),
                                          ("albedo", "color_t"), ("ior", "float"), ("tex", "uint"),
                                          ("extra", "ivec2")])
    assert scene.layout(fancy) == (18, [("tint", 0, 3, 4, None), ("frame", 3, 9, 14, None), ("rough", 12, 1, 0, None),
                                        ("albedo", 13, 1, -1, color), ("ior", 14, 1, 0, None),
                                        ("tex", 15, 1, 2, None), ("extra", 16, 2, 7, None)])
    mt = scene.push_material_type(fancy, {"roughness": "rough", "albedo_texture": "tex"})
    x = scene.push_material(mt)
    scene.set_parameter(x, "tint", [1.0, 2.0, 3.0])
    scene.set_parameter(x, "rough", 0.3)
    scene.set_parameter(x, "albedo", [1.0, 0.5, 0.0])           # float form: round(255 x), alpha 1
    scene.set_parameter(x, "ior", np.float32(1.33))
    scene.set_parameter(x, "extra", [-1, 7])
    scene.set_parameter(x, "frame", np.arange(9, dtype=np.float32))
    want = [f32(1.0), f32(2.0), f32(3.0)] + [f32(float(i)) for i in range(9)] + [
        f32(0.3), 0xFF0080FF, f32(1.33), 0, 0xFFFFFFFF, 7]
    assert list(scene.material_blocks()) == want
    r = scene.get_material(x)[0]
    assert int(r["albedo_rgba8"]) == 0xFF0080FF and np.float32(r["roughness"]) == np.float32(0.3)
    assert np.float32(r["ior"]) == np.float32(1.33)
    # Unbound inputs keep their defaults: metallic and transmission 0.
    assert float(r["metallic"]) == 0.0 and float(r["transmission"]) == 0.0 and int(r["texture"]) == 0
    # Explicit path through a layout, and an explicitly unbound input.
    mt2 = scene.push_material_type(fancy, {"albedo": "albedo.value", "ior": "", "roughness": "rough"})
    y = scene.push_material(mt2)
    assert scene.material_slot(y) == (mt2, 18)
    scene.set_parameter(y, "ior", np.float32(2.0))
    assert float(scene.get_material(y)[0]["ior"]) == 1.5  # unbound: the default


def test_binding_type_errors(scene):
    scene.register_type("color_t", [("value", "vec4unorm8")])
    pair = scene.register_type("pair", [("a", "float"), ("b", "float")])
    lay = scene.register_type("l", [("tint", "vec3"), ("rough", "float"), ("alb", "vec4"), ("p", "pair"),
                                    ("handle", "uvec3"), ("ior", "vec2")])
    with pytest.raises(m.MrtError, match="roughness input cannot take a parameter of type vec3"):
        scene.push_material_type(lay, {"roughness": "tint", "ior": ""})
    with pytest.raises(m.MrtError, match="albedo input cannot take a parameter of type vec4"):
        scene.push_material_type(lay, {"albedo": "alb", "ior": ""})
    with pytest.raises(m.MrtError, match="albedo_texture input cannot take"):
        scene.push_material_type(lay, {"albedo_texture": "handle", "ior": ""})
    with pytest.raises(m.MrtError, match="names a structured parameter"):
        scene.push_material_type(lay, {"metallic": "p", "ior": ""})
    with pytest.raises(m.MrtError, match="has no parameter 'nope'"):
        scene.push_material_type(lay, {"metallic": "nope", "ior": ""})
    with pytest.raises(m.MrtError, match="has no members"):
        scene.push_material_type(lay, {"metallic": "rough.x", "ior": ""})
    # Binding by name finds "ior", a vec2: an error unless the input is unbound or rebound.
    with pytest.raises(m.MrtError, match="ior input cannot take a parameter of type vec2"):
        scene.push_material_type(lay)
    mt = scene.push_material_type(lay, {"ior": "p.b", "metallic": "p.a", "roughness": "rough"})
    x = scene.push_material(mt)
    scene.set_parameter(x, "p", np.float32([0.75, 1.25]))
    r = scene.get_material(x)[0]
    assert float(r["metallic"]) == 0.75 and float(r["ior"]) == 1.25
    with pytest.raises(m.MrtError, match="structured"):
        scene.set_parameter(x, "p", [0.5, 0.5])


def test_layouts_freeze_once_used_and_lookup(scene):
    # Before any registration, "color_t" is no type (the built-in record layout is not visible).
    with pytest.raises(m.MrtError, match="unknown type"):
        scene.find_type("color_t")
    assert scene.find_type("float") == ("constant", 0)
    assert scene.find_type("vec4unorm8") == ("constant", 6)
    assert scene.find_type("uvec2") == ("constant", 10)
    assert scene.find_type("mat4") == ("constant", 15)
    color = scene.register_type("color_t", [("value", "vec4unorm8")])
    assert scene.find_type("color_t") == ("layout", color)
    # types_registry::find looks at layouts first: a layout named "float" shadows the constant.
    shadow = scene.register_type("float", [("x", "vec2")])
    assert scene.find_type("float") == ("layout", shadow)
    with pytest.raises(m.MrtError, match="unknown parameter type"):
        scene.register_type("bad", [("x", "vec5")])
    # color_t is in use by a layout, so it cannot grow; a material type freezes its layout too.
    outer = scene.register_type("outer", [("c", "color_t")])
    with pytest.raises(m.MrtError, match="in use"):
        m._check(m.lib().mrt_scene_layout_add_parameter(scene.h, color, b"more", b"float"))
    scene.push_material_type(outer, {"ior": ""})
    with pytest.raises(m.MrtError, match="in use"):
        m._check(m.lib().mrt_scene_layout_add_parameter(scene.h, outer, b"more", b"float"))


def test_float_form_checks(scene):
    lay = scene.register_type("t", [("tex", "uvec2"), ("n", "int"), ("c", "vec4unorm8"), ("ior", "float")])
    mt = scene.push_material_type(lay, {"albedo_texture": "tex", "albedo": "c"})
    x = scene.push_material(mt)
    with pytest.raises(m.MrtError, match="not a texture id"):
        scene.set_parameter(x, "tex", 1.0)  # no textures pushed
    t = scene.push_texture(np.ones((1, 1, 3), np.float32))
    scene.set_parameter(x, "tex", float(t))  # a texture input takes the id alone
    assert list(scene.material_blocks()[:2]) == [t, 0]
    with pytest.raises(m.MrtError, match="integral"):
        scene.set_parameter(x, "n", 2.5)
    with pytest.raises(m.MrtError, match="takes 1 values"):
        scene.set_parameter(x, "ior", [1.0, 2.0])
    with pytest.raises(m.MrtError, match="3 or 4 values"):
        scene.set_parameter(x, "c", [1.0, 0.0])
    with pytest.raises(m.MrtError, match="unknown material parameter"):
        scene.set_parameter(x, "nope", 1.0)
    # A uvec2 handle with a nonzero high word names no texture of this scene.
    scene.set_parameter(x, "tex", np.array([t, 1], np.uint32))
    assert int(scene.get_material(x)[0]["texture"]) == 0


def test_record_entry_points_on_a_generic_type(scene):
    lay = scene.register_type("t", [("roughness", "float"), ("albedo", "vec4unorm8")])
    mt = scene.push_material_type(lay, {"ior": ""})
    x = scene.push_material(mt)
    rec = m.material(albedo=(1.0, 0.0, 0.0), roughness=0.4)
    rec["ior"] = 1.5
    m._check(m.lib().mrt_scene_set_material(scene.h, x, m._ptr(rec)))
    assert list(scene.material_blocks()) == [f32(0.4), 0xFF0000FF]
    rec["metallic"] = 0.5  # no parameter feeds metallic
    with pytest.raises(m.MrtError, match="no parameter for the metallic input"):
        m._check(m.lib().mrt_scene_set_material(scene.h, x, m._ptr(rec)))
    # Records pushed with mrt_scene_push_material use the built-in 7-block layout.
    y = scene.push_material(m.material(albedo=(0.0, 1.0, 0.0), ior=1.4, roughness=0.2, metallic=0.1,
                                       transmission=0.3))
    t, off = scene.material_slot(y)
    assert off == 2
    assert list(scene.material_blocks()[2:]) == [0xFF00FF00, f32(1.4), f32(0.2), f32(0.1), f32(0.3), 0, 0]
