"""OBJ/MTL ingest: the library's parser (mrt_obj_*) and scene push (mrt_scene_push_obj) against
hand-derived results and against the Python restatement in oracle/obj_ref.py.

Bar: indices, vertex arrays, group names and MTL fields bit-exact; material records exact except
the roughness power, within one float32 ulp. The OBJ fixtures in tests/golden/obj are the
reference's own scene files (res/podium.obj, res/plane.obj with their MTL libraries)."""
import os

import numpy as np
import pytest

import myrt_amd as m
from myrt_amd import obj as mobj
from oracle import obj_ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "obj")
PODIUM = os.path.join(GOLDEN, "podium.obj")
PLANE = os.path.join(GOLDEN, "plane.obj")


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def groups_of(path):
    f = mobj.ObjFile(path)
    try:
        return f.groups
    finally:
        f.close()


def test_quad_closed_form(ensure_built, tmp_path):
    p = write(tmp_path, "quad.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    (g,) = groups_of(p)
    assert g["name"] == "/Default" and g["material_name"] == "Default"
    assert g["indices"].tolist() == [[0, 1, 2], [0, 2, 3]]
    assert g["positions"].tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    # No vn / vt lines: every vertex takes the default normal (0, 1, 0) and uv (0, 0).
    assert g["normals"].tolist() == [[0, 1, 0]] * 4
    assert g["texcoords"].tolist() == [[0, 0]] * 4
    assert g["diffuse"] == (1.0, 0.0, 1.0) and g["specularity"] == 100.0 and g["dissolve"] == 1.0
    assert abs(g["ior"] - 1.45) < 1e-6
    compare_with_oracle(p)


def test_fan_dedup_and_attribute_forms(ensure_built, tmp_path):
    text = """o Fan
v 0 0 0
v 2 0 0
v 3 1 0
v 1 3 0
v -1 1 0
vt 0 0
vt 1 0
vt 0.5 1
vn 0 0 1
vn 0 0 -1
g pent
f 1/1/1 2/2/1 3/3/1 4/1/1 5/2/1
g forms
f 1//1 2//1 3//1
f 1//2 3//2 4//2
f 1/3 2/3 3/3
f 1 2 3
f 1 2
"""
    pent, forms = compare_with_oracle(write(tmp_path, "fan.obj", text))[0]
    assert pent["name"] == "Fan/pent"
    # Fan around the first vertex: (v0, v1, v2), (v0, v2, v3), (v0, v3, v4).
    assert pent["indices"].tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]
    assert pent["texcoords"].tolist() == [[0, 0], [1, 0], [0.5, 1], [0, 0], [1, 0]]
    assert forms["name"] == "Fan/forms"
    # f 1//1 ... : uv index absent = uv 0; the second face shares position 1 and 3 but with
    # normal 2, so it adds three new vertices; "1/3" takes uv 3 and normal 0; "1" takes both 0.
    # The two-vertex face is dropped.
    assert forms["indices"].tolist() == [[0, 1, 2], [3, 4, 5], [6, 7, 8], [0, 1, 2]]
    assert forms["normals"][:3].tolist() == [[0, 0, 1]] * 3
    assert forms["normals"][3:6].tolist() == [[0, 0, -1]] * 3
    assert forms["positions"][3:6].tolist() == [[0, 0, 0], [3, 1, 0], [1, 3, 0]]
    assert forms["texcoords"][6:9].tolist() == [[0.5, 1]] * 3


def test_global_and_negative_indices(ensure_built, tmp_path):
    # Positive indices count over the whole file; negative ones from the end of the object's list.
    text = """o A
v 0 0 0
v 1 0 0
v 0 1 0
f 1 2 3
o B
v 5 0 0
v 6 0 0
v 5 1 0
v 7 7 7
f 4 5 6
f -1 -2 -4
"""
    a, b = compare_with_oracle(write(tmp_path, "multi.obj", text))[0]
    assert a["name"] == "A/Default" and b["name"] == "B/Default"
    assert b["positions"].tolist() == [[5, 0, 0], [6, 0, 0], [5, 1, 0], [7, 7, 7]]
    assert b["indices"].tolist() == [[0, 1, 2], [3, 2, 0]]


def test_mtl_fields_groups_and_material_mapping(ensure_built, tmp_path):
    write(tmp_path, "a.mtl", """newmtl Red
Ns 225.000000
Kd 0.092716 0.800000 0.5
Ni 1.330000
d 0.25
map_Kd tex/red.png
newmtl Plain
Kd 0.2 0.3 0.4
""")
    write(tmp_path, "b.mtl", "newmtl Red\nKd 1 1 1\nnewmtl Blue\nKd 0 0 1\nNs 1000000\n")
    text = """mtllib a.mtl b.mtl
o Thing
v 0 0 0
v 1 0 0
v 0 1 0
usemtl Unknown
usemtl Red
f 1 2 3
usemtl Blue
f 1 3 2
usemtl Plain
"""
    groups, ref = compare_with_oracle(write(tmp_path, "m.obj", text))
    # "usemtl Red" after an empty group reuses it (keeping the name "Unknown"); the trailing
    # "usemtl Plain" opens a group without faces.
    assert [g["name"] for g in groups] == ["Thing/Unknown", "Thing/Blue", "Thing/Plain"]
    red, blue, plain = groups
    # A name defined by two libraries keeps the first library's definition.
    assert red["material_name"] == "Red" and red["diffuse"] == tuple(np.float32([0.092716, 0.8, 0.5]))
    assert red["specularity"] == 225.0 and np.float32(red["ior"]) == np.float32(1.33)
    assert red["dissolve"] == 0.25 and red["albedo_texture"] == str(tmp_path / "tex" / "red.png")
    assert blue["diffuse"] == (0.0, 0.0, 1.0) and blue["albedo_texture"] == ""
    assert plain["indices"].size == 0

    s = m.Scene()
    f = mobj.ObjFile(str(tmp_path / "m.obj"))
    geoms = np.zeros(3, np.uint32)
    mats = np.zeros(3, np.uint32)
    m._check(m.lib().mrt_scene_push_obj(s.h, f.h, m._ptr(geoms), m._ptr(mats), 3))
    f.close()
    assert geoms[2] == mobj.NO_GROUP and mats[2] == mobj.NO_GROUP
    r = s.get_material(int(mats[0]))[0]
    # albedo = unorm8 by truncation: 255 * 0.092716 = 23.6 -> 23, 255 * 0.8 = 204, 255 * 0.5 = 127.
    assert r["albedo_rgba8"] == 23 | 204 << 8 | 127 << 16 | 255 << 24
    # roughness = (1 / 225)^(1 / pi) = exp(-ln 225 / pi) = 0.17824...
    assert abs(r["roughness"] - np.exp(-np.log(225.0) / np.pi)) < 1e-6
    assert r["transmission"] == 0.75 and r["metallic"] == 0.0 and r["texture"] == 0
    assert np.float32(r["ior"]) == np.float32(1.33)
    b = s.get_material(int(mats[1]))[0]
    assert b["roughness"] == np.float32(0.054)  # (1e-6)^(1/pi) = 0.0123 is below the floor
    assert b["albedo_rgba8"] == 255 << 16 | 255 << 24
    # The oracle's mapping gives the same records.
    for rec, g in ((r, ref[0]), (b, ref[1])):
        rgba, ior, rough, trans = obj_ref.material_record(g["material"])
        assert (rec["albedo_rgba8"], rec["transmission"]) == (rgba, trans) and np.float32(rec["ior"]) == np.float32(ior)
        assert abs(rec["roughness"] - rough) < 1e-7
    s.close()


def test_errors(ensure_built, tmp_path):
    with pytest.raises(m.MrtError, match="cannot open"):
        mobj.ObjFile(str(tmp_path / "missing.obj"))
    with pytest.raises(m.MrtError, match="out of range"):
        mobj.ObjFile(write(tmp_path, "bad.obj", "v 0 0 0\nv 1 0 0\nf 1 2 3\n"))
    with pytest.raises(m.MrtError, match="out of range"):
        mobj.ObjFile(write(tmp_path, "bad2.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nf 1/2 2 3\n"))


def compare_with_oracle(path):
    got = groups_of(path)
    ref = obj_ref.load_groups(path)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert g["name"] == r["name"]
        assert g["material_name"] == r["material"]["name"]
        assert np.array_equal(g["indices"], r["indices"])
        for key in ("positions", "normals", "texcoords"):
            assert g[key].view(np.uint32).tolist() == r[key].view(np.uint32).tolist(), key
        assert g["diffuse"] == tuple(np.float32(r["material"]["diffuse"]))
        assert np.float32(g["specularity"]) == np.float32(r["material"]["specularity"])
        assert np.float32(g["dissolve"]) == np.float32(r["material"]["dissolve"])
    return got, ref


def test_reference_scene_files_match_oracle(ensure_built):
    got, _ = compare_with_oracle(PODIUM)
    # podium.obj: 4 objects of one material each; 32 triangles and 481 quads.
    assert len(got) == 4
    assert sum(len(g["indices"]) for g in got) == 32 + 2 * 481
    # One group per object, named by its "o" line and its single "usemtl" line.
    lines = open(PODIUM).read().split("\n")
    names = [l.split()[1] for l in lines if l.startswith("o ")]
    used = [l.split()[1] for l in lines if l.startswith("usemtl ")]
    assert [g["name"] for g in got] == ["%s/%s" % nu for nu in zip(names, used)]
    got, _ = compare_with_oracle(PLANE)
    assert len(got) == 5
    assert sum(len(g["indices"]) for g in got) == 192 + 2 * 1369
    # Every triangle's corners are positions listed in the file.
    listed = {tuple(np.float32(l.split()[1:4]).tolist()) for l in open(PLANE) if l.startswith("v ")}
    for g in got:
        assert {tuple(p) for p in g["positions"].tolist()} <= listed


def random_obj(rng, tmp_path, k):
    """A random OBJ: several objects and groups, polygons of 3-7 vertices, every index form,
    repeated attribute values (so that vertex sharing matters) and a material library."""
    mats = ["M%d" % i for i in range(4)]
    mtl = "".join("newmtl %s\nNs %g\nKd %g %g %g\nd %g\nNi %g\n" % (
        n, rng.uniform(1, 900), *rng.uniform(0, 1, 3), rng.uniform(0, 1), rng.uniform(1, 2)) for n in mats)
    write(tmp_path, "r%d.mtl" % k, mtl)
    lines = ["mtllib r%d.mtl" % k]
    off = [0, 0, 0]  # sizes of the earlier objects' position, uv and normal lists
    for o in range(int(rng.integers(1, 4))):
        lines.append("o obj%d" % o)
        nv, nt, nn = (int(x) for x in rng.integers(3, 12, 3))
        grid = np.round(rng.uniform(-2, 2, (nv, 3)) * 4) / 4
        lines += ["v %s" % " ".join(repr(float(x)) for x in p) for p in grid]
        lines += ["vt %s %s" % tuple(repr(float(x)) for x in np.round(rng.uniform(0, 1, 2), 2)) for _ in range(nt)]
        lines += ["vn %s %s %s" % tuple(repr(float(x)) for x in rng.choice([-1.0, 0.0, 1.0], 3)) for _ in range(nn)]
        for g in range(int(rng.integers(1, 4))):
            lines.append(("g grp%d" % g) if rng.random() < 0.5 else ("usemtl " + str(rng.choice(mats + ["none"]))))
            form = int(rng.integers(4))
            for _ in range(int(rng.integers(1, 6))):
                verts = []
                for _ in range(int(rng.integers(3, 8))):
                    p, t, n = int(rng.integers(nv)), int(rng.integers(nt)), int(rng.integers(nn))
                    neg = rng.random() < 0.3
                    P = (p - nv) if neg else None
                    T = (t - nt) if neg else None
                    N = (n - nn) if neg else None
                    verts.append((P, T, N, p, t, n))
                toks = []
                for P, T, N, p, t, n in verts:
                    # positive indices are global: add the sizes of the earlier objects' lists
                    ip = P if P is not None else p + 1 + off[0]
                    it = T if T is not None else t + 1 + off[1]
                    inn = N if N is not None else n + 1 + off[2]
                    toks.append([str(ip), "%d/%d" % (ip, it), "%d//%d" % (ip, inn), "%d/%d/%d" % (ip, it, inn)][form])
                lines.append("f " + " ".join(toks))
        off[0] += nv
        off[1] += nt
        off[2] += nn
    return write(tmp_path, "r%d.obj" % k, "\n".join(lines) + "\n")


def test_random_files_match_oracle(ensure_built, tmp_path):
    rng = np.random.default_rng(7)
    for k in range(20):
        compare_with_oracle(random_obj(rng, tmp_path, k))


def test_load_object_file_pushes_geometry_and_materials(ensure_built):
    s = m.Scene()
    objs = mobj.load_object_file(s, PODIUM, import_scale=6.0)
    ref = obj_ref.load_groups(PODIUM)
    assert [o.name for o in objs] == [r["name"] for r in ref]
    for o, r in zip(objs, ref):
        got = s.get_material(o.material)[0]
        rgba, ior, rough, trans = obj_ref.material_record(r["material"])
        assert got["albedo_rgba8"] == rgba and got["ior"] == np.float32(ior) and got["transmission"] == trans
        assert abs(int(np.float32(got["roughness"]).view(np.uint32)) - int(np.float32(rough).view(np.uint32))) <= 1
        assert np.array_equal(o.transformation, np.diag([6.0, 6.0, 6.0, 1.0]).astype(np.float32))
    mobj.enqueue_objects(s, objs)
    res = s.build_host()
    assert res["num_geometries"] == 4 and res["num_instances"] == 4
    # split_aabbs may reference a triangle from several leaves, never from none.
    assert res["blas_tri_records"] >= 32 + 2 * 481
    s.close()


def test_map_kd_texture_is_loaded(ensure_built, tmp_path):
    PIL = pytest.importorskip("PIL.Image")
    img = np.zeros((2, 3, 3), np.uint8)
    img[0, 0] = (255, 0, 0)
    img[1, 2] = (0, 0, 255)
    (tmp_path / "tex").mkdir()
    PIL.fromarray(img).save(str(tmp_path / "tex" / "t.png"))
    write(tmp_path, "t.mtl", "newmtl T\nKd 1 1 1\nmap_Kd tex/t.png\n")
    p = write(tmp_path, "t.obj", "mtllib t.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nusemtl T\nf 1/1 2/1 3/1\n")
    rgba = mobj.load_ldr_texture(str(tmp_path / "tex" / "t.png"))
    assert rgba.shape == (2, 3, 4) and rgba[0, 0].tolist() == [1, 0, 0, 1] and rgba[1, 2].tolist() == [0, 0, 1, 1]
    s = m.Scene()
    (o,) = mobj.load_object_file(s, p, import_scale=1.0)
    assert s.get_material(o.material)[0]["texture"] == 1
    s.close()
