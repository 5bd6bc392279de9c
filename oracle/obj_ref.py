"""Pure-Python restatement of the OBJ/MTL ingest (test infrastructure only).

load_mtllib follows src/obj.cpp:9-80, load_obj src/obj.cpp:96-271, triangulate
src/obj.cpp:297-361 and material_record the MTL -> material mapping of src/main.cpp:919-927.
Where the reference's behaviour is undefined (a face or `usemtl` before any object, `usemtl` of a
name no library defines, a face vertex without all three fields, an index outside its list) this
restatement takes the defined choice the library documents: the object is created, the default
material is used, an absent field is 0, an out-of-range index raises IndexError.
Numbers are parsed to float32 the way an istream does (decimal -> nearest float32); a token that is
not a number leaves 0 and ends the statement, as a failed extraction does.
"""
import math
import os
import struct

import numpy as np


def _default_material():
    # make_default_material (src/obj.cpp:84-95)
    return {"name": "Default", "specularity": 100.0, "diffuse": (1.0, 0.0, 1.0), "ior": 1.45, "dissolve": 1.0,
            "map_diffuse": ""}


def _f32(tok):
    return float(np.float32(float(tok)))


def _floats(toks, n):
    out = [0.0] * n
    for k in range(min(n, len(toks))):
        try:
            out[k] = _f32(toks[k])
        except ValueError:
            break
    return out


def load_mtllib(path):
    lib = {}
    cur = None
    if not os.path.isfile(path):
        return lib
    with open(path, "rb") as f:
        for raw in f.read().decode("latin-1").split("\n"):
            toks = raw.split()
            if not toks:
                continue
            ident, rest = toks[0], toks[1:]
            if ident == "newmtl":
                cur = _default_material()
                cur["name"] = rest[0] if rest else ""
                lib[cur["name"]] = cur
            elif cur is None:
                continue
            elif ident == "Ns":
                cur["specularity"] = _floats(rest, 1)[0]
            elif ident == "Ni":
                cur["ior"] = _floats(rest, 1)[0]
            elif ident == "d":
                cur["dissolve"] = _floats(rest, 1)[0]
            elif ident == "Kd":
                cur["diffuse"] = tuple(_floats(rest, 3))
            elif ident == "map_Kd":
                cur["map_diffuse"] = rest[0] if rest else ""
    return lib


def _strtol(s, i):
    """strtol(s + i, &end, 10) over ASCII: returns (value, end index)."""
    j = i
    if j < len(s) and s[j] in "+-":
        j += 1
    k = j
    while k < len(s) and s[k].isdigit():
        k += 1
    if k == j:
        return 0, i
    return int(s[i:k]), k


def _parse_vertex(tok, obj, offset):
    fields = [0, 0, 0]
    i = 0
    for k in range(3):
        fields[k], end = _strtol(tok, i)
        if end >= len(tok) or tok[end] != "/":
            break
        i = end + 1
    sizes = (len(obj["positions"]), len(obj["texcoords"]), len(obj["normals"]))
    return tuple(sizes[k] + f if f < 0 else (f - offset[k] if f != 0 else 0) for k, f in enumerate(fields))


def load_obj(path):
    with open(path, "rb") as f:
        text = f.read().decode("latin-1")
    result = []
    offset = [1, 1, 1]
    pend = {"positions": [], "texcoords": [], "normals": []}
    lib = {}
    fallback = _default_material()
    material = fallback

    def new_object(name=""):
        result.append({"name": name, "groups": [], "positions": [], "texcoords": [], "normals": []})
        return result[-1]

    def current():
        return result[-1] if result else new_object()

    for raw in text.split("\n"):
        toks = raw.split()
        if not toks:
            continue
        ident, rest = toks[0], toks[1:]
        if ident == "mtllib":
            lib = {}
            for name in rest:
                for k, v in load_mtllib(os.path.join(os.path.dirname(path), name)).items():
                    lib.setdefault(k, v)
        elif ident == "usemtl":
            name = rest[0] if rest else ""
            material = lib.get(name, fallback)
            o = current()
            if not o["groups"] or o["groups"][-1]["faces"]:
                o["groups"].append({"name": name, "faces": [], "material": None})
            o["groups"][-1]["material"] = material
        elif ident == "o":
            if result:
                offset[0] += len(result[-1]["positions"])
                offset[1] += len(result[-1]["texcoords"])
                offset[2] += len(result[-1]["normals"])
            new_object(rest[0] if rest else "")
        elif ident == "g":
            name = rest[0] if rest else ""
            if not result:
                new_object(name)
            result[-1]["groups"].append({"name": name, "faces": [], "material": material})
        elif ident in ("v", "vn"):
            pend["positions" if ident == "v" else "normals"].append(tuple(_floats(rest, 3)))
        elif ident == "vt":
            pend["texcoords"].append(tuple(_floats(rest, 2)))
        elif ident == "f":
            o = current()
            if not o["groups"]:
                o["groups"].append({"name": "Default", "faces": [], "material": material})
            for key in pend:
                o[key].extend(pend[key])
                pend[key] = []
            o["groups"][-1]["faces"].append([_parse_vertex(t, o, offset) for t in rest])
    for o in result:
        if not o["texcoords"]:
            o["texcoords"].append((0.0, 0.0))
        if not o["normals"]:
            o["normals"].append((0.0, 1.0, 0.0))
    return result


def _key(x):
    # Floats compare by value: -0.0 == 0.0; a NaN equals nothing, so it is made unique.
    return tuple(0.0 if v == 0 else (object() if v != v else v) for v in x)


def triangulate(obj):
    out = []
    for g in obj["groups"]:
        m = {"name": obj["name"] + "/" + g["name"], "material": g["material"], "indices": [], "positions": [],
             "normals": [], "texcoords": []}
        seen = {}

        def vertex(v):
            p, t, n = v
            for idx, lst in ((p, "positions"), (t, "texcoords"), (n, "normals")):
                if idx < 0 or idx >= len(obj[lst]):
                    raise IndexError("face index out of range")
            pos, tex, nor = obj["positions"][p], obj["texcoords"][t], obj["normals"][n]
            key = _key(pos + nor + tex)
            if key not in seen:
                seen[key] = len(m["positions"])
                m["positions"].append(pos)
                m["normals"].append(nor)
                m["texcoords"].append(tex)
            return seen[key]

        for face in g["faces"]:
            if len(face) <= 2:
                continue
            first = vertex(face[0])
            last = vertex(face[1])
            for v in face[2:]:
                nxt = vertex(v)
                m["indices"] += [first, last, nxt]
                last = nxt
        out.append(m)
    return out


def load_groups(path):
    """load_obj then triangulate every object: the groups in file order, arrays as numpy."""
    groups = []
    for o in load_obj(path):
        for m in triangulate(o):
            groups.append({"name": m["name"], "material": m["material"],
                           "indices": np.asarray(m["indices"], np.uint32).reshape(-1, 3),
                           "positions": np.asarray(m["positions"], np.float32).reshape(-1, 3),
                           "normals": np.asarray(m["normals"], np.float32).reshape(-1, 3),
                           "texcoords": np.asarray(m["texcoords"], np.float32).reshape(-1, 2)})
    return groups


def material_record(mat):
    """(albedo_rgba8, ior, roughness, transmission) as load_object_file sets them: albedo
    vec4ui8(255 * (Kd, 1)) with float -> uint8 truncation; roughness max(0.054, (1/Ns)^(1/pi))
    evaluated in float32; transmission 1 - d."""
    def unorm(v):
        return int(min(max(np.float32(255.0) * np.float32(v), 0.0), 255.0))
    d = mat["diffuse"]
    rgba = unorm(d[0]) | unorm(d[1]) << 8 | unorm(d[2]) << 16 | 255 << 24
    inv = np.float32(1.0) / np.float32(mat["specularity"])
    rough = max(np.float32(0.054), np.float32(math.pow(float(inv), float(np.float32(1.0 / 3.1415926535897)))))
    trans = np.float32(1.0) - np.float32(mat["dissolve"])
    return rgba, float(np.float32(mat["ior"])), float(rough), float(trans)


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]
