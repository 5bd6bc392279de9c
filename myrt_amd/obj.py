"""Wavefront OBJ/MTL ingest over the C ABI.

ObjFile mirrors myrt::obj::load_obj + triangulate (src/obj.hpp, src/obj.cpp:96-361);
load_object_file mirrors the app's load_object_file (src/main.cpp:902-949): every non-empty
group becomes a geometry with a material mapped from its MTL entry, a map_Kd image becomes the
material's albedo texture, and each object carries the transformation scale(import_scale).
"""
import ctypes
from collections import namedtuple

import numpy as np

from . import MrtError, _check, _ptr, _SZ, lib

NO_GROUP = 0xFFFFFFFF


class _GroupInfo(ctypes.Structure):
    _fields_ = [("n_idx", ctypes.c_uint64), ("n_vtx", ctypes.c_uint64), ("diffuse", ctypes.c_float * 3),
                ("specularity", ctypes.c_float), ("ior", ctypes.c_float), ("dissolve", ctypes.c_float),
                ("name", ctypes.c_char * 128), ("material_name", ctypes.c_char * 64),
                ("albedo_texture", ctypes.c_char * 256)]


assert ctypes.sizeof(_GroupInfo) == 488

# myrt::geometric_object as load_object_file fills it (src/main.cpp:941-945).
GeometricObject = namedtuple("GeometricObject", "name geometry material transformation")


class ObjFile:
    """A parsed and triangulated OBJ file. `groups` lists, per group, its name ("object/group"),
    the MTL fields (diffuse, specularity, ior, dissolve, material_name, albedo_texture) and the
    arrays indices (n, 3) uint32, positions / normals (v, 3) and texcoords (v, 2) float32."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        _check(lib().mrt_obj_load(str(path).encode(), ctypes.byref(h)))
        self.h = h
        n = _SZ()
        _check(lib().mrt_obj_num_groups(self.h, ctypes.byref(n)))
        self.groups = [self._group(i) for i in range(n.value)]

    def _group(self, i):
        info = _GroupInfo()
        _check(lib().mrt_obj_group(self.h, i, ctypes.byref(info)))
        idx = np.zeros(info.n_idx, np.uint32)
        pos = np.zeros((info.n_vtx, 3), np.float32)
        nrm = np.zeros((info.n_vtx, 3), np.float32)
        uv = np.zeros((info.n_vtx, 2), np.float32)
        _check(lib().mrt_obj_group_copy(self.h, i, _ptr(idx), _ptr(pos), _ptr(nrm), _ptr(uv)))
        return {"name": info.name.decode(), "material_name": info.material_name.decode(),
                "diffuse": tuple(info.diffuse), "specularity": info.specularity, "ior": info.ior,
                "dissolve": info.dissolve, "albedo_texture": info.albedo_texture.decode(),
                "indices": idx.reshape(-1, 3), "positions": pos, "normals": nrm, "texcoords": uv}

    def close(self):
        if self.h:
            lib().mrt_obj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_ldr_texture(path):
    """An 8-bit image as the reference's ldr_texture reads it (src/main.cpp:77-100): RGB with
    alpha 1, top row first, unorm to float. Needs PIL; raises MrtError without it."""
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - PIL is part of the image
        raise MrtError("loading texture '%s' needs PIL" % path) from e
    with Image.open(path) as im:
        rgb = np.asarray(im.convert("RGB"), np.float32) / 255.0
    rgba = np.ones(rgb.shape[:2] + (4,), np.float32)
    rgba[..., :3] = rgb
    return rgba


def load_object_file(scene, path, import_scale=6.0, load_textures=True):
    """load_object_file (src/main.cpp:902-949): pushes every non-empty group of the OBJ at `path`
    into `scene` and returns the GeometricObjects; enqueue them with enqueue_objects."""
    f = ObjFile(path)
    try:
        n = len(f.groups)
        geoms = np.zeros(max(n, 1), np.uint32)
        mats = np.zeros(max(n, 1), np.uint32)
        _check(lib().mrt_scene_push_obj(scene.h, f.h, _ptr(geoms), _ptr(mats), n))
        scale = np.diag([import_scale, import_scale, import_scale, 1.0]).astype(np.float32)
        out = []
        for g, geo, mat in zip(f.groups, geoms[:n], mats[:n]):
            if geo == NO_GROUP:
                continue
            if load_textures and g["albedo_texture"]:
                tex = scene.push_texture(load_ldr_texture(g["albedo_texture"]))
                scene.set_parameter(int(mat), "albedo_texture", float(tex))
            out.append(GeometricObject(g["name"], int(geo), int(mat), scale.copy()))
        return out
    finally:
        f.close()


def enqueue_objects(scene, objects):
    """Enqueues each GeometricObject with its own transformation, as the app does every frame."""
    return [scene.enqueue(o.geometry, o.transformation, o.material) for o in objects]
