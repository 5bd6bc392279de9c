"""Python binding of libmyrt_amd.so (ctypes over the C ABI in include/myrt.h).

The library owns its device memory; numpy arrays are host buffers. Nothing here falls back to a
CPU path: loading fails loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MYRT_AMD_LIBRARY selects another build of the library (launch-shape tuning builds).
LIB_PATH = os.environ.get("MYRT_AMD_LIBRARY") or os.path.join(HERE, "libmyrt_amd.so")

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("bary", "<u4")])
NODE_DTYPE = np.dtype([("min", "<f4", 3), ("type_and_parent", "<u4"), ("max", "<f4", 3), ("first", "<u4"),
                       ("second", "<u4"), ("pad", "<u4", 3)])
AABB_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("centroid", "<f4", 3), ("has_centroid", "<u4")])
WIDE_NODE_DTYPE = np.dtype([("origin", "<f4", 3), ("exp", "i1", 3), ("nchildren", "u1"), ("qlo", "u1", (3, 4)),
                            ("qhi", "u1", (3, 4)), ("pad", "<u4", 2), ("entry", "<u4", 4)])
TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("prim", "<u4"), ("v1", "<f4", 3), ("pad1", "<u4"), ("v2", "<f4", 3),
                      ("pad2", "<u4")])
INSTANCE_DTYPE = np.dtype([("inv", "<f4", 12), ("blas_root", "<u4"), ("geom", "<u4"), ("inst", "<u4"),
                           ("pad", "<u4")])
assert WIDE_NODE_DTYPE.itemsize == 64 and TRI_DTYPE.itemsize == 48 and INSTANCE_DTYPE.itemsize == 64

MATERIAL_DTYPE = np.dtype([("albedo_rgba8", "<u4"), ("ior", "<f4"), ("roughness", "<f4"), ("metallic", "<f4"),
                           ("transmission", "<f4"), ("texture", "<u4"), ("pad", "<u4", 2)])
LIGHT_DTYPE = np.dtype([("corner", "<f4", 3), ("u", "<f4", 3), ("v", "<f4", 3), ("emission", "<f4", 3)])
SDF_NODE_DTYPE = np.dtype([("op", "<u4"), ("material", "<u4"), ("p", "<f4", 6)])
SDF = dict(SPHERE=1, BOX=2, TORUS=3, CAPSULE=4, MENGER=5, UNION=16, SUBTRACT=17, INTERSECT=18, SMOOTH_UNION=19,
           SMOOTH_SUBTRACT=20, SMOOTH_INTERSECT=21, PUSH_TRANSLATE=32, PUSH_MIRROR=33, POP=34)
SDF_HIT = 0x80000000
SDF_DRAWABLE_DTYPE = np.dtype([("inv", "<f4", 12), ("box_lo", "<f4", 3), ("tape_offset", "<u4"),
                               ("box_hi", "<f4", 3), ("tape_len", "<u4")])
QUERY_DTYPE = np.dtype([("pos", "<f4", 3), ("roughness", "<f4"), ("dir", "<f4", 3), ("pad0", "<f4"),
                        ("nrm", "<f4", 3), ("pad1", "<f4"), ("albedo", "<f4", 3), ("pad2", "<f4")])
CACHE_PARAMS = 20672

SPLIT_NONE = 0
SPLIT_AABBS = 1
MISS = 0xFFFFFFFF


class MrtError(RuntimeError):
    pass


class _PrepareResult(ctypes.Structure):
    _fields_ = [("geometries_changed", ctypes.c_int32), ("drawables_changed", ctypes.c_int32),
                ("num_geometries", ctypes.c_uint32), ("num_instances", ctypes.c_uint32),
                ("blas_wide_nodes", ctypes.c_uint32), ("blas_tri_records", ctypes.c_uint32),
                ("tlas_wide_nodes", ctypes.c_uint32), ("max_stack_entries", ctypes.c_uint32),
                ("tlas_refitted", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]


class _RenderStats(ctypes.Structure):
    _fields_ = [("samples", ctypes.c_uint64), ("milliseconds", ctypes.c_double), ("rays", ctypes.c_uint64),
                ("nonfinite", ctypes.c_uint64)]


class _TraceStats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint64), ("node_visits", ctypes.c_uint64), ("tri_tests", ctypes.c_uint64),
                ("instance_visits", ctypes.c_uint64)]


_lib = None

# Every exported entry point and its argument types; include/myrt.h declares exactly these.
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_U32P = ctypes.POINTER(ctypes.c_uint32)
SIGNATURES = {
    "mrt_last_error": ([], ctypes.c_char_p),
    "mrt_device_info": ([ctypes.c_char_p, _SZ, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "mrt_device_alloc": ([_SZ, ctypes.POINTER(_P)], ctypes.c_int),
    "mrt_device_free": ([_P], ctypes.c_int),
    "mrt_memcpy_h2d": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_memcpy_d2h": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_device_synchronize": ([], ctypes.c_int),
    "mrt_device_copy": ([_P, _P, _SZ, _P], ctypes.c_int),
    "mrt_timer_create": ([ctypes.POINTER(_P)], ctypes.c_int),
    "mrt_timer_destroy": ([_P], ctypes.c_int),
    "mrt_timer_start": ([_P, _P], ctypes.c_int),
    "mrt_timer_stop": ([_P, _P], ctypes.c_int),
    "mrt_timer_elapsed_ms": ([_P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "mrt_generate_triangle_bounds": ([_P, _SZ, _P, _SZ, _P], ctypes.c_int),
    "mrt_bvh_build": ([_P, _SZ, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "mrt_bvh_build_triangles": ([_P, _SZ, _P, _SZ, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
    "mrt_bvh_sizes": ([_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], ctypes.c_int),
    "mrt_bvh_copy": ([_P, _P, _P], ctypes.c_int),
    "mrt_bvh_destroy": ([_P], ctypes.c_int),
    "mrt_scene_create": ([ctypes.POINTER(_P)], ctypes.c_int),
    "mrt_scene_destroy": ([_P], ctypes.c_int),
    "mrt_scene_push_geometry": ([_P, _P, _SZ, _P, _P, _P, _SZ, _U32P], ctypes.c_int),
    "mrt_scene_enqueue_geometry": ([_P, ctypes.c_uint32, ctypes.c_uint32, _P, _U32P], ctypes.c_int),
    "mrt_scene_prepare": ([_P, ctypes.POINTER(_PrepareResult)], ctypes.c_int),
    "mrt_scene_instance_inverses": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_scene_set_tlas_refit": ([_P, ctypes.c_uint32], ctypes.c_int),
    "mrt_scene_update_transforms_device": ([_P, _P, _SZ, _P], ctypes.c_int),
    "mrt_scene_update_status": ([_P, _U32P], ctypes.c_int),
    "mrt_scene_build_host": ([_P, ctypes.POINTER(_PrepareResult)], ctypes.c_int),
    "mrt_scene_copy_wide": ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
    "mrt_trace_closest": ([_P, _P, _SZ, _P], ctypes.c_int),
    "mrt_trace_any": ([_P, _P, _SZ, _P], ctypes.c_int),
    "mrt_trace_closest_device": ([_P, _P, _SZ, _P, _P], ctypes.c_int),
    "mrt_trace_any_device": ([_P, _P, _SZ, _P, _P], ctypes.c_int),
    "mrt_trace_closest_device_noalloc": ([_P, _P, _SZ, _P, _P], ctypes.c_int),
    "mrt_trace_stats_device": ([_P, _P, _SZ, ctypes.c_int, ctypes.POINTER(_TraceStats), _P], ctypes.c_int),
    "mrt_scene_pick": ([_P, _P, _P, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "mrt_scene_push_material": ([_P, _P, _U32P], ctypes.c_int),
    "mrt_scene_set_material": ([_P, ctypes.c_uint32, _P], ctypes.c_int),
    "mrt_scene_get_material": ([_P, ctypes.c_uint32, _P], ctypes.c_int),
    "mrt_scene_set_material_param": ([_P, ctypes.c_uint32, ctypes.c_char_p, _P, _SZ], ctypes.c_int),
    "mrt_scene_register_layout": ([_P, ctypes.c_char_p, _U32P], ctypes.c_int),
    "mrt_scene_layout_add_parameter": ([_P, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_char_p], ctypes.c_int),
    "mrt_scene_find_type": ([_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), _U32P], ctypes.c_int),
    "mrt_scene_layout_info": ([_P, ctypes.c_uint32, _U32P, _U32P], ctypes.c_int),
    "mrt_scene_layout_parameter": ([_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p, _SZ, _U32P, _U32P,
                                    ctypes.POINTER(ctypes.c_int32), _U32P], ctypes.c_int),
    "mrt_scene_push_material_type": ([_P, ctypes.c_uint32, _P, _U32P], ctypes.c_int),
    "mrt_scene_push_material_of_type": ([_P, ctypes.c_uint32, _U32P], ctypes.c_int),
    "mrt_scene_set_parameter": ([_P, ctypes.c_uint32, ctypes.c_char_p, _P, _SZ, ctypes.POINTER(ctypes.c_int)],
                                ctypes.c_int),
    "mrt_scene_material_slot": ([_P, ctypes.c_uint32, _U32P, _U32P], ctypes.c_int),
    "mrt_scene_material_blocks": ([_P, _P, _SZ, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "mrt_scene_set_lights": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_scene_set_environment": ([_P, ctypes.c_float], ctypes.c_int),
    "mrt_scene_push_sdf": ([_P, _P, _SZ, _U32P], ctypes.c_int),
    "mrt_scene_enqueue_sdf": ([_P, ctypes.c_uint32, _P, _U32P], ctypes.c_int),
    "mrt_scene_set_sdf_marching": ([_P, ctypes.c_int, ctypes.c_float], ctypes.c_int),
    "mrt_scene_set_sdf_param": ([_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P, _SZ], ctypes.c_int),
    "mrt_scene_copy_sdf_grid": ([_P, ctypes.c_uint32, _P, _P, _U32P, _P, _SZ], ctypes.c_int),
    "mrt_scene_sdf_drawables": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_renderer_create": ([ctypes.POINTER(_P)], ctypes.c_int),
    "mrt_renderer_destroy": ([_P], ctypes.c_int),
    "mrt_renderer_set_camera": ([_P, _P, _P], ctypes.c_int),
    "mrt_renderer_set_num_bounces": ([_P, ctypes.c_int], ctypes.c_int),
    "mrt_renderer_set_material_sort": ([_P, ctypes.c_int], ctypes.c_int),
    "mrt_renderer_invalidate": ([_P], ctypes.c_int),
    "mrt_renderer_set_dof_enabled": ([_P, ctypes.c_int], ctypes.c_int),
    "mrt_renderer_set_lens_size": ([_P, ctypes.c_float, ctypes.c_float], ctypes.c_int),
    "mrt_renderer_set_focus": ([_P, ctypes.c_float], ctypes.c_int),
    "mrt_renderer_set_bokeh_mask": ([_P, _P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mrt_renderer_set_cubemap": ([_P, _P, ctypes.c_int], ctypes.c_int),
    "mrt_scene_push_texture": ([_P, _P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "mrt_renderer_set_cache": ([_P, _P, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mrt_renderer_set_cache_training": ([_P, ctypes.c_int, ctypes.c_float], ctypes.c_int),
    "mrt_renderer_set_sample_base": ([_P, ctypes.c_uint32], ctypes.c_int),
    "mrt_renderer_sample_count": ([_P, _U32P], ctypes.c_int),
    "mrt_render": ([_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "mrt_renderer_read": ([_P, _P], ctypes.c_int),
    "mrt_renderer_stats": ([_P, ctypes.POINTER(_RenderStats)], ctypes.c_int),
    "mrt_cache_create": ([ctypes.POINTER(_P), ctypes.c_uint64], ctypes.c_int),
    "mrt_cache_destroy": ([_P], ctypes.c_int),
    "mrt_cache_set_bounds": ([_P, _P, _P], ctypes.c_int),
    "mrt_cache_set_adam": ([_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float], ctypes.c_int),
    "mrt_cache_get_params": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_cache_set_params": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_cache_infer": ([_P, _P, _SZ, _P], ctypes.c_int),
    "mrt_cache_infer_device": ([_P, _P, _SZ, _P, _P], ctypes.c_int),
    "mrt_cache_train": ([_P, _P, _P, _SZ, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "mrt_cache_train_device": ([_P, _P, _P, _SZ, _P], ctypes.c_int),
    "mrt_cache_gradient": ([_P, _P, _P, _SZ, _P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "mrt_denoiser_create": ([ctypes.POINTER(_P), ctypes.c_uint64], ctypes.c_int),
    "mrt_denoiser_destroy": ([_P], ctypes.c_int),
    "mrt_denoiser_num_params": ([ctypes.POINTER(_SZ)], ctypes.c_int),
    "mrt_denoiser_get_params": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_denoiser_set_params": ([_P, _P, _SZ], ctypes.c_int),
    "mrt_denoiser_set_adam": ([_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float], ctypes.c_int),
    "mrt_denoiser_forward": ([_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "mrt_denoiser_gradient": ([_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P,
                               ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "mrt_denoiser_train": ([_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
                           ctypes.c_int),
    "mrt_denoiser_train_device": ([_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_float), _P], ctypes.c_int),
    "mrt_denoiser_denoise": ([_P, _P, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
    "mrt_denoiser_denoise_device": ([_P, _P, ctypes.c_int, ctypes.c_int, _P, _P], ctypes.c_int),
    "mrt_renderer_read_device": ([_P, _P, _P], ctypes.c_int),
    "mrt_renderer_read_aux": ([_P, _P], ctypes.c_int),
    "mrt_post_pass_device": ([ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _P],
                             ctypes.c_int),
    "mrt_post_process": ([_P, _P, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, ctypes.c_float, ctypes.c_float],
                         ctypes.c_int),
    "mrt_renderer_denoise": ([_P, _P, _P], ctypes.c_int),
    "mrt_obj_load": ([ctypes.c_char_p, _P], ctypes.c_int),
    "mrt_obj_destroy": ([_P], ctypes.c_int),
    "mrt_obj_num_groups": ([_P, _P], ctypes.c_int),
    "mrt_obj_group": ([_P, _SZ, _P], ctypes.c_int),
    "mrt_obj_group_copy": ([_P, _SZ, _P, _P, _P, _P], ctypes.c_int),
    "mrt_scene_push_obj": ([_P, _P, _P, _P, _SZ], ctypes.c_int),
}


def lib():
    """Loads libmyrt_amd.so once; raises if it is missing (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MrtError("libmyrt_amd.so is not built (run python -m myrt_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().mrt_last_error()
        raise MrtError("myrt error %d: %s" % (rc, msg.decode() if msg else ""))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def device_info():
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int(0)
    _check(lib().mrt_device_info(name, 256, ctypes.byref(cus)))
    return name.value.decode(), cus.value


def synchronize():
    _check(lib().mrt_device_synchronize())


class DeviceBuffer:
    """A block of device memory owned by libmyrt_amd."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().mrt_device_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().mrt_memcpy_h2d(b.ptr, _ptr(a), a.nbytes))
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        _check(lib().mrt_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().mrt_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """Stream-ordered event timer (start/stop recorded on a HIP stream, NULL = default)."""

    def __init__(self):
        p = ctypes.c_void_p()
        _check(lib().mrt_timer_create(ctypes.byref(p)))
        self.h = p

    def start(self, stream=None):
        _check(lib().mrt_timer_start(self.h, stream))

    def stop(self, stream=None):
        _check(lib().mrt_timer_stop(self.h, stream))

    def elapsed_ms(self):
        ms = ctypes.c_float()
        _check(lib().mrt_timer_elapsed_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().mrt_timer_destroy(self.h)
        except Exception:
            pass


def triangle_bounds(indices, points):
    idx = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(idx.size // 3, dtype=AABB_DTYPE)
    _check(lib().mrt_generate_triangle_bounds(_ptr(idx), idx.size, _ptr(pts), pts.shape[0], _ptr(out)))
    return out


def _bvh_result(h):
    nn, ni = ctypes.c_size_t(), ctypes.c_size_t()
    try:
        _check(lib().mrt_bvh_sizes(h, ctypes.byref(nn), ctypes.byref(ni)))
        nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
        idx = np.zeros(ni.value, dtype=np.uint32)
        _check(lib().mrt_bvh_copy(h, _ptr(nodes), _ptr(idx)))
    finally:
        lib().mrt_bvh_destroy(h)
    return nodes, idx


def bvh_build(boxes, split=SPLIT_NONE):
    """bvh(aabbs, split) -> (nodes, reordered_indices)."""
    boxes = np.ascontiguousarray(boxes, dtype=AABB_DTYPE)
    h = ctypes.c_void_p()
    _check(lib().mrt_bvh_build(_ptr(boxes), boxes.size, split, ctypes.byref(h)))
    return _bvh_result(h)


def bvh_build_triangles(indices, points, split=SPLIT_AABBS):
    idx = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    h = ctypes.c_void_p()
    _check(lib().mrt_bvh_build_triangles(_ptr(idx), idx.size, _ptr(pts), pts.shape[0], split, ctypes.byref(h)))
    return _bvh_result(h)


def make_rays(origins, directions, tmin=0.0, tmax=np.inf):
    o = np.asarray(origins, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(directions, dtype=np.float32).reshape(-1, 3)
    r = np.zeros(o.shape[0], dtype=RAY_DTYPE)
    r["o"] = o
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def unpack_bary(hits):
    b = hits["bary"].astype(np.uint32)
    return (b & 0xFFFF).astype(np.float32) / 65535.0, (b >> 16).astype(np.float32) / 65535.0


class Scene:
    """myrt::scene (src/pathtracer/scene.hpp:60-255) over the C ABI."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().mrt_scene_create(ctypes.byref(h)))
        self.h = h
        self.last_prepare = None

    def close(self):
        if self.h:
            lib().mrt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push_geometry(self, indices, points, normals=None, uvs=None):
        idx = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        uv = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        out = ctypes.c_uint32()
        _check(lib().mrt_scene_push_geometry(self.h, _ptr(idx), idx.size, _ptr(pts), _ptr(nrm), _ptr(uv),
                                             pts.shape[0], ctypes.byref(out)))
        return out.value

    def enqueue(self, geometry, transformation=None, material=0):
        """scene::enqueue(geometry, material, mat4): `transformation` is a 4x4 array in the
        usual row/column sense (m[r][c]); it is passed column-major like rnu::mat4."""
        m = np.eye(4, dtype=np.float32) if transformation is None else np.asarray(transformation, np.float32)
        colmajor = np.ascontiguousarray(m.T).ravel()
        out = ctypes.c_uint32()
        _check(lib().mrt_scene_enqueue_geometry(self.h, int(geometry), int(material), _ptr(colmajor),
                                                ctypes.byref(out)))
        return out.value

    def _result(self, r):
        return {k: getattr(r, k) for k, _ in _PrepareResult._fields_}

    def prepare(self):
        r = _PrepareResult()
        _check(lib().mrt_scene_prepare(self.h, ctypes.byref(r)))
        self.last_prepare = self._result(r)
        return self.last_prepare

    def build_host(self):
        """BVH builds and wide collapse only (no GPU)."""
        r = _PrepareResult()
        _check(lib().mrt_scene_build_host(self.h, ctypes.byref(r)))
        self.last_prepare = self._result(r)
        return self.last_prepare

    def wide_arrays(self):
        p = self.last_prepare
        blas = np.zeros(p["blas_wide_nodes"], WIDE_NODE_DTYPE)
        tris = np.zeros(p["blas_tri_records"], TRI_DTYPE)
        tlas = np.zeros(p["tlas_wide_nodes"], WIDE_NODE_DTYPE)
        n_items = p["num_instances"]
        items = np.zeros(max(n_items, 1), np.uint32)
        inst = np.zeros(max(p["num_instances"], 1), INSTANCE_DTYPE)
        _check(lib().mrt_scene_copy_wide(self.h, _ptr(blas), _ptr(tris), _ptr(tlas), _ptr(items), _ptr(inst)))
        return blas, tris, tlas, items[:n_items], inst[:p["num_instances"]]

    def set_tlas_refit(self, max_refits):
        """Refit the TLAS on the GPU for frames that only move instances (0: always rebuild)."""
        _check(lib().mrt_scene_set_tlas_refit(self.h, int(max_refits)))

    def update_transforms_device(self, d_mats, n, stream=None):
        """Move the prepared instances to n device-resident column-major 4x4 matrices."""
        _check(lib().mrt_scene_update_transforms_device(self.h, d_mats.ptr, n, stream))

    def update_status(self):
        out = ctypes.c_uint32()
        _check(lib().mrt_scene_update_status(self.h, ctypes.byref(out)))
        return out.value

    def instance_inverses(self, n):
        out = np.zeros((max(n, 1), 12), np.float32)
        _check(lib().mrt_scene_instance_inverses(self.h, _ptr(out), max(n, 1)))
        return out[:n]

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(rays.size, HIT_DTYPE)
        _check(lib().mrt_trace_closest(self.h, _ptr(rays), rays.size, _ptr(hits)))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(rays.size, np.uint8)
        _check(lib().mrt_trace_any(self.h, _ptr(rays), rays.size, _ptr(occ)))
        return occ

    def trace_closest_device(self, d_rays, n, d_hits, stream=None):
        _check(lib().mrt_trace_closest_device(self.h, d_rays.ptr, n, d_hits.ptr, stream))

    def trace_closest_device_noalloc(self, d_rays, n, d_hits, stream=None):
        _check(lib().mrt_trace_closest_device_noalloc(self.h, d_rays.ptr, n, d_hits.ptr, stream))

    def trace_any_device(self, d_rays, n, d_occ, stream=None):
        _check(lib().mrt_trace_any_device(self.h, d_rays.ptr, n, d_occ.ptr, stream))

    def trace_stats(self, d_rays, n, any_hit=False):
        s = _TraceStats()
        _check(lib().mrt_trace_stats_device(self.h, d_rays.ptr, n, 1 if any_hit else 0, ctypes.byref(s), None))
        return {k: getattr(s, k) for k, _ in _TraceStats._fields_}

    def pick(self, origin, direction, length=np.inf):
        """scene::pick(ray_t) -> None or the hit record."""
        r = make_rays([origin], [direction], 0.0, length)
        h = np.zeros(1, HIT_DTYPE)
        found = ctypes.c_int()
        _check(lib().mrt_scene_pick(self.h, _ptr(r), _ptr(h), ctypes.byref(found)))
        return h[0] if found.value else None


def pack_rgba8(rgb, a=1.0):
    c = [int(round(min(max(float(x), 0.0), 1.0) * 255)) for x in list(rgb) + [a]][:4]
    return c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24)


def material(albedo=(0.8, 0.8, 0.8), ior=1.5, roughness=0.5, metallic=0.0, transmission=0.0):
    m = np.zeros(1, MATERIAL_DTYPE)
    m["albedo_rgba8"] = pack_rgba8(albedo)
    m["ior"], m["roughness"], m["metallic"], m["transmission"] = ior, roughness, metallic, transmission
    return m


def quad_light(corner, u, v, emission):
    l = np.zeros(1, LIGHT_DTYPE)
    l["corner"], l["u"], l["v"], l["emission"] = corner, u, v, emission
    return l


def _scene_push_material(self, mat):
    """scene::push_material: `mat` is a material type id (from push_material_type; the new
    material's parameters start at zero) or a MATERIAL_DTYPE record of the built-in layout."""
    out = ctypes.c_uint32()
    if isinstance(mat, (int, np.integer)):
        _check(lib().mrt_scene_push_material_of_type(self.h, int(mat), ctypes.byref(out)))
        return out.value
    mat = np.ascontiguousarray(mat, MATERIAL_DTYPE)
    _check(lib().mrt_scene_push_material(self.h, _ptr(mat), ctypes.byref(out)))
    return out.value


def _scene_register_type(self, name, parameters):
    """object_type(name) with add_parameter(name, type) for each (name, type name) pair, type
    names resolved like types_registry::find (layouts first, then "float", "vec4unorm8", ...)."""
    out = ctypes.c_uint32()
    _check(lib().mrt_scene_register_layout(self.h, name.encode(), ctypes.byref(out)))
    for pname, tname in parameters:
        _check(lib().mrt_scene_layout_add_parameter(self.h, out.value, pname.encode(), tname.encode()))
    return out.value


def _scene_find_type(self, name):
    """types_registry::find: ("constant", parameter_type_id) or ("layout", layout id)."""
    c, l = ctypes.c_int32(), ctypes.c_uint32()
    _check(lib().mrt_scene_find_type(self.h, name.encode(), ctypes.byref(c), ctypes.byref(l)))
    return ("constant", c.value) if c.value >= 0 else ("layout", l.value)


def _scene_layout(self, layout):
    """(blocks, [(name, offset, blocks, constant id or -1, inner layout or None), ...])."""
    blocks, n = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib().mrt_scene_layout_info(self.h, int(layout), ctypes.byref(blocks), ctypes.byref(n)))
    params = []
    for i in range(n.value):
        name = ctypes.create_string_buffer(256)
        off, b, inner = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        c = ctypes.c_int32()
        _check(lib().mrt_scene_layout_parameter(self.h, int(layout), i, name, 256, ctypes.byref(off),
                                                ctypes.byref(b), ctypes.byref(c), ctypes.byref(inner)))
        params.append((name.value.decode(), off.value, b.value, c.value, None if c.value >= 0 else inner.value))
    return blocks.value, params


def _scene_push_material_type(self, layout, inputs=None):
    """scene::push_material_type: `inputs` maps GGX inputs (albedo, ior, roughness, metallic,
    transmission, albedo_texture) to parameter paths; missing inputs bind by name, "" unbinds."""
    names = ("albedo", "ior", "roughness", "metallic", "transmission", "albedo_texture")
    arr = None
    keep = []
    if inputs is not None:
        unknown = set(inputs) - set(names)
        if unknown:
            raise MrtError("unknown shading inputs %s" % sorted(unknown))
        arr = (ctypes.c_char_p * 6)()
        for i, k in enumerate(names):
            if k in inputs:
                keep.append(inputs[k].encode())
                arr[i] = keep[-1]
    out = ctypes.c_uint32()
    _check(lib().mrt_scene_push_material_type(self.h, int(layout), arr, ctypes.byref(out)))
    return out.value


def _scene_material_slot(self, mid):
    t, o = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib().mrt_scene_material_slot(self.h, int(mid), ctypes.byref(t), ctypes.byref(o)))
    return t.value, o.value


def _scene_material_blocks(self):
    """material_buffer::buffer() as uint32 blocks (view floats with .view(np.float32))."""
    total = ctypes.c_size_t()
    _check(lib().mrt_scene_material_blocks(self.h, None, 0, ctypes.byref(total)))
    out = np.zeros(max(total.value, 1), np.uint32)
    _check(lib().mrt_scene_material_blocks(self.h, _ptr(out), out.size, ctypes.byref(total)))
    return out[:total.value]


def _scene_get_material(self, mid):
    out = np.zeros(1, MATERIAL_DTYPE)
    _check(lib().mrt_scene_get_material(self.h, int(mid), _ptr(out)))
    return out


def _scene_set_material_param(self, mid, name, value):
    """scene::set_parameter(material, name, value). A numpy value (array or scalar) or bytes is
    written as its object representation, like the reference's set_parameter<T>: its size must
    match the parameter, and a name the layout lacks is ignored (the call returns False). Python
    numbers and lists go through the float form (vec4unorm8 from 3 or 4 floats in [0, 1],
    integral values for integer types), where an unknown name is an error."""
    if isinstance(value, (bytes, bytearray, memoryview, np.ndarray, np.generic)):
        raw = bytes(value) if not isinstance(value, (np.ndarray, np.generic)) else np.ascontiguousarray(value).tobytes()
        buf = ctypes.create_string_buffer(raw, max(len(raw), 1))
        found = ctypes.c_int()
        _check(lib().mrt_scene_set_parameter(self.h, int(mid), name.encode(), buf, len(raw), ctypes.byref(found)))
        return bool(found.value)
    v = np.atleast_1d(np.asarray(value, np.float32))
    _check(lib().mrt_scene_set_material_param(self.h, int(mid), name.encode(), _ptr(v), v.size))
    return True


def _scene_set_lights(self, lights):
    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    _check(lib().mrt_scene_set_lights(self.h, _ptr(lights) if lights.size else None, lights.size))


def _scene_set_environment(self, scale):
    _check(lib().mrt_scene_set_environment(self.h, float(scale)))


def _scene_push_sdf(self, nodes):
    t = np.ascontiguousarray(nodes, SDF_NODE_DTYPE)
    out = ctypes.c_uint32()
    _check(lib().mrt_scene_push_sdf(self.h, _ptr(t), t.size, ctypes.byref(out)))
    return out.value


def _scene_enqueue_sdf(self, sdf, transformation=None):
    m = np.eye(4, dtype=np.float32) if transformation is None else np.asarray(transformation, np.float32)
    colmajor = np.ascontiguousarray(m.T).ravel()
    out = ctypes.c_uint32()
    _check(lib().mrt_scene_enqueue_sdf(self.h, int(sdf), _ptr(colmajor), ctypes.byref(out)))
    return out.value


def _scene_set_sdf_param(self, sdf, node, slot, values):
    """scene::set_parameter(sdf, parameter, value): p[slot:slot + len(values)] of tape node `node`."""
    v = np.atleast_1d(np.asarray(values, np.float32))
    _check(lib().mrt_scene_set_sdf_param(self.h, int(sdf), int(node), int(slot), _ptr(v), v.size))


def _scene_sdf_grid(self, sdf, with_masks=True):
    """(dim, lo, inv, masks or None) of a program's culling grid."""
    lo, inv = np.zeros(3, np.float32), np.zeros(3, np.float32)
    dim = ctypes.c_uint32()
    _check(lib().mrt_scene_copy_sdf_grid(self.h, int(sdf), _ptr(lo), _ptr(inv), ctypes.byref(dim), None, 0))
    masks = None
    if with_masks and dim.value:
        masks = np.zeros(dim.value ** 3, np.uint64)
        _check(lib().mrt_scene_copy_sdf_grid(self.h, int(sdf), _ptr(lo), _ptr(inv), ctypes.byref(dim), _ptr(masks),
                                             masks.size))
    return dim.value, lo, inv, masks


def _scene_set_sdf_marching(self, steps=400, eps=1e-5):
    _check(lib().mrt_scene_set_sdf_marching(self.h, int(steps), float(eps)))


def sdf_program(nodes):
    """Builds a tape from (op_name, material, params...) tuples."""
    t = np.zeros(len(nodes), SDF_NODE_DTYPE)
    for i, nd in enumerate(nodes):
        t[i]["op"] = SDF[nd[0]]
        t[i]["material"] = nd[1] if len(nd) > 1 else 0
        ps = list(nd[2:]) if len(nd) > 2 else []
        t[i]["p"][:len(ps)] = ps
    return t


def _scene_sdf_drawables(self, n):
    out = np.zeros(max(n, 1), SDF_DRAWABLE_DTYPE)
    _check(lib().mrt_scene_sdf_drawables(self.h, _ptr(out), out.size))
    return out[:n]


def _scene_push_texture(self, image):
    """image: (h, w, 3 or 4) floats, row 0 at v = 0. Returns the texture id (>= 1)."""
    im = np.asarray(image, np.float32)
    rgba = np.ones(im.shape[:2] + (4,), np.float32)
    rgba[..., :im.shape[2]] = im
    rgba = np.ascontiguousarray(rgba)
    out = ctypes.c_uint32()
    _check(lib().mrt_scene_push_texture(self.h, _ptr(rgba), rgba.shape[1], rgba.shape[0], ctypes.byref(out)))
    return out.value


Scene.push_texture = _scene_push_texture
Scene.sdf_drawables = _scene_sdf_drawables
Scene.push_sdf = _scene_push_sdf
Scene.enqueue_sdf = _scene_enqueue_sdf
Scene.set_sdf_marching = _scene_set_sdf_marching
Scene.set_sdf_param = _scene_set_sdf_param
Scene.sdf_grid = _scene_sdf_grid
Scene.push_material = _scene_push_material
Scene.get_material = _scene_get_material
Scene.set_parameter = _scene_set_material_param
Scene.register_type = _scene_register_type
Scene.find_type = _scene_find_type
Scene.layout = _scene_layout
Scene.push_material_type = _scene_push_material_type
Scene.material_slot = _scene_material_slot
Scene.material_blocks = _scene_material_blocks
Scene.set_lights = _scene_set_lights
Scene.set_environment = _scene_set_environment


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """View matrix (world -> camera, camera looks down -z), 4x4 m[r][c]."""
    e = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - e
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ e
    return m.astype(np.float32)


def perspective(fovy_deg, aspect, near=0.1, far=1000.0):
    f = 1.0 / np.tan(np.radians(fovy_deg) / 2)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = (far + near) / (near - far)
    m[2, 3] = 2 * far * near / (near - far)
    m[3, 2] = -1.0
    return m.astype(np.float32)


class Renderer:
    """sequential_pathtracer (src/pathtracer/sequential_pathtracer.hpp:26-58) over the C ABI."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().mrt_renderer_create(ctypes.byref(h)))
        self.h = h
        self.size = (0, 0)

    def close(self):
        if self.h:
            lib().mrt_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_camera(self, view, proj):
        v = np.ascontiguousarray(np.asarray(view, np.float32).T).ravel()
        p = np.ascontiguousarray(np.asarray(proj, np.float32).T).ravel()
        _check(lib().mrt_renderer_set_camera(self.h, _ptr(v), _ptr(p)))

    def set_num_bounces(self, n):
        _check(lib().mrt_renderer_set_num_bounces(self.h, int(n)))

    def set_material_sort(self, enabled=True):
        """Shade each bounce's paths in material order; the image does not change."""
        _check(lib().mrt_renderer_set_material_sort(self.h, 1 if enabled else 0))

    def invalidate(self):
        _check(lib().mrt_renderer_invalidate(self.h))

    def set_cubemap(self, faces):
        """faces: (6, n, n, 3 or 4) floats in +x, -x, +y, -y, +z, -z order, or None for the sky."""
        if faces is None:
            _check(lib().mrt_renderer_set_cubemap(self.h, None, 0))
            return
        f = np.asarray(faces, np.float32)
        rgba = np.ones(f.shape[:3] + (4,), np.float32)
        rgba[..., :f.shape[3]] = f
        rgba = np.ascontiguousarray(rgba)
        _check(lib().mrt_renderer_set_cubemap(self.h, _ptr(rgba), rgba.shape[1]))

    def set_dof(self, enabled=True, lens_size=(10.0, 10.0), focus=10.0, bokeh_mask=None):
        """set_dof_enabled / set_lens_size / set_focus / set_bokeh_mask; mask: (h, w, 3 or 4) floats."""
        _check(lib().mrt_renderer_set_dof_enabled(self.h, int(bool(enabled))))
        _check(lib().mrt_renderer_set_lens_size(self.h, float(lens_size[0]), float(lens_size[1])))
        _check(lib().mrt_renderer_set_focus(self.h, float(focus)))
        if bokeh_mask is None:
            _check(lib().mrt_renderer_set_bokeh_mask(self.h, None, 0, 0))
        else:
            mk = np.asarray(bokeh_mask, np.float32)
            rgba = np.ones(mk.shape[:2] + (4,), np.float32)
            rgba[..., :mk.shape[2]] = mk
            rgba = np.ascontiguousarray(rgba)
            _check(lib().mrt_renderer_set_bokeh_mask(self.h, _ptr(rgba), rgba.shape[1], rgba.shape[0]))

    def set_cache(self, cache, mode=2, train_fraction=1.0 / 16, train_steps=8, min_bounce=2):
        _check(lib().mrt_renderer_set_cache(self.h, cache.h if cache is not None else None, int(mode),
                                            float(train_fraction), int(train_steps), int(min_bounce)))

    def set_cache_training(self, train_end=4, unbiased_fraction=0.0):
        """Every path ends in the cache at vertex train_end (0: training paths never do), training
        paths learning their tail from its prediction; an unbiased_fraction of them runs to full
        depth."""
        _check(lib().mrt_renderer_set_cache_training(self.h, int(train_end), float(unbiased_fraction)))

    def set_sample_base(self, base):
        """Sampler index of the accumulation's first sample; restarts the accumulation."""
        _check(lib().mrt_renderer_set_sample_base(self.h, int(base)))

    def sample_count(self):
        out = ctypes.c_uint32()
        _check(lib().mrt_renderer_sample_count(self.h, ctypes.byref(out)))
        return out.value

    def render(self, scene, width, height, spp=1):
        _check(lib().mrt_render(self.h, scene.h, int(width), int(height), int(spp)))
        self.size = (width, height)

    def image(self):
        w, h = self.size
        out = np.zeros((h, w, 4), np.float32)
        _check(lib().mrt_renderer_read(self.h, _ptr(out)))
        return out

    def image_device(self, d_rgba, stream=None):
        """Writes the mean image into the DeviceBuffer d_rgba (H * W * 16 bytes) on the device."""
        _check(lib().mrt_renderer_read_device(self.h, d_rgba.ptr, stream))

    def stats(self):
        s = _RenderStats()
        _check(lib().mrt_renderer_stats(self.h, ctypes.byref(s)))
        return {"samples": s.samples, "milliseconds": s.milliseconds, "rays": s.rays, "nonfinite": s.nonfinite}

    def aux(self):
        """Per-pixel means of the primary hits: (albedo (H, W, 3), normal (H, W, 3), distance (H, W))."""
        w, h = self.size
        out = np.zeros((h, w, 8), np.float32)
        _check(lib().mrt_renderer_read_aux(self.h, _ptr(out)))
        return out[..., 0:3], out[..., 3:6], out[..., 6]

    def raw_features(self):
        """The denoiser's 10 raw features per pixel (H, W, 10): radiance, albedo, normal, distance."""
        albedo, normal, depth = self.aux()
        return np.concatenate([self.image()[..., :3], albedo, normal, depth[..., None]], axis=-1)

    def denoise(self, denoiser):
        """The accumulated image through the denoiser (H, W, 4), computed on the device."""
        w, h = self.size
        out = np.zeros((h, w, 4), np.float32)
        _check(lib().mrt_renderer_denoise(self.h, denoiser.h, _ptr(out)))
        return out


POST_BILATERAL, POST_TONEMAP_FILMIC, POST_FXAA = 1, 2, 3


def post_process(rgba, ops, exponent=10.0, strength=0.05):
    """Runs the post passes `ops` (sequence of POST_*) over an (h, w, 3 or 4) float image on the GPU;
    returns (h, w, 4)."""
    im = np.asarray(rgba, np.float32)
    x = np.ones(im.shape[:2] + (4,), np.float32)
    x[..., :im.shape[2]] = im
    x = np.ascontiguousarray(x)
    out = np.zeros_like(x)
    o = np.ascontiguousarray(np.asarray(list(ops), np.int32))
    _check(lib().mrt_post_process(_ptr(x), _ptr(out), x.shape[1], x.shape[0], _ptr(o) if o.size else None, o.size,
                                  float(exponent), float(strength)))
    return out


class Denoiser:
    """The U-Net denoiser over the C ABI (mrt_denoiser_*)."""

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        _check(lib().mrt_denoiser_create(ctypes.byref(h), int(seed)))
        self.h = h
        n = _SZ()
        _check(lib().mrt_denoiser_num_params(ctypes.byref(n)))
        self.n_params = n.value

    def close(self):
        if self.h:
            lib().mrt_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        _check(lib().mrt_denoiser_get_params(self.h, _ptr(out), out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().mrt_denoiser_set_params(self.h, _ptr(p), p.size))

    def set_adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(lib().mrt_denoiser_set_adam(self.h, lr, beta1, beta2, eps))

    def forward(self, raw, half=False):
        """Log-space residual y (B, H, W, 3) of raw features (B, H, W, 10)."""
        raw = np.ascontiguousarray(raw, np.float32)
        B, H, W, _ = raw.shape
        y = np.zeros((B, H, W, 3), np.float32)
        _check(lib().mrt_denoiser_forward(self.h, _ptr(raw), B, H, W, int(bool(half)), _ptr(y)))
        return y

    def gradient(self, raw, target):
        raw = np.ascontiguousarray(raw, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        B, H, W, _ = raw.shape
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(lib().mrt_denoiser_gradient(self.h, _ptr(raw), _ptr(target), B, H, W, _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, raw, target):
        raw = np.ascontiguousarray(raw, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        B, H, W, _ = raw.shape
        loss = ctypes.c_float()
        _check(lib().mrt_denoiser_train(self.h, _ptr(raw), _ptr(target), B, H, W, ctypes.byref(loss)))
        return loss.value

    def train_device(self, d_raw, d_target, B, H, W, want_loss=True):
        loss = ctypes.c_float()
        _check(lib().mrt_denoiser_train_device(self.h, d_raw.ptr, d_target.ptr, B, H, W,
                                               ctypes.byref(loss) if want_loss else None, None))
        return loss.value if want_loss else None

    def denoise(self, raw):
        """Denoised rgba (H, W, 4) of one raw feature image (H, W, 10), any size (fp16 kernels)."""
        raw = np.ascontiguousarray(raw, np.float32)
        H, W, _ = raw.shape
        out = np.zeros((H, W, 4), np.float32)
        _check(lib().mrt_denoiser_denoise(self.h, _ptr(raw), H, W, _ptr(out)))
        return out

    def save(self, path):
        np.save(path, self.params(), allow_pickle=False)

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))


class RadianceCache:
    """The MLP radiance cache over the C ABI (mrt_cache_*)."""

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        _check(lib().mrt_cache_create(ctypes.byref(h), int(seed)))
        self.h = h

    def close(self):
        if self.h:
            lib().mrt_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bounds(self, lo, hi):
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        _check(lib().mrt_cache_set_bounds(self.h, _ptr(lo), _ptr(hi)))

    def set_adam(self, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8):
        _check(lib().mrt_cache_set_adam(self.h, lr, beta1, beta2, eps))

    def params(self):
        out = np.zeros(CACHE_PARAMS, np.float32)
        _check(lib().mrt_cache_get_params(self.h, _ptr(out), out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().mrt_cache_set_params(self.h, _ptr(p), p.size))

    def infer(self, queries):
        q = np.ascontiguousarray(queries, QUERY_DTYPE)
        out = np.zeros((q.size, 3), np.float32)
        _check(lib().mrt_cache_infer(self.h, _ptr(q), q.size, _ptr(out)))
        return out

    def train(self, queries, targets):
        q = np.ascontiguousarray(queries, QUERY_DTYPE)
        t = np.ascontiguousarray(targets, np.float32)
        loss = ctypes.c_float()
        _check(lib().mrt_cache_train(self.h, _ptr(q), _ptr(t), q.size, ctypes.byref(loss)))
        return loss.value

    def gradient(self, queries, targets):
        q = np.ascontiguousarray(queries, QUERY_DTYPE)
        t = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(CACHE_PARAMS, np.float32)
        loss = ctypes.c_float()
        _check(lib().mrt_cache_gradient(self.h, _ptr(q), _ptr(t), q.size, _ptr(g), ctypes.byref(loss)))
        return g, loss.value
