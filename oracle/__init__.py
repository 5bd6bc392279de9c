"""CPU oracle (ctypes over liboracle.so). TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the checker or
the timed reference baseline; the product (myrt_amd) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MYRT_ORACLE_LIBRARY") or os.path.join(HERE, "liboracle.so")

AABB_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("centroid", "<f4", 3), ("has_centroid", "<u4")])
NODE_DTYPE = np.dtype([("min", "<f4", 3), ("type_and_parent", "<u4"), ("max", "<f4", 3), ("first", "<u4"),
                       ("second", "<u4"), ("pad", "<u4", 3)])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("b1", "<f4"), ("b2", "<f4")])


class _Scene(ctypes.Structure):
    _fields_ = [("idx", ctypes.c_void_p), ("pos", ctypes.c_void_p), ("geo_idx_off", ctypes.c_void_p),
                ("geo_tri_cnt", ctypes.c_void_p), ("geo_vtx_off", ctypes.c_void_p), ("n_geoms", ctypes.c_uint32),
                ("inst_geom", ctypes.c_void_p), ("inst_inv", ctypes.c_void_p), ("n_inst", ctypes.c_uint32)]


class _PtScene(ctypes.Structure):
    _fields_ = [("geo", ctypes.POINTER(_Scene)), ("nrm", ctypes.c_void_p), ("geo_has_nrm", ctypes.c_void_p),
                ("inst_material", ctypes.c_void_p), ("materials", ctypes.c_void_p), ("n_materials", ctypes.c_uint32),
                ("lights", ctypes.c_void_p), ("n_lights", ctypes.c_uint32), ("env_scale", ctypes.c_double),
                ("sdf_tape", ctypes.c_void_p), ("sdf_drawables", ctypes.c_void_p), ("n_sdf", ctypes.c_uint32),
                ("sdf_steps", ctypes.c_int), ("sdf_eps", ctypes.c_float), ("dof", ctypes.c_int),
                ("lens", ctypes.c_float * 2), ("focus", ctypes.c_float), ("bokeh", ctypes.c_void_p),
                ("bokeh_w", ctypes.c_int), ("bokeh_h", ctypes.c_int), ("uv", ctypes.c_void_p),
                ("texels", ctypes.c_void_p), ("tex_info", ctypes.c_void_p), ("n_textures", ctypes.c_uint32),
                ("cubemap", ctypes.c_void_p), ("cube_size", ctypes.c_int)]


MATERIAL_DTYPE = np.dtype([("albedo_rgba8", "<u4"), ("ior", "<f4"), ("roughness", "<f4"), ("metallic", "<f4"),
                           ("transmission", "<f4"), ("texture", "<u4"), ("pad", "<u4", 2)])
LIGHT_DTYPE = np.dtype([("corner", "<f4", 3), ("u", "<f4", 3), ("v", "<f4", 3), ("emission", "<f4", 3)])
# The product's material when none was pushed.
DEFAULT_MATERIAL = np.array([(0xffcccccc, 1.5, 1.0, 0.0, 0.0, 0, (0, 0))], MATERIAL_DTYPE)

_lib = None
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle.so is not built (run python oracle/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_triangle_bounds": ([_P, _SZ, _P, _P], None),
            "orc_bvh_build": ([_P, _SZ, ctypes.c_int, ctypes.POINTER(_P)], ctypes.c_int),
            "orc_bvh_sizes": ([_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], None),
            "orc_bvh_copy": ([_P, _P, _P, _P], None),
            "orc_bvh_destroy": ([_P], None),
            "orc_invert": ([_P, _P], ctypes.c_int),
            "orc_transform_box": ([_P, _P, _P, _P, _P], None),
            "orc_watertight": ([_P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
            "orc_moller_trumbore": ([_P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
            "orc_trace_brute": ([ctypes.POINTER(_Scene), _P, _SZ, ctypes.c_int, _P, ctypes.c_int], None),
            "orc_accel_build": ([ctypes.POINTER(_Scene), ctypes.POINTER(_P)], ctypes.c_int),
            "orc_accel_trace": ([_P, _P, _SZ, ctypes.c_int, _P, ctypes.c_int], None),
            "orc_accel_destroy": ([_P], None),
            "orc_reference_trace": ([_P, _P, _SZ, ctypes.c_int, _P], None),
            "orc_reference_time": ([_P, _P, _SZ, ctypes.c_int, ctypes.c_int, _SZ, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)], ctypes.c_double),
            "orc_sdf_map": ([_P, ctypes.c_uint32, _P, _P], ctypes.c_float),
            "orc_sdf_map_seeded": ([_P, ctypes.c_uint32, _P, ctypes.c_uint32, _P], ctypes.c_float),
            "orc_sdf_blend_seed": ([_P, _P], ctypes.c_uint32),
            "orc_sdf_trace": ([_P, _P, ctypes.c_uint32, _P, _SZ, ctypes.c_int, ctypes.c_int, ctypes.c_float, _P,
                               _P, ctypes.c_int], None),
            "orc_reference_render_time": ([ctypes.POINTER(_PtScene), _P, _P, _P, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)], ctypes.c_double),
            "orc_render": ([ctypes.POINTER(_PtScene), _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, _P, ctypes.c_int], ctypes.c_int),
            "orc_fetch_counts": ([_P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_uint32, _P, _SZ, ctypes.c_int, _P, _P,
                                  ctypes.c_int], ctypes.c_int),
            "orc_bsdf_eval": ([_P, _P, _P, _P, _SZ, _P, _P], ctypes.c_int),
            "orc_bsdf_sample": ([_P, _P, _P, ctypes.c_int, _P, _SZ, _P, _P, _P, _P, _P], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def triangle_bounds(indices, points):
    idx = np.ascontiguousarray(indices, np.uint32).ravel()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros(idx.size // 3, AABB_DTYPE)
    lib().orc_triangle_bounds(_p(idx), idx.size, _p(pts), _p(out))
    return out


def bvh_build(boxes, split_aabbs):
    """Literal restatement of bvh(aabbs, split): returns (nodes, indices, final_boxes)."""
    boxes = np.ascontiguousarray(boxes, AABB_DTYPE)
    h = ctypes.c_void_p()
    if lib().orc_bvh_build(_p(boxes), boxes.size, 1 if split_aabbs else 0, ctypes.byref(h)) != 0:
        raise RuntimeError("oracle bvh build failed")
    nn, ni = ctypes.c_size_t(), ctypes.c_size_t()
    lib().orc_bvh_sizes(h, ctypes.byref(nn), ctypes.byref(ni))
    nodes = np.zeros(nn.value, NODE_DTYPE)
    idx = np.zeros(ni.value, np.uint32)
    fb = np.zeros(ni.value, AABB_DTYPE)
    lib().orc_bvh_copy(h, _p(nodes), _p(idx), _p(fb))
    lib().orc_bvh_destroy(h)
    return nodes, idx, fb


def invert(m4):
    """Rows 0-2 of the inverse of a 4x4 matrix m4[r][c] (float32)."""
    col = np.ascontiguousarray(np.asarray(m4, np.float32).T).ravel()
    out = np.zeros(12, np.float32)
    if not lib().orc_invert(_p(col), _p(out)):
        raise ValueError("singular matrix")
    return out


def transform_box(bmin, bmax, m4):
    col = np.ascontiguousarray(np.asarray(m4, np.float32).T).ravel()
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().orc_transform_box(_p(np.asarray(bmin, np.float32)), _p(np.asarray(bmax, np.float32)), _p(col), _p(lo),
                            _p(hi))
    return lo, hi


def watertight(o, d, v0, v1, v2):
    args = [np.ascontiguousarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    t, b1, b2 = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    hit = lib().orc_watertight(*[_p(a) for a in args], ctypes.byref(t), ctypes.byref(b1), ctypes.byref(b2))
    return bool(hit), t.value, b1.value, b2.value


def moller_trumbore(o, d, v1, v2, v3):
    args = [np.ascontiguousarray(x, np.float32) for x in (o, d, v1, v2, v3)]
    t, bx, by = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    hit = lib().orc_moller_trumbore(*[_p(a) for a in args], ctypes.byref(t), ctypes.byref(bx), ctypes.byref(by))
    return bool(hit), t.value, bx.value, by.value


class OracleScene:
    """Flat scene arrays for the brute-force / BVH oracle queries."""

    def __init__(self, geometries, instances):
        """geometries: list of (indices, points); instances: list of (geom, 4x4 matrix)."""
        idx, pos, io, tc, vo = [], [], [], [], []
        n_idx = n_vtx = 0
        for ind, pts in geometries:
            ind = np.asarray(ind, np.uint32).ravel()
            pts = np.asarray(pts, np.float32).reshape(-1, 3)
            idx.append(ind)
            pos.append(pts)
            io.append(n_idx)
            tc.append(ind.size // 3)
            vo.append(n_vtx)
            n_idx += ind.size
            n_vtx += pts.shape[0]
        self.idx = np.ascontiguousarray(np.concatenate(idx))
        self.pos = np.ascontiguousarray(np.concatenate(pos))
        self.io = np.asarray(io, np.uint32)
        self.tc = np.asarray(tc, np.uint32)
        self.vo = np.asarray(vo, np.uint32)
        self.inst_geom = np.asarray([g for g, _ in instances], np.uint32)
        self.inst_inv = np.ascontiguousarray(np.stack([invert(m) for _, m in instances])
                                             if instances else np.zeros((1, 12), np.float32))
        self.c = _Scene(self.idx.ctypes.data, self.pos.ctypes.data, self.io.ctypes.data, self.tc.ctypes.data,
                        self.vo.ctypes.data, len(geometries), self.inst_geom.ctypes.data,
                        self.inst_inv.ctypes.data, len(instances))
        self._accel = None

    def brute(self, rays, any_hit=False, threads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.size, HIT_DTYPE)
        lib().orc_trace_brute(ctypes.byref(self.c), _p(rays), rays.size, 1 if any_hit else 0, _p(out), threads)
        return out

    def accel(self):
        if self._accel is None:
            h = ctypes.c_void_p()
            if lib().orc_accel_build(ctypes.byref(self.c), ctypes.byref(h)) != 0:
                raise RuntimeError("oracle accel build failed")
            self._accel = h
        return self._accel

    def bvh(self, rays, any_hit=False, threads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.size, HIT_DTYPE)
        lib().orc_accel_trace(self.accel(), _p(rays), rays.size, 1 if any_hit else 0, _p(out), threads)
        return out

    def reference(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.size, HIT_DTYPE)
        lib().orc_reference_trace(self.accel(), _p(rays), rays.size, 1 if any_hit else 0, _p(out))
        return out

    def reference_time(self, rays, any_hit=False, threads=0, chunk=65536, reps=5):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        used = ctypes.c_int()
        rate = lib().orc_reference_time(self.accel(), _p(rays), rays.size, 1 if any_hit else 0, threads, chunk,
                                        reps, ctypes.byref(used))
        return rate, used.value

    def __del__(self):
        try:
            if self._accel is not None:
                lib().orc_accel_destroy(self._accel)
        except Exception:
            pass


SDF_NODE_DTYPE = np.dtype([("op", "<u4"), ("material", "<u4"), ("p", "<f4", 6)])
SDF_DRAWABLE_DTYPE = np.dtype([("inv", "<f4", 12), ("box_lo", "<f4", 3), ("tape_offset", "<u4"),
                               ("box_hi", "<f4", 3), ("tape_len", "<u4")])


def sdf_map(tape, p):
    tape = np.ascontiguousarray(tape, SDF_NODE_DTYPE)
    mat = ctypes.c_uint32()
    d = lib().orc_sdf_map(_p(tape), tape.size, _p(np.ascontiguousarray(p, np.float32)), ctypes.byref(mat))
    return d, mat.value


def sdf_map_seeded(tape, p, seed):
    tape = np.ascontiguousarray(tape, SDF_NODE_DTYPE)
    mat = ctypes.c_uint32()
    d = lib().orc_sdf_map_seeded(_p(tape), tape.size, _p(np.ascontiguousarray(p, np.float32)), int(seed),
                                 ctypes.byref(mat))
    return d, mat.value


def sdf_blend_seed(o, d):
    return int(lib().orc_sdf_blend_seed(_p(np.ascontiguousarray(o, np.float32)), _p(np.ascontiguousarray(d, np.float32))))


def sdf_trace(tape, drawables, rays, hits=None, any_hit=False, steps=400, eps=1e-5, threads=0):
    """SDF march on top of `hits` (from OracleScene.brute/bvh); returns (hits, world normals)."""
    tape = np.ascontiguousarray(tape, SDF_NODE_DTYPE)
    drawables = np.ascontiguousarray(drawables, SDF_DRAWABLE_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    if hits is None:
        hits = np.zeros(rays.size, HIT_DTYPE)
        hits["t"] = np.inf
        hits["prim"] = hits["inst"] = 0xFFFFFFFF
    hits = np.ascontiguousarray(hits.copy(), HIT_DTYPE)
    normals = np.zeros((rays.size, 3), np.float32)
    lib().orc_sdf_trace(_p(tape), _p(drawables), drawables.size, _p(rays), rays.size, 1 if any_hit else 0, steps,
                        eps, _p(hits), _p(normals), threads)
    return hits, normals


class OracleRender:
    """FP64 render of a scene description with the product's sampler sequence."""

    def __init__(self, geometries, instances, materials=None, lights=None, env_scale=1.0, sdf_tape=None,
                 sdf_drawables=None, sdf_steps=400, sdf_eps=1e-5, textures=None):
        """geometries: list of (indices, points, normals or None[, uvs]); instances: (geom, 4x4,
        material); sdf_drawables: the product's prepared drawable records (mrt_scene_sdf_drawables);
        textures: list of (h, w, 3 or 4) float images (ids 1, 2, ... in order)."""
        self.scene = OracleScene([(g[0], g[1]) for g in geometries], [(g, mt) for g, mt, _ in instances])
        nrm = []
        for g in geometries:
            pts, nr = g[1], g[2]
            n = np.zeros((np.asarray(pts).reshape(-1, 3).shape[0], 3), np.float32) if nr is None else nr
            nrm.append(np.asarray(n, np.float32).reshape(-1, 3))
        self.nrm = np.ascontiguousarray(np.concatenate(nrm))
        self.has_nrm = np.array([0 if g[2] is None else 1 for g in geometries], np.uint32)
        uvs = []
        for g in geometries:
            npts = np.asarray(g[1]).reshape(-1, 3).shape[0]
            uv = g[3] if len(g) > 3 and g[3] is not None else np.zeros((npts, 2), np.float32)
            uvs.append(np.asarray(uv, np.float32).reshape(-1, 2))
        self.uv = np.ascontiguousarray(np.concatenate(uvs))
        texels, info, off = [], [], 0
        for im in textures or []:
            im = np.asarray(im, np.float32)
            rgba = np.ones(im.shape[:2] + (4,), np.float32)
            rgba[..., :im.shape[2]] = im
            texels.append(rgba.reshape(-1, 4))
            info.append((off, im.shape[1], im.shape[0], 0))
            off += im.shape[0] * im.shape[1]
        self.texels = np.ascontiguousarray(np.concatenate(texels) if texels else np.ones((1, 4), np.float32))
        self.tex_info = np.ascontiguousarray(np.array(info or [(0, 1, 1, 0)], np.uint32))
        self.n_textures = len(info)
        self.inst_mat = np.array([mi for _, _, mi in instances] or [0], np.uint32)
        mats = DEFAULT_MATERIAL if materials is None or len(materials) == 0 else np.ascontiguousarray(materials,
                                                                                                      MATERIAL_DTYPE)
        self.mats = np.ascontiguousarray(mats, MATERIAL_DTYPE)
        self.lights = np.ascontiguousarray(np.zeros(0, LIGHT_DTYPE) if lights is None else lights, LIGHT_DTYPE)
        self._lights_buf = self.lights if self.lights.size else np.zeros(1, LIGHT_DTYPE)
        self.sdf_tape = None if sdf_tape is None else np.ascontiguousarray(sdf_tape, SDF_NODE_DTYPE)
        self.sdf_dr = None if sdf_drawables is None else np.ascontiguousarray(sdf_drawables, SDF_DRAWABLE_DTYPE)
        n_sdf = 0 if self.sdf_dr is None else self.sdf_dr.size
        self.c = _PtScene(ctypes.pointer(self.scene.c), self.nrm.ctypes.data, self.has_nrm.ctypes.data,
                          self.inst_mat.ctypes.data, self.mats.ctypes.data, self.mats.size,
                          self._lights_buf.ctypes.data, self.lights.size, float(env_scale),
                          None if not n_sdf else self.sdf_tape.ctypes.data,
                          None if not n_sdf else self.sdf_dr.ctypes.data, n_sdf, int(sdf_steps), float(sdf_eps))
        self.c.uv = self.uv.ctypes.data
        self.c.texels = self.texels.ctypes.data
        self.c.tex_info = self.tex_info.ctypes.data
        self.c.n_textures = self.n_textures
        self._cube = None

    def set_cubemap(self, faces):
        """faces: (6, n, n, 3 or 4) floats (+x, -x, +y, -y, +z, -z), or None for the sky gradient."""
        self._cube = None
        self.c.cubemap, self.c.cube_size = None, 0
        if faces is not None and np.asarray(faces).shape[1] > 10:
            f = np.asarray(faces, np.float32)
            rgba = np.ones(f.shape[:3] + (4,), np.float32)
            rgba[..., :f.shape[3]] = f
            self._cube = np.ascontiguousarray(rgba)
            self.c.cubemap, self.c.cube_size = self._cube.ctypes.data, rgba.shape[1]

    def set_dof(self, enabled=True, lens_size=(10.0, 10.0), focus=10.0, bokeh_mask=None):
        """The product's thin-lens camera; mask (h, w, 3 or 4) floats, narrower than 10 = none."""
        self.c.dof = int(bool(enabled))
        self.c.lens[0], self.c.lens[1] = float(lens_size[0]), float(lens_size[1])
        self.c.focus = float(focus)
        self._bokeh = None
        self.c.bokeh, self.c.bokeh_w, self.c.bokeh_h = None, 0, 0
        if bokeh_mask is not None and np.asarray(bokeh_mask).shape[1] >= 10:
            mk = np.asarray(bokeh_mask, np.float32)
            rgba = np.ones(mk.shape[:2] + (4,), np.float32)
            rgba[..., :mk.shape[2]] = mk
            self._bokeh = np.ascontiguousarray(rgba)
            self.c.bokeh, self.c.bokeh_w, self.c.bokeh_h = self._bokeh.ctypes.data, rgba.shape[1], rgba.shape[0]

    def reference_render_time(self, view, proj, width, height, max_depth, rows=0, threads=0, reps=5):
        """CPU baseline: the reference's own plain renderer (pt_generate / pt_trace / pt_color
        restated, its two sphere lights, its stackless traversal) on rows [0, rows) at 1 spp under
        the thread_pool clone. Returns (samples/s, rays/s, threads used, radiance checksum)."""
        v = np.ascontiguousarray(np.asarray(view, np.float32).T).ravel()
        p = np.ascontiguousarray(np.asarray(proj, np.float32).T).ravel()
        used, rps, chk = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        sps = lib().orc_reference_render_time(ctypes.byref(self.c), self.scene.accel(), _p(v), _p(p), width, height,
                                              rows, max_depth, threads, reps, ctypes.byref(used), ctypes.byref(rps),
                                              ctypes.byref(chk))
        return sps, rps.value, used.value, chk.value

    def render(self, view, proj, width, height, spp, max_depth, sample_base=0, threads=0):
        """Mean radiance per pixel, shape (height, width, 3), float64 (rows from the bottom)."""
        v = np.ascontiguousarray(np.asarray(view, np.float32).T).ravel()
        p = np.ascontiguousarray(np.asarray(proj, np.float32).T).ravel()
        out = np.zeros(width * height * 3, np.float64)
        rc = lib().orc_render(ctypes.byref(self.c), self.scene.accel(), _p(v), _p(p), width, height, spp,
                              sample_base, max_depth, _p(out), threads)
        if rc < 0:
            raise RuntimeError("oracle render failed")
        self.last_nonfinite = rc  # non-finite values rejected by this render
        return out.reshape(height, width, 3) / spp


def bsdf_eval(material, n, wo, wi):
    """The renderer's material model: BSDF value f (N, 3) and sampling density pdf (N,) of the
    non-delta lobes for directions wi (N, 3), shading normal n on wo's side."""
    mat = np.ascontiguousarray(material, MATERIAL_DTYPE).reshape(1)
    n = np.ascontiguousarray(n, np.float64)
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.ascontiguousarray(wi, np.float64).reshape(-1, 3)
    f = np.zeros_like(wi)
    pdf = np.zeros(wi.shape[0])
    lib().orc_bsdf_eval(_p(mat), _p(n), _p(wo), _p(wi), wi.shape[0], _p(f), _p(pdf))
    return f, pdf


def bsdf_sample(material, n, wo, u, entering=True):
    """One renderer BSDF sample per uniform triple u (N, 3): dict(wi, weight, pdf, side, delta);
    side 0 = path ends, 1 = reflected, 2 = refracted."""
    mat = np.ascontiguousarray(material, MATERIAL_DTYPE).reshape(1)
    n = np.ascontiguousarray(n, np.float64)
    wo = np.ascontiguousarray(wo, np.float64)
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 3)
    k = u.shape[0]
    wi, w = np.zeros((k, 3)), np.zeros((k, 3))
    pdf = np.zeros(k)
    side = np.zeros(k, np.int32)
    delta = np.zeros(k, np.int32)
    lib().orc_bsdf_sample(_p(mat), _p(n), _p(wo), 1 if entering else 0, _p(u), k, _p(wi), _p(w), _p(pdf), _p(side),
                          _p(delta))
    return dict(wi=wi, weight=w, pdf=pdf, side=side, delta=delta)


def fetch_counts(wide, rays, any_hit=False, threads=0):
    """Per-ray fetches of the product traversal kernel, replayed on the CPU over the product's
    device-format arrays `wide` = (blas, tris, tlas, items, instances) from Scene.wide_arrays():
    returns (counts (N, 3) uint32: wide nodes, triangle records, instance records; hits)."""
    blas, tris, tlas, items, inst = wide
    rays = np.ascontiguousarray(rays)
    n = rays.size
    counts = np.zeros((n, 3), np.uint32)
    hits = np.zeros(n, HIT_DTYPE)
    keep = [np.ascontiguousarray(a) for a in (blas, tris, tlas, items if items.size else np.zeros(1, np.uint32),
                                              inst)]
    single = 0 if inst.size == 1 else 0xFFFFFFFF
    lib().orc_fetch_counts(_p(keep[0]), _p(keep[1]), _p(keep[2]), _p(keep[3]), _p(keep[4]), 1 if tlas.size else 0,
                           single, _p(rays), n, 1 if any_hit else 0, _p(counts), _p(hits), threads)
    return counts, hits
