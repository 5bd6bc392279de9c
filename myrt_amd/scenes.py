"""Seeded synthetic workloads: the fBm heightfield, icosphere fields and ray sets.

All randomness comes from numpy's PCG64 generator with the seeds named here; the noise itself is
integer-hashed value noise using only +, -, *, / so every platform produces the same vertices.
"""
import numpy as np

# ---- geometry ---------------------------------------------------------------------------------


def _hash01(ix, iz, seed):
    """Integer lattice hash -> float in [0, 1) (a multiply/xorshift mix on uint32)."""
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iz.astype(np.uint32) * np.uint32(0xD8163841))
    h = h ^ np.uint32((seed * 0x9E3779B1) & 0xFFFFFFFF)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x2C1B3C6D)
    h ^= h >> np.uint32(12)
    h *= np.uint32(0x297A2D39)
    h ^= h >> np.uint32(15)
    return (h >> np.uint32(8)).astype(np.float64) / float(1 << 24)


def value_noise(x, z, seed):
    ix = np.floor(x)
    iz = np.floor(z)
    fx = x - ix
    fz = z - iz
    ix = ix.astype(np.int64)
    iz = iz.astype(np.int64)
    sx = fx * fx * (3.0 - 2.0 * fx)
    sz = fz * fz * (3.0 - 2.0 * fz)
    a = _hash01(ix, iz, seed)
    b = _hash01(ix + 1, iz, seed)
    c = _hash01(ix, iz + 1, seed)
    d = _hash01(ix + 1, iz + 1, seed)
    return (a + (b - a) * sx) * (1.0 - sz) + (c + (d - c) * sx) * sz


def fbm(x, z, seed, octaves=6, lacunarity=2.0, gain=0.5):
    total = np.zeros_like(x)
    amp, freq = 1.0, 1.0
    for o in range(octaves):
        total += amp * (value_noise(x * freq, z * freq, seed + o) * 2.0 - 1.0)
        amp *= gain
        freq *= lacunarity
    return total


def heightfield(nx=1000, nz=500, seed=3, extent=(20.0, 10.0), height=1.5, freq=0.35):
    """nx*nz cells, two triangles each, on (nx+1)*(nz+1) vertices over x in [-ex/2, ex/2],
    z in [-ez/2, ez/2] with y = height * fBm(6 octaves, lacunarity 2, gain 0.5)."""
    ex, ez = extent
    xs = np.linspace(-ex / 2, ex / 2, nx + 1)
    zs = np.linspace(-ez / 2, ez / 2, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="xy")  # shape (nz+1, nx+1)
    Y = height * fbm(X * freq, Z * freq, seed)
    pts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    # Normals from central differences of the height grid.
    gy, gx = np.gradient(Y, zs, xs)
    n = np.stack([-gx, np.ones_like(Y), -gy], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nrm = n.reshape(-1, 3).astype(np.float32)
    i = np.arange(nx)[None, :]
    j = np.arange(nz)[:, None]
    v00 = j * (nx + 1) + i
    v10 = v00 + 1
    v01 = v00 + (nx + 1)
    v11 = v01 + 1
    t0 = np.stack([v00, v01, v11], axis=-1)
    t1 = np.stack([v00, v11, v10], axis=-1)
    idx = np.stack([t0, t1], axis=2).reshape(-1, 3).astype(np.uint32)
    uv = np.stack([(X - X.min()) / ex, (Z - Z.min()) / ez], axis=-1).reshape(-1, 2).astype(np.float32)
    return idx, pts, nrm, uv


def icosphere(level):
    """Unit icosphere with 20 * 4**level faces (level 2: 320, level 3: 1280)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    pts = np.array(verts, np.float32)
    return np.array(faces, np.uint32), pts, pts.copy()


def translate_scale(center, radius):
    m = np.eye(4, dtype=np.float32)
    m[0, 0] = m[1, 1] = m[2, 2] = radius
    m[:3, 3] = center
    return m


def sphere_field(n, level, seed=1, radius=(0.1, 0.5), box=5.0):
    """n instances of one icosphere: radius U[r0, r1], centre U[-box, box]^3."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(radius[0], radius[1], n)
    centers = rng.uniform(-box, box, (n, 3))
    return icosphere(level), [translate_scale(c, r) for c, r in zip(centers, radii)]


def ground_quad(size=6.0, y=-5.0):
    pts = np.array([[-size, y, -size], [size, y, -size], [size, y, size], [-size, y, size]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    return idx, pts, nrm


def textured_ground(size=8.0, y=-5.0, tiles=4.0):
    """ground_quad with uvs running 0..tiles across (for textures that wrap)."""
    idx, pts, nrm = ground_quad(size, y)
    uv = np.array([[0, 0], [tiles, 0], [tiles, tiles], [0, tiles]], np.float32)
    return idx, pts, nrm, uv


def checker_texture(n=64, cells=8, a=(0.9, 0.85, 0.8), b=(0.15, 0.2, 0.3)):
    """An n x n rgb checkerboard of cells x cells squares."""
    ys, xs = np.mgrid[0:n, 0:n]
    on = ((xs * cells // n) + (ys * cells // n)) % 2 == 0
    img = np.where(on[..., None], np.asarray(a, np.float32), np.asarray(b, np.float32))
    return img.astype(np.float32)


def gradient_cubemap(n=32):
    """A six-face cube map: each face a distinct tint times a vertical ramp (row index)."""
    tints = np.array([[1.0, 0.3, 0.2], [0.2, 1.0, 0.3], [0.3, 0.2, 1.0], [1.0, 1.0, 0.2], [0.2, 1.0, 1.0],
                      [1.0, 0.2, 1.0]], np.float32)
    ramp = (np.arange(n, dtype=np.float32) + 0.5) / n
    return (tints[:, None, None, :] * (0.5 + ramp)[None, :, None, None]).repeat(n, axis=2).astype(np.float32)


# ---- rays -------------------------------------------------------------------------------------

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])


def look_at_basis(eye, target, up=(0.0, 1.0, 0.0)):
    f = np.asarray(target, np.float64) - np.asarray(eye, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return f, r, u


def primary_rays(width, height, spp, seed, eye=(0.0, 12.0, -7.0), target=(0.0, -1.0, 0.8), fov_deg=30.0):
    """Pinhole camera, `spp` uniformly jittered samples per pixel, pixel-major order."""
    rng = np.random.default_rng(seed)
    f, r, u = look_at_basis(eye, target)
    tan = np.tan(np.radians(fov_deg) / 2)
    aspect = width / height
    px = np.repeat(np.arange(width * height), spp)
    jx = rng.random(px.size)
    jy = rng.random(px.size)
    x = ((px % width) + jx) / width * 2 - 1
    y = 1 - ((px // width) + jy) / height * 2
    d = f[None, :] + (x * tan * aspect)[:, None] * r[None, :] + (y * tan)[:, None] * u[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(px.size, RAY_DTYPE)
    rays["o"] = np.asarray(eye, np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def _onb(n):
    """Orthonormal basis around unit normals n (Duff et al. 2017)."""
    s = np.where(n[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1 + s * n[:, 0] ** 2 * a, s * b, -s * n[:, 0]], axis=1)
    bt = np.stack([b, s + n[:, 1] ** 2 * a, -n[:, 1]], axis=1)
    return t, bt


def bounce_rays(idx, pts, count, seed, offset=1e-4):
    """Cosine-distributed rays about the upward geometric normal at uniformly area-sampled points."""
    rng = np.random.default_rng(seed)
    p0, p1, p2 = (pts[idx[:, k]].astype(np.float64) for k in range(3))
    cr = np.cross(p1 - p0, p2 - p0)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    cdf = np.cumsum(area)
    tri = np.searchsorted(cdf, rng.random(count) * cdf[-1], side="right")
    tri = np.minimum(tri, idx.shape[0] - 1)
    r1 = np.sqrt(rng.random(count))
    r2 = rng.random(count)
    b0 = 1 - r1
    b1 = r1 * (1 - r2)
    b2 = r1 * r2
    p = b0[:, None] * p0[tri] + b1[:, None] * p1[tri] + b2[:, None] * p2[tri]
    n = cr[tri] / np.linalg.norm(cr[tri], axis=1, keepdims=True)
    n = np.where(n[:, 1:2] < 0, -n, n)
    u1 = rng.random(count)
    u2 = rng.random(count)
    rad = np.sqrt(u1)
    phi = 2 * np.pi * u2
    lx, ly, lz = rad * np.cos(phi), rad * np.sin(phi), np.sqrt(np.maximum(0.0, 1 - u1))
    t, bt = _onb(n)
    d = lx[:, None] * t + ly[:, None] * bt + lz[:, None] * n
    rays = np.zeros(count, RAY_DTYPE)
    rays["o"] = (p + offset * n).astype(np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def shadow_rays(rays, hits, light=(3.0, 10.0, 2.0), pull=1e-3):
    """One segment ray per closest hit toward a point light: o = hit point pulled back along the
    incoming ray, d = light - o (unnormalised), t in (1e-4, 1 - 1e-4). Misses get a zero-length ray."""
    hit = np.isfinite(hits["t"])
    t = np.where(hit, hits["t"], 0.0).astype(np.float64)
    o = rays["o"].astype(np.float64)
    d = rays["d"].astype(np.float64)
    p = o + (t * (1.0 - pull))[:, None] * d
    out = np.zeros(rays.size, RAY_DTYPE)
    out["o"] = p.astype(np.float32)
    out["d"] = np.where(hit[:, None], np.asarray(light)[None, :] - p, 0.0).astype(np.float32)
    out["tmin"] = 1e-4
    out["tmax"] = 1.0 - 1e-4
    return out


def traversal_workload(n_coherent=4194304, n_incoherent=4194304, width=1024, spp=8):
    """The 1M-triangle heightfield and its coherent + incoherent ray sets (seeds 3 and 4)."""
    idx, pts, nrm, uv = heightfield(seed=3)
    height = max(1, n_coherent // (width * spp))
    coh = primary_rays(width, height, spp, seed=5)[:n_coherent]
    inc = bounce_rays(idx, pts, n_incoherent, seed=4)
    return (idx, pts, nrm, uv), coh, inc


# ---- path-tracing scenes ----------------------------------------------------------------------

MATERIAL_DTYPE = np.dtype([("albedo_rgba8", "<u4"), ("ior", "<f4"), ("roughness", "<f4"), ("metallic", "<f4"),
                           ("transmission", "<f4"), ("texture", "<u4"), ("pad", "<u4", 2)])
LIGHT_DTYPE = np.dtype([("corner", "<f4", 3), ("u", "<f4", 3), ("v", "<f4", 3), ("emission", "<f4", 3)])


def _rgba8(rgb):
    c = [int(round(min(max(float(x), 0.0), 1.0) * 255)) for x in list(rgb) + [1.0]]
    return c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24)


def sphere_field_scene(n_spheres=200, level=2, seed=1, glass_every=0, metal_every=0, ground=True):
    """Ground quad plus n icospheres (radius U[0.1, 0.5], centre U[-5, 5]^3, seed), each with its
    own material: GGX roughness U[0.05, 1], albedo U[0.1, 0.9] (seed + 1); every `glass_every`-th
    sphere is a dielectric (transmission 1) and every `metal_every`-th a metal. One quad light
    above the field. Returns dict(geometries, instances, materials, lights, view, proj)."""
    (fi, fp, fn), mats = sphere_field(n_spheres, level, seed=seed)
    rng = np.random.default_rng(seed + 1)
    materials = np.zeros(n_spheres + 1, MATERIAL_DTYPE)
    for k in range(n_spheres):
        albedo = rng.uniform(0.1, 0.9, 3)
        rough = rng.uniform(0.05, 1.0)
        materials[k] = (_rgba8(albedo), 1.5, rough, 0.0, 0.0, 0, (0, 0))
        if glass_every and k % glass_every == 0:
            materials[k]["transmission"] = 1.0
            materials[k]["roughness"] = 0.05
        elif metal_every and k % metal_every == 1:
            materials[k]["metallic"] = 1.0
    materials[n_spheres] = (_rgba8((0.7, 0.7, 0.7)), 1.5, 0.9, 0.0, 0.0, 0, (0, 0))
    geometries = [(fi, fp, fn)]
    instances = [(0, mt, k) for k, mt in enumerate(mats)]
    if ground:
        gi, gp, gn = ground_quad(size=8.0, y=-5.0)
        geometries.append((gi, gp, gn))
        instances.append((1, np.eye(4, dtype=np.float32), n_spheres))
    lights = np.zeros(1, LIGHT_DTYPE)
    lights[0] = ((-1.5, 7.0, -1.5), (3.0, 0.0, 0.0), (0.0, 0.0, 3.0), (12.0, 11.0, 10.0))
    # cross(u, v) = (0, 0, 3) x ... -> u = +x, v = +z gives cross = -y: emits downward.
    return dict(geometries=geometries, instances=instances, materials=materials, lights=lights,
                eye=(0.0, 4.0, -14.0), target=(0.0, -1.0, 0.0), fovy=45.0)


def wavefront_scene(n_spheres=2000, level=3, seed=7, ground_cells=(1000, 750), n_sdf=64):
    """The 4M-triangle wavefront scene: n_spheres instanced icospheres (level 3: 1280 faces; radius
    U[0.1, 0.5], centre U[-5, 5]^3) over a displaced ground (fBm heightfield of ground_cells,
    24 x 18 units, 1.2 high, at y = -5.5) and an SDF cluster of n_sdf spheres and boxes joined by
    smooth-min k = 0.1, scaled 1.5x at the origin. Materials: GGX (roughness U[0.05, 1], albedo
    U[0.1, 0.9]); every 6th sphere a dielectric, every 5th a metal, every 7th Lambert (roughness 1).
    One quad light above the field and the sky."""
    (fi, fp, fn), mats = sphere_field(n_spheres, level, seed=seed)
    rng = np.random.default_rng(seed + 1)
    materials = np.zeros(n_spheres + 5, MATERIAL_DTYPE)
    for k in range(n_spheres):
        albedo = rng.uniform(0.1, 0.9, 3)
        materials[k] = (_rgba8(albedo), 1.5, rng.uniform(0.05, 1.0), 0.0, 0.0, 0, (0, 0))
        if k % 6 == 0:
            materials[k]["transmission"] = 1.0
            materials[k]["roughness"] = 0.05
        elif k % 5 == 1:
            materials[k]["metallic"] = 1.0
        elif k % 7 == 2:
            materials[k]["roughness"] = 1.0
            materials[k]["ior"] = 1.0
    ground_mat = n_spheres
    materials[ground_mat] = (_rgba8((0.6, 0.6, 0.55)), 1.0, 1.0, 0.0, 0.0, 0, (0, 0))
    for j, rgb in enumerate([(0.8, 0.3, 0.2), (0.2, 0.6, 0.8), (0.9, 0.9, 0.9), (0.3, 0.8, 0.3)]):
        materials[n_spheres + 1 + j] = (_rgba8(rgb), 1.5, 0.3 + 0.2 * j, 0.0, 0.0, 0, (0, 0))
    gi, gp, gn, _ = heightfield(ground_cells[0], ground_cells[1], seed=seed + 2, extent=(24.0, 18.0), height=1.2)
    gp = gp + np.array([0.0, -5.5, 0.0], np.float32)
    geometries = [(fi, fp, fn), (gi, gp, gn)]
    instances = [(0, mt, k) for k, mt in enumerate(mats)]
    instances.append((1, np.eye(4, dtype=np.float32), ground_mat))
    tape = sdf_cluster(n_sdf, seed=seed + 3, k=0.1, n_materials=4)
    tape["material"] += np.uint32(n_spheres + 1)
    lights = np.zeros(1, LIGHT_DTYPE)
    lights[0] = ((-2.0, 8.0, -2.0), (4.0, 0.0, 0.0), (0.0, 0.0, 4.0), (10.0, 9.5, 9.0))
    return dict(geometries=geometries, instances=instances, materials=materials, lights=lights, sdf_tape=tape,
                sdf_transform=translate_scale((0.0, 0.0, 0.0), 1.5), eye=(0.0, 3.0, -15.0), target=(0.0, -1.5, 0.0),
                fovy=45.0)


def training_scene(seed):
    """A seeded random scene for denoiser training data: 20-1500 icospheres (level 2 or 3, radius
    U[0.1, 0.6], centre U[-5, 5]^3) with GGX, metal, glass and Lambert materials over a flat or
    displaced ground, optionally an SDF cluster, one quad light of random size, place and power,
    a random sky level and a camera on a random orbit. About 10K to 2M triangles."""
    rng = np.random.default_rng(10_000 + seed)
    n = int(rng.integers(20, 1500))
    level = int(rng.integers(2, 4))
    (fi, fp, fn), mats = sphere_field(n, level, seed=int(rng.integers(1 << 30)), radius=(0.1, 0.6))
    materials = np.zeros(n + 5, MATERIAL_DTYPE)
    for k in range(n):
        albedo = rng.uniform(0.05, 0.95, 3)
        materials[k] = (_rgba8(albedo), 1.5, rng.uniform(0.02, 1.0), 0.0, 0.0, 0, (0, 0))
        u = rng.random()
        if u < 0.12:
            materials[k]["transmission"] = 1.0
            materials[k]["roughness"] = rng.uniform(0.0, 0.3)
        elif u < 0.3:
            materials[k]["metallic"] = 1.0
        elif u < 0.45:
            materials[k]["roughness"] = 1.0
            materials[k]["ior"] = 1.0
    ground_mat = n
    materials[ground_mat] = (_rgba8(rng.uniform(0.2, 0.9, 3)), float(rng.choice([1.0, 1.5])),
                             rng.uniform(0.3, 1.0), 0.0, 0.0, 0, (0, 0))
    for j in range(4):
        materials[n + 1 + j] = (_rgba8(rng.uniform(0.1, 0.9, 3)), 1.5, rng.uniform(0.1, 1.0), 0.0, 0.0, 0, (0, 0))
    geometries = [(fi, fp, fn)]
    instances = [(0, mt, k) for k, mt in enumerate(mats)]
    if rng.random() < 0.5:
        gi, gp, gn = ground_quad(size=12.0, y=-5.5)
    else:
        cells = int(rng.integers(100, 600))
        gi, gp, gn, _ = heightfield(cells, cells, seed=int(rng.integers(1 << 30)), extent=(24.0, 24.0),
                                    height=float(rng.uniform(0.3, 2.0)))
        gp = gp + np.array([0.0, -5.5, 0.0], np.float32)
    geometries.append((gi, gp, gn))
    instances.append((1, np.eye(4, dtype=np.float32), ground_mat))
    lights = np.zeros(1, LIGHT_DTYPE)
    size = rng.uniform(1.0, 5.0)
    corner = (rng.uniform(-6, 3), rng.uniform(5, 10), rng.uniform(-6, 3))
    power = rng.uniform(4.0, 25.0) * rng.uniform(0.7, 1.0, 3)
    lights[0] = (corner, (size, 0.0, 0.0), (0.0, 0.0, size), tuple(power))
    az, el, dist = rng.uniform(0, 2 * np.pi), rng.uniform(0.05, 0.9), rng.uniform(9, 20)
    eye = (dist * np.cos(el) * np.sin(az), dist * np.sin(el) - 1.0, dist * np.cos(el) * np.cos(az))
    out = dict(geometries=geometries, instances=instances, materials=materials, lights=lights, eye=eye,
               target=tuple(rng.uniform(-2, 2, 3)), fovy=float(rng.uniform(30, 60)), env=float(rng.uniform(0.0, 1.5)))
    if rng.random() < 0.4:
        tape = sdf_cluster(int(rng.integers(8, 65)), seed=int(rng.integers(1 << 30)), k=0.1, n_materials=4)
        tape["material"] += np.uint32(n + 1)
        out["sdf_tape"] = tape
        out["sdf_transform"] = translate_scale(tuple(rng.uniform(-2, 2, 3)), float(rng.uniform(0.8, 2.0)))
    return out


def world_triangles(geometries, instances):
    """World-space triangles of every instance (n, 3, 3) float64 and, per triangle, the point its
    outward normal faces away from: the instance's origin for closed meshes (geometry 0), a point far
    below for open ones (the ground)."""
    tris, refs = [], []
    for g, mt, _ in instances:
        idx, pts = geometries[g][0], geometries[g][1]
        m = np.asarray(mt, np.float64)
        w = pts.astype(np.float64) @ m[:3, :3].T + m[:3, 3]
        t = w[idx]
        tris.append(t)
        ref = m[:3, 3] if g == 0 else np.array([0.0, -1e6, 0.0])
        refs.append(np.broadcast_to(ref, (t.shape[0], 3)))
    return np.concatenate(tris), np.concatenate(refs)


def scene_bounce_rays(tris, refs, count, seed, offset=1e-4):
    """Cosine-distributed rays about the outward normal at uniformly area-sampled points of
    world-space triangles (see world_triangles)."""
    rng = np.random.default_rng(seed)
    cr = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    cdf = np.cumsum(area)
    tri = np.minimum(np.searchsorted(cdf, rng.random(count) * cdf[-1], side="right"), tris.shape[0] - 1)
    r1 = np.sqrt(rng.random(count))
    r2 = rng.random(count)
    t = tris[tri]
    p = (1 - r1)[:, None] * t[:, 0] + (r1 * (1 - r2))[:, None] * t[:, 1] + (r1 * r2)[:, None] * t[:, 2]
    n = cr[tri] / np.linalg.norm(cr[tri], axis=1, keepdims=True)
    n = np.where((np.einsum("ij,ij->i", n, t.mean(axis=1) - refs[tri]) < 0)[:, None], -n, n)
    u1, u2 = rng.random(count), rng.random(count)
    rad, phi = np.sqrt(u1), 2 * np.pi * u2
    tx, bt = _onb(n)
    d = (rad * np.cos(phi))[:, None] * tx + (rad * np.sin(phi))[:, None] * bt + np.sqrt(np.maximum(0.0, 1 - u1))[:, None] * n
    rays = np.zeros(count, RAY_DTYPE)
    rays["o"] = (p + offset * n).astype(np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def camera_matrices(eye, target, fovy, aspect):
    e = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - e
    f /= np.linalg.norm(f)
    s = np.cross(f, (0.0, 1.0, 0.0))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ e
    t = 1.0 / np.tan(np.radians(fovy) / 2)
    near, far = 0.05, 1000.0
    proj = np.zeros((4, 4))
    proj[0, 0] = t / aspect
    proj[1, 1] = t
    proj[2, 2] = (far + near) / (near - far)
    proj[2, 3] = 2 * far * near / (near - far)
    proj[3, 2] = -1.0
    return view.astype(np.float32), proj.astype(np.float32)


# ---- SDF programs -----------------------------------------------------------------------------

SDF_NODE_DTYPE = np.dtype([("op", "<u4"), ("material", "<u4"), ("p", "<f4", 6)])
_SDF_OPS = dict(SPHERE=1, BOX=2, TORUS=3, CAPSULE=4, MENGER=5, UNION=16, SUBTRACT=17, INTERSECT=18, SMOOTH_UNION=19,
                SMOOTH_SUBTRACT=20, SMOOTH_INTERSECT=21, PUSH_TRANSLATE=32, PUSH_MIRROR=33, POP=34)


def sdf_tape(nodes):
    """Tape from (op_name, material, *params) tuples."""
    t = np.zeros(len(nodes), SDF_NODE_DTYPE)
    for i, nd in enumerate(nodes):
        t[i]["op"] = _SDF_OPS[nd[0]]
        t[i]["material"] = nd[1] if len(nd) > 1 else 0
        ps = list(nd[2:])
        t[i]["p"][:len(ps)] = ps
    return t


def sdf_cluster(n=64, seed=12, k=0.1, box=1.5, n_materials=4):
    """n spheres and boxes (alternating) at centres U[-box, box]^3, sizes U[0.15, 0.45], each
    wrapped in a translation, joined left to right by smooth_union(k)."""
    rng = np.random.default_rng(seed)
    nodes = []
    for i in range(n):
        c = rng.uniform(-box, box, 3)
        s = rng.uniform(0.15, 0.45)
        mat = int(rng.integers(0, n_materials))
        nodes.append(("PUSH_TRANSLATE", 0, *c))
        if i % 2 == 0:
            nodes.append(("SPHERE", mat, s))
        else:
            nodes.append(("BOX", mat, s, s * 0.8, s * 0.6))
        nodes.append(("POP", 0))
        if i > 0:
            nodes.append(("SMOOTH_UNION", 0, k))
    return sdf_tape(nodes)
