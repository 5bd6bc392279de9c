"""FP64 restatement of the post-process passes (test infrastructure only): the bilateral filter of
src/pt/pp_denoise.comp:24-63, the filmic curve of src/pt/pp_tonemap_filmic.cl:1-5 and FXAA of
src/pt/pp_fxaa.comp:31-105, over (h, w, 4) images with rows in memory order."""
import numpy as np


def normpdf(x, sigma):
    return 0.39894 * np.exp(-0.5 * x * x / (sigma * sigma)) / sigma


def bilateral(img, exponent=10.0, strength=0.05):
    """pp_denoise.comp:24-63 with the range weights relative to the window's smallest colour
    distance q_min: exp(-(q - q_min) / (2 sigma_r^2)) times the spatial kernel. The reference's
    factors 0.39894 / sigma_r * bZ and exp(-q_min / (2 sigma_r^2)) are common to the window and
    cancel in the normalised sum; the shift keeps the sum non-zero where every unshifted weight
    underflows."""
    img = np.asarray(img, np.float64)
    h, w = img.shape[:2]
    k = 4
    kern = normpdf(np.arange(-k, k + 1, dtype=np.float64), exponent)
    c = img[..., :3]
    ys, xs = np.mgrid[0:h, 0:w]
    offs = [(di, dj) for di in range(-k, k + 1) for dj in range(-k, k + 1)]
    samples = [np.clip(img[np.clip(ys + dj, 0, h - 1), np.clip(xs + di, 0, w - 1), :3], 0.0, 1.0) for di, dj in offs]
    q = [((s - c) ** 2).sum(-1) for s in samples]
    qmin = np.min(q, axis=0)
    z = np.zeros((h, w))
    acc = np.zeros((h, w, 3))
    for (di, dj), s, qq in zip(offs, samples, q):
        f = np.exp(-0.5 * (qq - qmin) / (strength * strength)) * kern[dj + k] * kern[di + k]
        z += f
        acc += f[..., None] * s
    out = np.ones((h, w, 4))
    out[..., :3] = acc / z[..., None]
    return out


def filmic(x):
    X = np.maximum(0.0, np.asarray(x, np.float64) - 0.004)
    return ((X * (6.2 * X + 0.5)) / (X * (6.2 * X + 1.7) + 0.06)) ** 1.3


def tonemap(img):
    img = np.asarray(img, np.float64)
    out = np.ones(img.shape[:2] + (4,))
    out[..., :3] = filmic(img[..., :3])
    return out


def _tex(img, u, v):
    """Bilinear rgb at normalised coordinates (arrays), texel centres at (i + 0.5) / size, clamped."""
    h, w = img.shape[:2]
    fx = np.clip(u * w - 0.5, 0, w - 1)
    fy = np.clip(v * h - 0.5, 0, h - 1)
    x0, y0 = fx.astype(int), fy.astype(int)
    x1, y1 = np.minimum(x0 + 1, w - 1), np.minimum(y0 + 1, h - 1)
    tx, ty = (fx - x0)[..., None], (fy - y0)[..., None]
    return (img[y0, x0, :3] * (1 - tx) * (1 - ty) + img[y0, x1, :3] * tx * (1 - ty) +
            img[y1, x0, :3] * (1 - tx) * ty + img[y1, x1, :3] * tx * ty)


def fxaa(img):
    img = np.asarray(img, np.float64)
    h, w = img.shape[:2]
    py, px = np.mgrid[0:h, 0:w].astype(np.float64)
    ivx, ivy = 1.0 / w, 1.0 / h
    luma = np.array([0.299, 0.587, 0.114])
    nw = _tex(img, (px - 1) * ivx, (py - 1) * ivy)
    ne = _tex(img, (px + 1) * ivx, (py - 1) * ivy)
    sw = _tex(img, (px - 1) * ivx, (py + 1) * ivy)
    se = _tex(img, (px + 1) * ivx, (py + 1) * ivy)
    m = _tex(img, px * ivx, py * ivy)
    lnw, lne, lsw, lse, lm = (c @ luma for c in (nw, ne, sw, se, m))
    lmin = np.minimum(lm, np.minimum(np.minimum(lnw, lne), np.minimum(lsw, lse)))
    lmax = np.maximum(lm, np.maximum(np.maximum(lnw, lne), np.maximum(lsw, lse)))
    dx = -((lnw + lne) - (lsw + lse))
    dy = (lnw + lsw) - (lne + lse)
    red = np.maximum((lnw + lne + lsw + lse) * (0.25 / 8.0), 1.0 / 128.0)
    rcp = 1.0 / (np.minimum(np.abs(dx), np.abs(dy)) + red)
    dx = np.clip(dx * rcp, -8, 8) * ivx
    dy = np.clip(dy * rcp, -8, 8) * ivy
    u, v = px * ivx, py * ivy
    ra = 0.5 * (_tex(img, u + dx * (1 / 3 - 0.5), v + dy * (1 / 3 - 0.5)) +
                _tex(img, u + dx * (2 / 3 - 0.5), v + dy * (2 / 3 - 0.5)))
    rb = ra * 0.5 + 0.25 * (_tex(img, u - 0.5 * dx, v - 0.5 * dy) + _tex(img, u + 0.5 * dx, v + 0.5 * dy))
    lb = rb @ luma
    out = np.ones((h, w, 4))
    out[..., :3] = np.where(((lb < lmin) | (lb > lmax))[..., None], ra, rb)
    return out
