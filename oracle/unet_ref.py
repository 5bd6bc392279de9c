"""FP64 restatement of the denoiser U-Net (test infrastructure only).

Architecture (the product's denoiser): 5 levels of 32/48/64/96/128 channels, two 3x3 convolutions
with bias and ReLU per encoder and decoder level, 2x2 average pooling down, nearest-neighbour
upsampling and concatenation [upsampled, skip] up, and a 1x1 output convolution to 3 channels.
Inputs are 10 channels per pixel: log1p(radiance) (3), albedo (3), normal (3), depth feature (1).
The output y is a residual in log space: z = log1p(radiance) + y, the denoised radiance is
expm1(max(z, 0)). Training loss: mean over batch, pixels and channels of |z - log1p(target)|.

Parameters are packed layer by layer in LAYERS order: W[cout][tap][cin] (tap = ky*3 + kx, cin in
concatenation order), then b[cout]. Adam (Kingma & Ba 2015) with bias correction.
"""
import numpy as np

LEVELS = (32, 48, 64, 96, 128)
IN_CH = 10
OUT_CH = 3

# (name, input channel parts in concatenation order, output channels, taps, relu)
LAYERS = []
_c_in = IN_CH
for _l, _c in enumerate(LEVELS):
    LAYERS.append(("enc%da" % _l, (_c_in,), _c, 9, True))
    LAYERS.append(("enc%db" % _l, (_c,), _c, 9, True))
    _c_in = _c
for _l in range(len(LEVELS) - 2, -1, -1):
    LAYERS.append(("dec%da" % _l, (LEVELS[_l + 1], LEVELS[_l]), LEVELS[_l], 9, True))
    LAYERS.append(("dec%db" % _l, (LEVELS[_l],), LEVELS[_l], 9, True))
LAYERS.append(("out", (LEVELS[0],), OUT_CH, 1, False))


def layer_sizes():
    return [(co * taps * sum(parts), co) for _, parts, co, taps, _ in LAYERS]


N_PARAMS = sum(w + b for w, b in layer_sizes())
OUT_SCALE = 0.1  # the output layer starts at a tenth of the He bound (initial residual near 0)


def unpack(params):
    out, k = {}, 0
    for (name, parts, co, taps, _), (nw, nb) in zip(LAYERS, layer_sizes()):
        W = np.asarray(params[k:k + nw], np.float64).reshape(co, taps, sum(parts))
        k += nw
        b = np.asarray(params[k:k + nb], np.float64)
        k += nb
        out[name] = (W, b)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([layers[name][0].ravel(), layers[name][1].ravel()])
                           for name, *_ in LAYERS])


def _splitmix_uniform(seed, n):
    """n values of the splitmix64 stream mapped to [0, 1) with 24 bits: state_i = seed + (i+1)*g."""
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) * g)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def he_uniform(seed):
    """The product's initialisation in binary32: one splitmix64 stream over all parameters in
    packing order; weights U(-s, s) with s = sqrt(6 / fan_in) (fan_in = taps * cin; times
    OUT_SCALE for the output layer); biases 0."""
    u = _splitmix_uniform(seed, N_PARAMS)
    out = np.zeros(N_PARAMS, np.float32)
    k = 0
    for (name, parts, co, taps, _), (nw, nb) in zip(LAYERS, layer_sizes()):
        s = np.float32(np.sqrt(np.float32(6.0) / np.float32(taps * sum(parts))))
        if name == "out":
            s = np.float32(s * np.float32(OUT_SCALE))
        out[k:k + nw] = (np.float32(2.0) * u[k:k + nw] - np.float32(1.0)) * s
        k += nw + nb
    return out


def prep(rgb, albedo, normal, depth):
    """The 10 input channels from per-pixel mean radiance, albedo, normal and hit distance
    (inf or 0 for a miss): log1p(max(rgb, 0)), albedo, normal, 1 / (1 + depth)."""
    d = np.asarray(depth, np.float64)
    feat = np.where(np.isfinite(d) & (d > 0), 1.0 / (1.0 + np.where(np.isfinite(d), d, 0.0)), 0.0)
    return np.concatenate([np.log1p(np.maximum(np.asarray(rgb, np.float64), 0.0)), np.asarray(albedo, np.float64),
                           np.asarray(normal, np.float64), feat[..., None]], axis=-1)


def conv(x, W, b, taps):
    """x (B, H, W, Cin), W (Cout, taps, Cin): out[y, x] = b + sum W[tap] x[y + ky - 1, x + kx - 1]."""
    if taps == 1:
        return x @ W[:, 0, :].T + b
    B, H, Wd, C = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1] = x
    out = np.broadcast_to(b, (B, H, Wd, W.shape[0])).copy()
    for t in range(9):
        ky, kx = divmod(t, 3)
        out += xp[:, ky:ky + H, kx:kx + Wd] @ W[:, t, :].T
    return out


def conv_backward(x, W, dy, taps):
    """Gradients of conv w.r.t. W, b and x."""
    B, H, Wd, C = x.shape
    if taps == 1:
        dW = np.einsum("bhwo,bhwi->oi", dy, x)[:, None, :]
        return dW, dy.sum((0, 1, 2)), dy @ W[:, 0, :]
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1] = x
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    for t in range(9):
        ky, kx = divmod(t, 3)
        dW[:, t, :] = np.einsum("bhwo,bhwi->oi", dy, xp[:, ky:ky + H, kx:kx + Wd])
        dxp[:, ky:ky + H, kx:kx + Wd] += dy @ W[:, t, :]
    return dW, dy.sum((0, 1, 2)), dxp[:, 1:-1, 1:-1]


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_backward(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(dy):
    B, H, W, C = dy.shape
    return dy.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(params, x):
    """x (B, H, W, 10) with H, W multiples of 16. Returns (y, cache) with y (B, H, W, 3)."""
    L = unpack(params) if not isinstance(params, dict) else params
    acts = {}
    h = x
    skips = []
    nlev = len(LEVELS)
    for l in range(nlev):
        if l > 0:
            h = pool(h)
        for s in "ab":
            name = "enc%d%s" % (l, s)
            acts[name + "_in"] = h
            h = np.maximum(conv(h, *L[name], 9), 0.0)
            acts[name] = h
        skips.append(h)
    for l in range(nlev - 2, -1, -1):
        h = np.concatenate([up(h), skips[l]], axis=-1)
        for s in "ab":
            name = "dec%d%s" % (l, s)
            acts[name + "_in"] = h
            h = np.maximum(conv(h, *L[name], 9), 0.0)
            acts[name] = h
    acts["out_in"] = h
    y = conv(h, *L["out"], 1)
    return y, (L, acts)


def loss_and_grad(params, x, target, loss="l1"):
    """Loss of the residual log-space prediction and its gradient w.r.t. the packed parameters.
    loss "l1" is the product's; "l2" (mean squared) is for the finite-difference check."""
    y, (L, acts) = forward(params, x)
    z = x[..., :3] + y
    t = np.log1p(np.maximum(np.asarray(target, np.float64), 0.0))
    n = z.size
    if loss == "l1":
        val = np.abs(z - t).sum() / n
        dy = np.sign(z - t) / n
    else:
        val = ((z - t) ** 2).sum() / n
        dy = 2.0 * (z - t) / n
    return val, _backward_in_order(L, acts, dy), y


def _backward_in_order(L, acts, dy):
    nlev = len(LEVELS)
    grads = {}
    dW, db, dh = conv_backward(acts["out_in"], L["out"][0], dy, 1)
    grads["out"] = (dW, db)
    dskips = [None] * nlev
    # Decoder, fine (level 0) to coarse (level nlev - 2): each decoder level's input is
    # [up(coarser output), skip].
    for l in range(0, nlev - 1):
        for s in "ba":
            name = "dec%d%s" % (l, s)
            dh = dh * (acts[name] > 0)
            dW, db, dh = conv_backward(acts[name + "_in"], L[name][0], dh, 9)
            grads[name] = (dW, db)
        cu = LEVELS[l + 1]
        dskips[l] = dh[..., cu:]
        dh = up_backward(dh[..., :cu])
    # dh: gradient of the bottleneck output enc(nlev-1)b. Encoder, coarse to fine.
    for l in range(nlev - 1, -1, -1):
        if l < nlev - 1:
            dh = dskips[l] + pool_backward(dh)
        for s in "ba":
            name = "enc%d%s" % (l, s)
            dh = dh * (acts[name] > 0)
            dW, db, dh = conv_backward(acts[name + "_in"], L[name][0], dh, 9)
            grads[name] = (dW, db)
    return pack(grads)


def denoise(params, x):
    """Denoised radiance expm1(max(log1p(rgb) + y, 0))."""
    y, _ = forward(params, x)
    return np.expm1(np.maximum(x[..., :3] + y, 0.0))


def adam(params, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
