"""FP64 restatement of the radiance-cache MLP (test infrastructure only).

encode -> 6 bias-free layers (ReLU after the first 5) -> 3 outputs; relative-L2 loss
mean over batch and channels of (y - t)^2 / (y^2 + 0.01) with y^2 treated as a constant in the
gradient; Adam (Kingma & Ba 2015) with bias correction.

`bf16=True` rounds inputs, weights and every stored activation/gradient to bfloat16 (round to
nearest even) exactly where the product kernels store them, so the GPU can be compared at a
tolerance set by fp32 accumulation order; `bf16=False` is plain binary64 (used for the finite-
difference check of the backward pass).
"""
import numpy as np

WIDTH = 64
LAYERS = 6
SHAPES = [(64, 64)] * 5 + [(3, 64)]
N_PARAMS = sum(a * b for a, b in SHAPES)
QUERY_DTYPE = np.dtype([("pos", "<f4", 3), ("roughness", "<f4"), ("dir", "<f4", 3), ("pad0", "<f4"),
                        ("nrm", "<f4", 3), ("pad1", "<f4"), ("albedo", "<f4", 3), ("pad2", "<f4")])


def to_bf16(x):
    """Round float64/float32 values to bfloat16 (RNE through binary32), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params):
    out, k = [], 0
    for r, c in SHAPES:
        out.append(np.asarray(params[k:k + r * c], np.float64).reshape(r, c))
        k += r * c
    return out


def pack(mats):
    return np.concatenate([m.ravel() for m in mats])


def _slot(s):
    """Input slot s -> (component, octave, cos, identity); components U = position in the cache
    box (3), direction (3), normal (3), roughness, albedo (3). Slots 0-7: position x octaves 2-5;
    8-15: position y 2-5; 16-23: position x, y octaves 0-1; 24-31: position z, direction x 0-1;
    32-39: direction y, z 0-1; 40-47: normal x, y 0-1; 48-55: position z 2-5; 56-59: normal z 0-1;
    60: roughness; 61-63: albedo (the kernel's order, mlp.hip slot_of)."""
    def lo_pair(c0, c1, i):
        return (c0 if i < 4 else c1, (i >> 1) & 1, i & 1, False)
    if s < 8:
        return (0, 2 + (s >> 1), s & 1, False)
    if s < 16:
        return (1, 2 + ((s - 8) >> 1), s & 1, False)
    if s < 48:
        c0, c1 = {16: (0, 1), 24: (2, 3), 32: (4, 5), 40: (6, 7)}[s - s % 8]
        return lo_pair(c0, c1, s % 8)
    if s < 56:
        return (2, 2 + ((s - 48) >> 1), s & 1, False)
    if s < 60:
        return lo_pair(8, 8, s - 56)
    return (9 + (s - 60), 0, 0, True)


SLOTS = [_slot(s) for s in range(64)]


def encode(q, lo, hi):
    """The 64 inputs of each query: position in [lo, hi] with sin/cos of 2^k*pi*x, k = 0..5 (36),
    direction and normal with k = 0, 1 (12 + 12), roughness, albedo (4), in the slot order SLOTS."""
    q = np.asarray(q)
    lo = np.asarray(lo, np.float64)
    ext = np.asarray(hi, np.float64) - lo
    p = (q["pos"].astype(np.float64) - lo) / ext
    U = np.concatenate([p, q["dir"].astype(np.float64), q["nrm"].astype(np.float64),
                        q["roughness"].astype(np.float64)[:, None], q["albedo"].astype(np.float64)], 1)
    x = np.zeros((q.size, 64))
    for s, (c, k, cos, ident) in enumerate(SLOTS):
        a = U[:, c] * (2 ** k) * np.pi
        x[:, s] = U[:, c] if ident else (np.cos(a) if cos else np.sin(a))
    return x


def forward(params, x, bf16=False):
    W = unpack(params)
    r = to_bf16 if bf16 else (lambda v: v)
    if bf16:
        W = [to_bf16(w) for w in W]
    hs = [r(x)]
    for l in range(LAYERS):
        z = hs[-1] @ W[l].T
        if l < LAYERS - 1:
            hs.append(r(np.maximum(z, 0.0)))
        else:
            y = z
    return y, hs, W


def loss_and_grad(params, x, t, bf16=False, den=None):
    """Mean relative-L2 loss and its gradient w.r.t. the packed parameters. The denominator
    y^2 + 0.01 is a constant of the gradient (stop-gradient); pass `den` to fix it explicitly."""
    y, hs, W = forward(params, x, bf16)
    n = x.shape[0]
    if den is None:
        den = y * y + 0.01
    e = y - t
    loss = float((e * e / den).sum() / (3 * n))
    r = to_bf16 if bf16 else (lambda v: v)
    g = r(2.0 * e / den / (3 * n))
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        grads[l] = g.T @ hs[l]
        if l > 0:
            dh = g @ W[l]
            g = r(dh * (hs[l] > 0))
    return loss, pack(grads)


def adam(params, grad, m, v, step, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-8):
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def he_uniform(seed):
    """The product's initialisation: splitmix64 stream, U(-sqrt(6/64), sqrt(6/64))."""
    x = np.uint64(seed)
    out = np.zeros(N_PARAMS, np.float32)
    bound = np.float32(np.sqrt(np.float32(6.0) / np.float32(64.0)))
    mask = (1 << 64) - 1
    xi = int(seed) & mask
    for i in range(N_PARAMS):
        xi = (xi + 0x9E3779B97F4A7C15) & mask
        z = xi
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        u = np.float32(z >> 40) * np.float32(1.0 / 16777216.0)
        out[i] = (np.float32(2.0) * u - np.float32(1.0)) * bound
    del x
    return out


def random_queries(n, seed, lo=(-1, -1, -1), hi=(1, 1, 1)):
    rng = np.random.default_rng(seed)
    q = np.zeros(n, QUERY_DTYPE)
    q["pos"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    q["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    q["nrm"] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q["roughness"] = rng.uniform(0.05, 1.0, n)
    q["albedo"] = rng.uniform(0.1, 0.9, (n, 3))
    return q
