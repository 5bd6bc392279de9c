"""GPU radiance cache (bf16 MFMA, fp32 accumulation) against the FP64 oracle.

Bar: initial weights bit-equal to the oracle's initialiser; outputs within 2e-2 relative in the
L2 norm of the oracle run with the same bf16 roundings; gradients within 5e-3, both in the L2
norm and entry by entry against the layer's largest entry (fp32 summation order and bf16 rounding
flips give at most 1.5e-3 on these inputs, while a weight-gradient block that loses one tile's
contribution is off by several per cent); a few Adam steps track the oracle; the network learns a
smooth radiance field."""
import numpy as np
import pytest

import myrt_amd as m
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu
TOL = 2e-2
GRAD_TOL = 5e-3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture
def cache(ensure_built):
    c = m.RadianceCache(seed=3)
    c.set_bounds((-1, -1, -1), (1, 1, 1))
    return c


def test_initial_params_match_oracle(cache):
    assert np.array_equal(cache.params(), M.he_uniform(3))


def test_inference_matches_oracle(cache):
    q = M.random_queries(1000, 5)
    y = cache.infer(q)
    ref, _, _ = M.forward(cache.params().astype(np.float64), M.encode(q, (-1, -1, -1), (1, 1, 1)), bf16=True)
    assert rel(y, ref) < TOL


def test_inference_at_the_per_frame_size(cache):
    """2^20 + 333 queries (the per-frame inference size, every wave of the persistent grid taking
    several tile pairs, a ragged end): a random subset of 8192 against the oracle, and each query's
    result independent of the batch it came in, bit for bit (the subset inferred alone)."""
    n = (1 << 20) + 333
    q = M.random_queries(n, 15)
    y = cache.infer(q)
    assert np.isfinite(y).all()
    pick = np.sort(np.random.default_rng(16).choice(n, 8192, replace=False))
    pick[-1] = n - 1  # the last query of the ragged end
    ref, _, _ = M.forward(cache.params().astype(np.float64), M.encode(q[pick], (-1, -1, -1), (1, 1, 1)), bf16=True)
    assert rel(y[pick], ref) < TOL
    assert np.array_equal(cache.infer(q[pick]).view(np.uint32), y[pick].view(np.uint32))


@pytest.mark.parametrize("n", [777, (1 << 17) + 777])
def test_gradient_matches_oracle(cache, n):
    """777 samples: one workgroup, a ragged last tile. 2^17 + 777: every workgroup sums several
    rounds of tiles into its partial, and the last round has waves past the end."""
    q = M.random_queries(n, 6)
    t = np.random.default_rng(7).uniform(0, 2, (n, 3)).astype(np.float32)
    g, loss = cache.gradient(q, t)
    p = cache.params().astype(np.float64)
    ref_loss, ref_g = M.loss_and_grad(p, M.encode(q, (-1, -1, -1), (1, 1, 1)), t, bf16=True)
    assert abs(loss - ref_loss) / ref_loss < TOL
    k = 0
    for l, (r, c) in enumerate(M.SHAPES):
        a, b = g[k:k + r * c], ref_g[k:k + r * c]
        assert rel(a, b) < GRAD_TOL, "layer %d" % l
        assert np.abs(a - b).max() < GRAD_TOL * np.abs(b).max(), "layer %d" % l
        k += r * c


def test_adam_steps_track_oracle(cache):
    cache.set_adam(1e-3, 0.9, 0.99, 1e-8)
    q = M.random_queries(512, 8)
    t = np.random.default_rng(9).uniform(0, 2, (512, 3)).astype(np.float32)
    x = M.encode(q, (-1, -1, -1), (1, 1, 1))
    p = cache.params().astype(np.float64)
    mm = np.zeros_like(p)
    vv = np.zeros_like(p)
    for step in range(1, 4):
        cache.train(q, t)
        _, g = M.loss_and_grad(p, x, t, bf16=True)
        p, mm, vv = M.adam(p, g, mm, vv, step, lr=1e-3)
    assert rel(cache.params(), p) < 1e-3


def test_learns_smooth_field(cache):
    rng = np.random.default_rng(10)
    q = M.random_queries(1 << 14, 11)
    pos = q["pos"].astype(np.float64)
    t = np.stack([1 + 0.5 * np.sin(2 * pos[:, 0]), 1 + 0.5 * np.cos(3 * pos[:, 1]), 0.5 + 0.25 * pos[:, 2]], 1)
    t = t.astype(np.float32)
    first = cache.train(q, t)
    for _ in range(300):
        idx = rng.integers(0, q.size, 4096)
        cache.train(q[idx], t[idx])
    _, last = cache.gradient(q, t)
    assert last < first / 10, (first, last)
