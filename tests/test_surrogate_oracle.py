"""The CPU reference of the surrogate MLP: its FP64 gradients are pinned by
central finite differences of its own loss, and its BF16 rounding helper by
exact cases."""
from __future__ import annotations

import numpy as np

import mlp_oracle as mo


def _params(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for i, (k, n) in enumerate(mo.SHAPES):
        fan = mo.NIN if i == 0 else k
        w = rng.normal(0, (np.sqrt(2.0) if i < 3 else 1.0) / np.sqrt(fan), (k, n))
        if i == 0:
            w[mo.NIN:] = 0.0
        parts += [w.ravel(), rng.normal(0, 0.1, n)]
    return np.concatenate(parts)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.0, 0.0], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1.0); 1 + 1.5 * 2^-8 rounds up
    assert mo.bf16(x).tolist() == [1.0, 1.0, 1.0078125, 1.0078125, -3.0, 0.0]


def test_fp64_gradients_match_finite_differences():
    rng = np.random.Generator(np.random.PCG64(5))
    p = _params(3)
    x = rng.normal(0, 1, (8, mo.NIN))
    t = rng.normal(0, 1, (8, mo.NOUT))
    _, g = mo.loss_and_grads(p, x, t, exact=True)
    offs = np.cumsum([0] + [k * n + n for k, n in mo.SHAPES])
    picks = []
    for li, (k, n) in enumerate(mo.SHAPES):
        kk = mo.NIN if li == 0 else k
        picks += [offs[li] + int(rng.integers(0, kk)) * n + int(rng.integers(0, n)) for _ in range(3)]
        picks += [offs[li] + k * n + int(rng.integers(0, n)) for _ in range(2)]
    for i in picks:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_and_grads(pp, x, t)[0] - mo.loss_and_grads(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_rounded_mode_close_to_exact():
    rng = np.random.Generator(np.random.PCG64(6))
    p = _params(4)
    x = rng.normal(0, 1, (32, mo.NIN))
    t = rng.normal(0, 1, (32, mo.NOUT))
    _, g64 = mo.loss_and_grads(p, x, t, exact=True)
    _, g16 = mo.loss_and_grads(p, x, t, exact=False)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert rel < 0.05, rel
