"""CPU reference of the IR surrogate MLP (include/mgcm_sg.h) -- test infrastructure.

Only tests/ and bench.py's checks use it; the product path never does.

Two modes:
  exact=True   plain FP64 forward/backward (pinned by finite differences in
               tests/test_surrogate_oracle.py);
  exact=False  the device kernels' rounding points: BF16 inputs, weights,
               activations, dY and dZ; FP32 pre-activations and outputs; the
               products themselves summed in FP64 (the device sums in FP32, so
               the two differ by accumulation error only).

Parameters use the flat layout of mgcm_sg.h: per layer W[k][n] then b[n];
W1 is 160 x 512 with rows 147..159 zero.
"""
from __future__ import annotations

import math

import numpy as np

NIN, KIN, NH, NOUT = 147, 160, 512, 36
SHAPES = [(KIN, NH), (NH, NH), (NH, NH), (NH, NOUT)]


def split(p: np.ndarray) -> list:
    out, o = [], 0
    for k, n in SHAPES:
        w = p[o:o + k * n].reshape(k, n)
        o += k * n
        b = p[o:o + n]
        o += n
        out.append((w, b))
    assert o == p.size
    return out


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32)


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _r(a, exact, kind):
    """Rounding point of the device path: 'b' BF16, 'f' FP32."""
    if exact:
        return np.asarray(a, dtype=np.float64)
    a32 = np.asarray(a, dtype=np.float64).astype(np.float32)
    return (bf16(a32) if kind == "b" else a32).astype(np.float64)


def forward(p: np.ndarray, x: np.ndarray, exact: bool = True) -> tuple:
    layers = split(np.asarray(p, dtype=np.float64) if exact else np.asarray(p, dtype=np.float32))
    h = np.zeros((x.shape[0], KIN))
    h[:, :NIN] = x
    h = _r(h, exact, "b")
    zs, hs = [], [h]
    for i, (w, b) in enumerate(layers):
        z = h @ _r(w, exact, "b") + np.asarray(b, dtype=np.float64)
        z = _r(z, exact, "f")
        if i < 3:
            zs.append(z)
            h = _r(gelu(z), exact, "b")
            hs.append(h)
        else:
            y = z
    return y, zs, hs


def loss_and_grads(p: np.ndarray, x: np.ndarray, t: np.ndarray, exact: bool = True) -> tuple:
    """Mean squared error over the 36 outputs and its gradient (flat layout)."""
    layers = split(np.asarray(p, dtype=np.float64) if exact else np.asarray(p, dtype=np.float32))
    y, zs, hs = forward(p, x, exact)
    B = x.shape[0]
    d = y - t
    loss = float((d * d).mean())
    dz = _r(2.0 * d / (B * NOUT), exact, "b")
    grads = []
    for i in range(3, -1, -1):
        w, _ = layers[i]
        grads.append((hs[i].T @ dz, dz.sum(axis=0)))
        if i > 0:
            dh = dz @ _r(w, exact, "b").T
            dz = _r(dh * gelu_d(zs[i - 1]), exact, "b")
    grads.reverse()
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    flat[:KIN * NH].reshape(KIN, NH)[NIN:] = 0.0
    return loss, flat
