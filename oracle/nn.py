"""Float64 restatement of the learned sinogram correction (test infrastructure).

The network has no counterpart in the reference: it is the 1-D convolutional
correction q' = q + f(q) applied along detector rows before backprojection
(f: width-9 convolutions 1 -> C -> ... -> C -> 1 channels, zero padding, GELU
between layers, GELU(x) = x Phi(x) with the exact normal CDF).  Arrays are
(rows, channels, n), n fastest, like the device layout.  Weights of layer l
are (cout, cin, 9); `round_bf16` uses them rounded to bfloat16 (RNE) as the
device kernels do.  The backward pass is written out by hand; it is pinned
against central finite differences in tests/test_oracle.py.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

K = 9
HALF = K // 2


def bf16_round(w):
    """Round float values to bfloat16 (round to nearest, ties to even)."""
    u = np.ascontiguousarray(np.asarray(w, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(w)).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def split_params(flat, layers, channels):
    """Per layer (w (cout, cin, 9), b (cout,)) from the flat parameter vector."""
    out, o = [], 0
    for l in range(layers):
        cin = 1 if l == 0 else channels
        cout = 1 if l == layers - 1 else channels
        w = np.asarray(flat[o:o + cout * cin * K], np.float64).reshape(cout, cin, K)
        o += cout * cin * K
        b = np.asarray(flat[o:o + cout], np.float64)
        o += cout
        out.append((w, b))
    assert o == len(flat)
    return out


def _windows(x):
    """(R, C, N) -> (R, C, N, K): x[r, c, n + k - 4] with zeros outside."""
    n = x.shape[2]
    xp = np.pad(x, ((0, 0), (0, 0), (HALF, HALF)))
    return np.stack([xp[:, :, k:k + n] for k in range(K)], axis=-1)


def conv(x, w, b):
    return np.einsum("rink,oik->ron", _windows(x), w) + b[None, :, None]


def forward(params, q, round_bf16=True):
    """q (R, N) -> (q + f(q) (R, N), cache)."""
    x = np.asarray(q, np.float64)[:, None, :]
    zs, inputs = [], []
    h = x
    for l, (w, b) in enumerate(params):
        wl = bf16_round(w) if round_bf16 else w
        inputs.append(h)
        z = conv(h, wl, b)
        zs.append(z)
        h = gelu(z) if l < len(params) - 1 else z
    out = x[:, 0, :] + zs[-1][:, 0, :]
    return out, (inputs, zs)


def backward(params, cache, gout, round_bf16=True):
    """Gradients [(dW, db)] of sum(gout * out) for the cached forward."""
    inputs, zs = cache
    grads = [None] * len(params)
    d = np.asarray(gout, np.float64)[:, None, :]               # dL/dz_L
    for l in range(len(params) - 1, -1, -1):
        w, _ = params[l]
        wl = bf16_round(w) if round_bf16 else w
        x = inputs[l]
        grads[l] = (np.einsum("ron,rink->oik", d, _windows(x)), d.sum(axis=(0, 2)))
        if l == 0:
            break
        n = d.shape[2]
        dxp = np.zeros((x.shape[0], x.shape[1], n + 2 * HALF))
        for k in range(K):
            dxp[:, :, k:k + n] += np.einsum("oi,ron->rin", wl[:, :, k], d)
        d = dxp[:, :, HALF:HALF + n] * gelu_grad(zs[l - 1])
    return grads


def flatten(grads):
    return np.concatenate([np.concatenate([dw.ravel(), db.ravel()]) for dw, db in grads])
