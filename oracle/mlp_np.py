"""NumPy restatement of the warm-start network (test infrastructure only).

The warm start has no counterpart in the reference: its anchor is the
default starting point of Fisher scoring (gp.py:169-177), which stays the
fallback.  This module is the checker for the HIP network kernels: the same
architecture and arithmetic definitions in float64, with every matrix-product
operand rounded to bfloat16 (round to nearest even) as the matrix cores read
it, so the only difference from the GPU is float32 accumulation order.

Architecture: x (n_in) -> [Linear -> GELU] x n_layers (width n_hidden) ->
Linear (n_out).  GELU is the exact form 0.5 z (1 + erf(z / sqrt 2)).
Parameters are one flat float32 vector: for each layer, W (fan_in x fan_out,
row-major) then b (fan_out).
Loss: masked mean squared error  sum(mask (out - t)^2) / sum(mask).

Pins (tests/test_mlp_oracle.py): central finite differences of the float64
loss (with bf16 rounding switched off), and bf16 rounding against its bit
definition.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

SQRT2 = math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even); returned as
    float64 holding the bfloat16 value exactly."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_shapes(n_in, n_hidden, n_layers, n_out):
    dims = [n_in] + [n_hidden] * n_layers + [n_out]
    return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]


def n_params(n_in, n_hidden, n_layers, n_out):
    return sum(a * b + b for a, b in layer_shapes(n_in, n_hidden, n_layers, n_out))


def unpack(params, shapes):
    out, o = [], 0
    for a, b in shapes:
        W = params[o:o + a * b].reshape(a, b)
        o += a * b
        bb = params[o:o + b]
        o += b
        out.append((W, bb))
    return out


def init_params(n_in, n_hidden, n_layers, n_out, seed=0):
    """He-style uniform init (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in layer_shapes(n_in, n_hidden, n_layers, n_out):
        lim = math.sqrt(6.0 / a) if b != n_out else math.sqrt(1.0 / a)
        parts.append(rng.uniform(-lim, lim, a * b))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / SQRT2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / SQRT2)) + z * np.exp(-0.5 * z * z) * INV_SQRT_2PI


def _mm(A, B, bf16):
    if bf16:
        return bf16_round(A) @ bf16_round(B)
    return np.asarray(A, np.float64) @ np.asarray(B, np.float64)


def forward(params, X, shapes, bf16=True):
    """Returns (out, cache) with cache = [(input, z)] per layer."""
    layers = unpack(np.asarray(params, np.float64), shapes)
    h = np.asarray(X, np.float64)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = _mm(h, W, bf16) + b
        cache.append((h, z))
        h = gelu(z) if li < len(layers) - 1 else z
    return h, cache


def loss_and_grad(params, X, T, M, shapes, bf16=True):
    """Masked MSE and its gradient with respect to the flat parameters."""
    out, cache = forward(params, X, shapes, bf16)
    msum = max(float(np.sum(M)), 1.0)
    r = (out - T) * M
    loss = float(np.sum(r * (out - T)) / msum)
    dz = 2.0 * r / msum
    layers = unpack(np.asarray(params, np.float64), shapes)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        h_in, z = cache[li]
        W, _ = layers[li]
        if li < len(layers) - 1:
            dz = dz * gelu_grad(z)
        gW = _mm(h_in.T, dz, bf16)
        gb = dz.sum(0)
        grads.append((gW, gb))
        if li > 0:
            dz = _mm(dz, W.T, bf16)     # gradient w.r.t. the previous activation
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam_step(params, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
    """One Adam(W) update in float32 (t counts from 1)."""
    g = grad.astype(np.float32)
    m[:] = beta1 * m + (1 - beta1) * g
    v[:] = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    params[:] = params - lr * (mh / (np.sqrt(vh) + eps) + wd * params)
    return params
