"""Dense network of the learned coupling surrogate (oracle; test infrastructure only).

A CPU restatement of what ``gm_mlp_fwd`` / ``gm_mlp_bwd`` / ``gm_adamw`` compute, with the
same bf16 rounding points (round to nearest even from FP32) and FP64 accumulation in place
of the kernels' FP32.  The reference has no learned component; this network follows the
layout of the task's surrogate (input features, 3 hidden SiLU layers, linear output).

Pinned by: central finite differences of the FP64 forward (``bf16=False``) against the
backward of every parameter and input; bf16 rounding against ``torch.bfloat16`` and hand
examples (ties to even, NaN); AdamW against its textbook update on scalars.
"""

import numpy as np


def bf16_round(x):
    """Round FP32 values to bf16 (nearest, ties to even), returned as float32."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def layer_forward(x, w, b, act, bf16=True):
    """``z = x w^T + b``, ``y = act(z)``; x, w are rounded to bf16 first when ``bf16``."""
    xr = bf16_round(x).astype(np.float64) if bf16 else np.asarray(x, np.float64)
    wr = bf16_round(w).astype(np.float64) if bf16 else np.asarray(w, np.float64)
    z = xr @ wr.T + (np.asarray(b, np.float64) if b is not None else 0.0)
    y = silu(z) if act == 1 else z
    return z, y


def layer_backward(dy, z, act, x, w, bf16=True):
    """``(dw, db, dx)`` of ``layer_forward``: dz = dy act'(z) (rounded to bf16 for the
    products when ``bf16``), dw = dz^T x, db = sum_b dz (unrounded), dx = dz w."""
    dz = np.asarray(dy, np.float64) * (silu_grad(z) if act == 1 else 1.0)
    dzr = bf16_round(dz).astype(np.float64) if bf16 else dz
    xr = bf16_round(x).astype(np.float64) if bf16 else np.asarray(x, np.float64)
    wr = bf16_round(w).astype(np.float64) if bf16 else np.asarray(w, np.float64)
    return dzr.T @ xr, dz.sum(axis=0), dzr @ wr


def forward(params, x, acts, bf16=True):
    """Whole network; returns the output and the per-layer ``(input, z)`` cache."""
    cache = []
    h = np.asarray(x, np.float64)
    for (w, b), act in zip(params, acts):
        z, y = layer_forward(h, w, b, act, bf16)
        cache.append((h, z))
        h = y
    return h, cache


def backward(params, cache, acts, dy, bf16=True):
    """Gradients ``[(dw, db)]`` per layer and the input gradient."""
    grads = [None] * len(params)
    g = np.asarray(dy, np.float64)
    for k in range(len(params) - 1, -1, -1):
        w, _ = params[k]
        x, z = cache[k]
        dw, db, g = layer_backward(g, z, acts[k], x, w, bf16)
        grads[k] = (dw, db)
    return grads, g


def adamw(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """Textbook AdamW (decoupled decay applied to the pre-step weights); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w * (1 - lr * wd) - lr * mh / (np.sqrt(vh) + eps), m, v
