"""Dense network trained and evaluated by this package's HIP kernels (``gm_mlp_fwd``,
``gm_mlp_bwd``, ``gm_adamw``, ``gm_mse_grad``): bf16 weights and activations on the matrix
cores, FP32 accumulation, FP32 master weights.

Used by the learned coupling surrogate (``gmaster_amd.learned``).  Hidden layers use SiLU,
the last layer is linear.
"""

import numpy as np
import torch

from . import _lib

SILU, LINEAR = 1, 0


def _bf16(t):
    """bf16 bit patterns (int16 tensor) of an FP32 tensor, by the library's rounding."""
    out = torch.empty(t.shape, dtype=torch.int16, device=t.device)
    _lib.call("gm_to_bf16", _lib.ptr(t), t.numel(), _lib.ptr(out), _lib.stream_ptr(t.device))
    return out


class MLP:
    """``sizes = [n_in, h1, ..., n_out]``; He-style initialisation from ``seed``."""

    def __init__(self, sizes, seed=0, device="cuda"):
        self.sizes = [int(s) for s in sizes]
        self.device = torch.device(device)
        rng = np.random.default_rng(seed)
        self.w, self.b = [], []
        for a, c in zip(self.sizes[:-1], self.sizes[1:]):
            w = rng.standard_normal((c, a)) * np.sqrt(2.0 / a)
            self.w.append(torch.as_tensor(w, dtype=torch.float32, device=self.device).contiguous())
            self.b.append(torch.zeros(c, dtype=torch.float32, device=self.device))
        self.acts = [SILU] * (len(self.w) - 1) + [LINEAR]
        self.refresh()
        self.m = [torch.zeros_like(p) for p in self.w + self.b]
        self.v = [torch.zeros_like(p) for p in self.w + self.b]
        self.step_count = 0

    def refresh(self):
        """Recompute the bf16 weight copies from the FP32 masters."""
        self.wbf = [_bf16(w) for w in self.w]

    def state_dict(self):
        return {"sizes": list(self.sizes), "w": [w.cpu() for w in self.w], "b": [b.cpu() for b in self.b]}

    @classmethod
    def from_state(cls, state, device="cuda"):
        net = cls(state["sizes"], device=device)
        net.w = [torch.as_tensor(w, dtype=torch.float32, device=net.device).contiguous() for w in state["w"]]
        net.b = [torch.as_tensor(b, dtype=torch.float32, device=net.device).contiguous() for b in state["b"]]
        net.refresh()
        return net

    def forward(self, x, keep=False):
        """``x``: (batch, n_in) FP32 device tensor.  Returns FP32 (batch, n_out)."""
        x = x.to(self.device, torch.float32).contiguous()
        if x.ndim != 2 or x.shape[1] != self.sizes[0]:
            raise ValueError(f"input must be (batch, {self.sizes[0]})")
        B = x.shape[0]
        st = _lib.stream_ptr(self.device)
        h = _bf16(x)
        cache = []
        y = None
        for k, (wbf, b, act) in enumerate(zip(self.wbf, self.b, self.acts)):
            n_in, n_out = self.sizes[k], self.sizes[k + 1]
            last = k == len(self.wbf) - 1
            z = torch.empty((B, n_out), dtype=torch.float32, device=self.device) if keep else None
            y = torch.empty((B, n_out), dtype=torch.float32, device=self.device) if last else None
            hn = None if last else torch.empty((B, n_out), dtype=torch.int16, device=self.device)
            _lib.call("gm_mlp_fwd", _lib.ptr(h), _lib.ptr(wbf), _lib.ptr(b), B, n_in, n_out, act, _lib.ptr(z),
                      _lib.ptr(y), _lib.ptr(hn), st)
            cache.append((h, z))
            h = hn
        return (y, cache) if keep else y

    def backward(self, cache, dy):
        """Gradients ``(dw list, db list)`` for the output gradient ``dy`` (FP32)."""
        st = _lib.stream_ptr(self.device)
        B = dy.shape[0]
        dws, dbs = [None] * len(self.w), [None] * len(self.w)
        g = dy.contiguous()
        for k in range(len(self.w) - 1, -1, -1):
            n_in, n_out = self.sizes[k], self.sizes[k + 1]
            x, z = cache[k]
            dz = torch.empty((B, n_out), dtype=torch.float32, device=self.device)
            dzb = torch.empty((B, n_out), dtype=torch.int16, device=self.device)
            dw = torch.empty((n_out, n_in), dtype=torch.float32, device=self.device)
            db = torch.empty(n_out, dtype=torch.float32, device=self.device)
            dx = torch.empty((B, n_in), dtype=torch.float32, device=self.device) if k > 0 else None
            _lib.call("gm_mlp_bwd", _lib.ptr(g), _lib.ptr(z), self.acts[k], _lib.ptr(x), _lib.ptr(self.wbf[k]), B,
                      n_in, n_out, _lib.ptr(dz), _lib.ptr(dzb), _lib.ptr(dw), _lib.ptr(db), _lib.ptr(dx), None, st)
            dws[k], dbs[k] = dw, db
            g = dx
        return dws, dbs

    def adam_step(self, dws, dbs, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
        self.step_count += 1
        st = _lib.stream_ptr(self.device)
        params = self.w + self.b
        grads = list(dws) + list(dbs)
        nl = len(self.w)
        for i, (p, g) in enumerate(zip(params, grads)):
            wbf = self.wbf[i] if i < nl else None
            decay = weight_decay if i < nl else 0.0
            _lib.call("gm_adamw", _lib.ptr(p), _lib.ptr(wbf), _lib.ptr(g), _lib.ptr(self.m[i]), _lib.ptr(self.v[i]),
                      p.numel(), lr, betas[0], betas[1], eps, decay, self.step_count, st)

    def mse_grad(self, y, t, wt=None):
        """``(dy, sum of weighted squared errors)`` for the mean over weighted entries."""
        n = y.numel()
        n_eff = float(wt.sum()) if wt is not None else float(n)
        dy = torch.empty_like(y)
        sq = torch.empty_like(y)
        _lib.call("gm_mse_grad", _lib.ptr(y), _lib.ptr(t), _lib.ptr(wt), n, max(n_eff, 1.0), _lib.ptr(dy), _lib.ptr(sq),
                  _lib.stream_ptr(self.device))
        return dy, sq

    def train_step(self, x, t, wt=None, lr=1e-3, weight_decay=0.0):
        y, cache = self.forward(x, keep=True)
        dy, sq = self.mse_grad(y, t, wt)
        dws, dbs = self.backward(cache, dy)
        self.adam_step(dws, dbs, lr=lr, weight_decay=weight_decay)
        return sq
