"""Pure-E/B spin-2 coupling blocks on the device (``gmaster/workspaces.py:928-1060``).

Purified spin-2 fields couple through the mask's spin-1 and spin-2 derivatives, so their
blocks are evaluated by Gauss-Legendre quadrature of order 2 lmax + 1 over Wigner
d-functions: the mask correlations ``c_mn`` / ``s_mn`` (``gm_wigner_d_corr``, sums over mask
multipoles up to 2 lmax with no table), the field d-tables (``gm_wigner_d_table``,
degree-major, the right-hand tables scaled by (2 l2 + 1) / 2), the left combinations with
the quadrature weights folded in (``gm_table_combine``), one FP64 GEMM per block
(``(L x N) @ (N x L)``, rocBLAS through ``torch.matmul``) and the parity split of the
spin-2 blocks (``gm_parity_split``).  The standard blocks of a purified workspace are the
exact Wigner-3j ones (``coupling_blocks``); the quadrature reproduces them to rounding.

Blocks: ``"02pure"`` (TE with the spin-2 field purified), ``"22p1"`` / ``"22m1"`` (M+ / M- with
one purified field in the slot's channels) and ``"22p2"`` / ``"22m2"`` (two).
"""

import ctypes
import hashlib

import numpy as np
import torch

from . import _lib
from ._array import device_tensor

_CACHE = {}
F64 = torch.float64


def gauss_legendre(order, device):
    """Ascending Gauss-Legendre nodes and weights (weights normalised to sum 2, as
    ``gmaster/workspaces.py:593-606`` does) on ``device``."""
    x = torch.empty(order, dtype=F64, device=device)
    w = torch.empty(order, dtype=F64, device=device)
    _lib.call("gm_gauss_legendre", int(order), _lib.ptr(x), _lib.ptr(w), _lib.stream_ptr(device))
    return x, w * (2.0 / w.sum())


def _need(t, n, what):
    """``t`` is a contiguous FP64 device tensor with at least ``n`` elements (the kernels index
    it up to n - 1)."""
    if t.dtype != F64 or not t.is_cuda or not t.is_contiguous() or t.numel() < n:
        raise ValueError(f"{what}: expected a contiguous float64 device tensor of at least {n} elements")


def wigner_d_table(x, m, n, lmax, lscale=None):
    """``(lmax + 1, len(x))`` table ``lscale[l] d^l_{mn}(beta_k)``, cos(beta_k) = x[k]."""
    _need(x, 1, "nodes")
    if lscale is not None:
        _need(lscale, lmax + 1, "lscale")
    out = torch.empty((lmax + 1, x.numel()), dtype=F64, device=x.device)
    _lib.call("gm_wigner_d_table", _lib.ptr(x), x.numel(), int(m), int(n), int(lmax), _lib.ptr(lscale), _lib.ptr(out),
              _lib.stream_ptr(x.device))
    return out


def wigner_d_corr(x, m, n, lmax, coef, signed=True):
    """``(c, s)``: ``sum_l coef[l] d^l_{mn}`` and the same with ``(-1)^l`` (None unless ``signed``)."""
    _need(x, 1, "nodes")
    coef = coef.contiguous()
    _need(coef, lmax + 1, "coef")
    c = torch.empty_like(x)
    s = torch.empty_like(x) if signed else None
    _lib.call("gm_wigner_d_corr", _lib.ptr(x), x.numel(), int(m), int(n), int(lmax), _lib.ptr(coef), _lib.ptr(c),
              _lib.ptr(s), _lib.stream_ptr(x.device))
    return c, s


def table_combine(terms, node_weight):
    """``sum_t coef_t u_t[k] v_t[l] T_t[l, k]`` times ``node_weight[k]``: terms are
    ``(coef, table, node_factor or None, degree_factor or None)``."""
    n = len(terms)
    L, nnodes = terms[0][1].shape
    for _, table, u, v in terms:
        if tuple(table.shape) != (L, nnodes):
            raise ValueError("table_combine: tables of different shapes")
        _need(table, L * nnodes, "table")
        if u is not None:
            _need(u, nnodes, "node factor")
        if v is not None:
            _need(v, L, "degree factor")
    if node_weight is not None:
        _need(node_weight, nnodes, "node weight")
    arr = lambda vals: (ctypes.c_void_p * n)(*[None if v is None else v.data_ptr() for v in vals])
    tables = arr([t[1] for t in terms])
    nodes = arr([t[2] for t in terms])
    degrees = arr([t[3] for t in terms])
    coef = (ctypes.c_double * n)(*[float(t[0]) for t in terms])
    out = torch.empty((L, nnodes), dtype=F64, device=terms[0][1].device)
    _lib.call("gm_table_combine", n, ctypes.cast(tables, ctypes.c_void_p), ctypes.cast(nodes, ctypes.c_void_p),
              ctypes.cast(degrees, ctypes.c_void_p), ctypes.cast(coef, ctypes.c_void_p), _lib.ptr(node_weight), L,
              nnodes, _lib.ptr(out), _lib.stream_ptr(out.device))
    return out


def parity_split(total, sgnd):
    L = total.shape[0]
    if tuple(total.shape) != (L, L) or tuple(sgnd.shape) != (L, L):
        raise ValueError("parity_split: expected two square matrices of the same size")
    total, sgnd = total.contiguous(), sgnd.contiguous()
    even = torch.empty_like(total)
    odd = torch.empty_like(total)
    _lib.call("gm_parity_split", _lib.ptr(total), _lib.ptr(sgnd), L, _lib.ptr(even), _lib.ptr(odd),
              _lib.stream_ptr(total.device))
    return even, odd


def pure_coupling_blocks(window, lmax):
    """``{"02pure", "22p1", "22m1", "22p2", "22m2"}`` ``(lmax+1, lmax+1)`` blocks of a mask
    window (its pseudo-C_ell; padded with zeros / cut to 2 lmax + 1).  Cached on
    (lmax, window bytes)."""
    lmax = int(lmax)
    win = device_tensor(window)[: 2 * lmax + 1]
    dev = win.device
    W = torch.zeros(2 * lmax + 1, dtype=F64, device=dev)
    W[: win.numel()] = win
    key = (lmax, str(dev), hashlib.sha1(W.cpu().numpy().tobytes()).hexdigest())
    hit = _CACHE.get(key)
    if hit is not None:
        return hit
    L, N = lmax + 1, 2 * lmax + 1
    x, wq = gauss_legendre(N, dev)
    ell = torch.arange(2 * lmax + 1, dtype=F64, device=dev)
    mf = torch.sqrt(ell * (ell + 1))
    ms = torch.sqrt(torch.clamp((ell + 2) * (ell + 1) * ell * (ell - 1), min=0))
    row = torch.arange(L, dtype=F64, device=dev)
    rf = torch.where(row > 1, 2 / torch.sqrt(torch.clamp((row + 2) * (row - 1), min=1)), torch.zeros_like(row))
    rs = torch.where(row > 1, 1 / torch.sqrt(torch.clamp((row + 2) * (row + 1) * row * (row - 1), min=1)),
                     torch.zeros_like(row))
    coef = (2 * ell + 1) * W / (4 * np.pi)
    c00, s00 = wigner_d_corr(x, 0, 0, 2 * lmax, coef)
    c10, s10 = wigner_d_corr(x, 1, 0, 2 * lmax, coef * mf)
    c20, s20 = wigner_d_corr(x, 2, 0, 2 * lmax, coef * ms)
    c11, s11 = wigner_d_corr(x, 1, 1, 2 * lmax, coef * mf * mf)
    c12, s12 = wigner_d_corr(x, 1, 2, 2 * lmax, coef * mf * ms)
    c22, s22 = wigner_d_corr(x, 2, 2, 2 * lmax, coef * ms * ms)
    c01, _ = wigner_d_corr(x, 0, 1, 2 * lmax, coef * mf, signed=False)
    c02, _ = wigner_d_corr(x, 0, 2, 2 * lmax, coef * ms, signed=False)
    half = ((2 * row + 1) / 2).contiguous()
    d = {mn: wigner_d_table(x, mn[0], mn[1], lmax) for mn in ((0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 2))}
    rf2, rfs, rs2 = rf * rf, rf * rs, rs * rs

    def block(terms, right):
        left = table_combine(terms, wq)
        out = torch.matmul(left, right.T)
        del left
        return out

    out = {}
    right = wigner_d_table(x, 0, -2, lmax, half)
    out["02pure"] = block([(1, d[(0, 2)], c00, None), (1, d[(0, 1)], c01, rf), (1, d[(0, 0)], c02, rs)], right)
    right = wigner_d_table(x, -2, -2, lmax, half)
    for name, ts in (("1", ((1, (2, 2), 0, None), (1, (1, 2), 1, rf), (1, (0, 2), 2, rs))),
                     ("2", ((1, (2, 2), 0, None), (2, (1, 2), 1, rf), (2, (0, 2), 2, rs), (1, (1, 1), 3, rf2),
                            (2, (1, 0), 4, rfs), (1, (0, 0), 5, rs2)))):
        cs = (c00, c10, c20, c11, c12, c22)
        ss = (s00, s10, s20, s11, s12, s22)
        total = block([(k, d[mn], cs[i], v) for k, mn, i, v in ts], right)
        sgnd = block([(k, d[mn], ss[i], v) for k, mn, i, v in ts], right)
        out["22p" + name], out["22m" + name] = parity_split(total, sgnd)
        del total, sgnd
    del d, right
    if len(_CACHE) >= 2:
        _CACHE.clear()
    _CACHE[key] = out
    return out


def clear_cache():
    _CACHE.clear()
