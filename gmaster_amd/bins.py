"""Bandpowers (``pymaster.NmtBin``), restating ``gmaster/bins.py:15-302``.

A band ``b`` holds multipoles with weights normalised to one per band; the binned
spectrum is ``C_b = sum_{l in b} w_l f_l C_l``.  Band membership is held in CSR form
(multipoles sorted by band, ``band_start``) on the GPU for the binning kernels.
"""

import numpy as np
import torch

from . import _lib
from ._array import device_tensor, wrap


def _linear_bands(lmax, nlb):
    """Bands of ``nlb`` from l = 2; l < 2 and an incomplete last band get -1 (``gmaster/bins.py:30-37``)."""
    ells = np.arange(lmax + 1, dtype=np.int32)
    bpws = ((ells - 2) // nlb).astype(np.int32)
    bpws[:2] = -1
    if np.sum(bpws == bpws[-1]) != nlb:
        bpws[bpws == bpws[-1]] = -1
    return ells, bpws


class NmtBin:
    """Bandpower definition; keyword-only like pymaster."""

    def __init__(self, *, bpws, ells, lmax=None, weights=None, f_ell=None):
        ells = np.asarray(ells, dtype=np.int32)
        bpws = np.asarray(bpws, dtype=np.int32)
        if ells.ndim != 1 or bpws.shape != ells.shape or not len(ells):
            raise ValueError("ells and bpws must be non-empty 1-D arrays of equal size")
        self.lmax = int(np.max(ells) if lmax is None else lmax)
        weights = np.ones(len(ells)) if weights is None else np.asarray(weights, dtype=np.float64)
        f_ell = np.ones(len(ells)) if f_ell is None else np.asarray(f_ell, dtype=np.float64)
        if weights.shape != ells.shape or f_ell.shape != ells.shape:
            raise ValueError("weights and f_ell must have the same shape as ells")
        keep = (ells <= self.lmax) & (bpws >= 0)
        ells, bpws = ells[keep], bpws[keep]
        weights, f_ell = weights[keep].astype(np.float64), f_ell[keep].astype(np.float64)
        n_bands = int(np.max(bpws, initial=0)) + 1
        totals = np.bincount(bpws, weights=weights, minlength=n_bands)
        if np.any(totals <= 0):
            raise RuntimeError(f"Weights in band {int(np.flatnonzero(totals <= 0)[0])} are wrong")
        weights = weights / totals[bpws]
        f_ell = np.where(f_ell > 0, f_ell, 1.0)
        order = np.argsort(bpws, kind="stable")
        self.n_bands = n_bands
        self.ell_max = self.lmax
        self._ells_np, self._bpws_np = ells[order], bpws[order]
        self._weights_np, self._f_ell_np = weights[order], f_ell[order]
        self._band_start_np = np.concatenate([[0], np.cumsum(np.bincount(self._bpws_np, minlength=n_bands))]).astype(np.int32)
        self._dev = {}

    # ------------------------------------------------------------------ constructors
    @classmethod
    def from_nside_linear(cls, nside, nlb, is_Dell=False, f_ell=None):
        return cls.from_lmax_linear(3 * nside - 1, nlb, is_Dell, f_ell)

    @classmethod
    def from_lmax_linear(cls, lmax, nlb, is_Dell=False, f_ell=None):
        ells, bpws = _linear_bands(lmax, nlb)
        if f_ell is None and is_Dell:
            f_ell = ells * (ells + 1.0) / (2 * np.pi)
        return cls(bpws=bpws, ells=ells, lmax=lmax, f_ell=f_ell)

    @classmethod
    def from_edges(cls, ell_ini, ell_end, is_Dell=False, f_ell=None):
        ell_ini = np.asarray(ell_ini, dtype=np.int32)
        ell_end = np.asarray(ell_end, dtype=np.int32)
        ells, bpws = [], []
        for b, (lo, hi) in enumerate(zip(ell_ini, ell_end)):
            ells.extend(range(lo, hi))
            bpws.extend([b] * (hi - lo))
        ells = np.asarray(ells, dtype=np.int32)
        bpws = np.asarray(bpws, dtype=np.int32)
        if f_ell is None and is_Dell:
            f_ell = ells * (ells + 1.0) / (2 * np.pi)
        elif f_ell is not None:
            f_ell = np.asarray(f_ell)[ells]
        return cls(bpws=bpws, ells=ells, lmax=int(ells.max()), f_ell=f_ell)

    # ------------------------------------------------------------------ accessors
    def get_n_bands(self):
        return self.n_bands

    def get_nell_list(self):
        return np.diff(self._band_start_np)

    def get_ell_list(self, b):
        return self._ells_np[self._bpws_np == int(b)].copy()

    def get_ell_min(self, b):
        return self.get_ell_list(b)[0]

    def get_ell_max(self, b):
        return self.get_ell_list(b)[-1]

    def get_weight_list(self, b):
        return self._weights_np[self._bpws_np == int(b)].copy()

    def get_fell_list(self, b):
        return self._f_ell_np[self._bpws_np == int(b)].copy()

    def get_effective_ells(self):
        return np.bincount(self._bpws_np, weights=self._ells_np * self._weights_np, minlength=self.n_bands)

    # ------------------------------------------------------------------ device views
    def _device(self, device):
        key = str(device)
        d = self._dev.get(key)
        if d is None:
            d = {
                "ells": torch.as_tensor(self._ells_np, dtype=torch.int32, device=device),
                "start": torch.as_tensor(self._band_start_np, dtype=torch.int32, device=device),
                "w": torch.as_tensor(self._weights_np * self._f_ell_np, dtype=torch.float64, device=device),
                "inv_f": torch.as_tensor(1.0 / self._f_ell_np, dtype=torch.float64, device=device),
            }
            self._dev[key] = d
        return d

    def bin_cell(self, cls_in):
        """``(n_cls, lmax+1) -> (n_cls, n_bands)`` (1-D in, 1-D out)."""
        cls = device_tensor(cls_in)
        oned = cls.ndim != 2
        if oned:
            cls = cls[None]
        if cls.ndim > 2 or cls.shape[1] != self.lmax + 1:
            raise ValueError("Input Cl has wrong size")
        cls = cls.contiguous()
        d = self._device(cls.device)
        out = torch.empty((cls.shape[0], self.n_bands), dtype=torch.float64, device=cls.device)
        _lib.call("gm_bin_cell", _lib.ptr(cls), cls.shape[0], cls.shape[1], _lib.ptr(d["ells"]), _lib.ptr(d["start"]),
                  self.n_bands, _lib.ptr(d["w"]), _lib.ptr(out), _lib.stream_ptr(cls.device))
        return wrap(out[0] if oned else out)

    def unbin_cell(self, cls_in):
        """``(n_cls, n_bands) -> (n_cls, lmax+1)``: each band's value / f_l on its multipoles."""
        cls = device_tensor(cls_in)
        oned = cls.ndim != 2
        if oned:
            cls = cls[None]
        if cls.ndim > 2 or cls.shape[1] != self.n_bands:
            raise ValueError("Input Cl has wrong size")
        d = self._device(cls.device)
        out = torch.zeros((cls.shape[0], self.lmax + 1), dtype=torch.float64, device=cls.device)
        bp = torch.as_tensor(self._bpws_np, dtype=torch.long, device=cls.device)
        out[:, d["ells"].long()] = cls[:, bp] * d["inv_f"]
        return wrap(out[0] if oned else out)
