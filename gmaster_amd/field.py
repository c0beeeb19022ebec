"""Curved-sky fields (``pymaster.NmtField``), restating ``gmaster/field.py:78-376``.

A field masks its maps, subtracts the best fit of its (masked) contaminant templates
(``gmaster/field.py:193-247``: covariance and products by ``gm_map_dots``, pseudo-inverse
on the host, subtraction by ``gm_map_deproject``) and analyses the result (``n_iter``
Jacobi steps).  A spin-0 field whose mask has the same band limit and ``n_iter`` analyses
the masked map and the mask in one shared latitudinal pass (``gmaster/field.py:273-289``);
other fields compute their mask alms lazily (``:348-376``).  Spin-2 fields may be E- and/or
B-purified (``gmaster/field.py:433-499``, ``gmaster_amd.purify``); their maps are then the
purified ones.  Anisotropic masks are outside this engine and raise ``NotImplementedError``.
"""

import torch

from . import _lib, batched, sht
from ._array import device_tensor, wrap
from .deproject import deproject
from .purify import purify
from ._config import APPROXIMATE_FORMS, nmt_params
from .utils import NmtAlmInfo, NmtMapInfo

# Mask alms shared by fields built on masks with the same content (a 128-bit fingerprint of
# the mask's bytes, computed on the device by gm_fingerprint), band limit, n_iter and
# precision, so that every workspace of one mask sees bit-identical mask spectra and shares
# one coupling computation.  Keying on content rather than on the tensor's address keeps a
# new mask that the caching allocator places at a freed mask's address from hitting the
# freed mask's alms.
_MASK_ALMS = {}


def mask_fingerprint(mask):
    """128-bit content fingerprint of a contiguous device tensor (two Python ints)."""
    out = torch.empty(2, dtype=torch.int64, device=mask.device)
    _lib.call("gm_fingerprint", _lib.ptr(mask), mask.numel() * mask.element_size(), _lib.ptr(out),
              _lib.stream_ptr(mask.device))
    return tuple(int(v) for v in out.cpu())


def _mask_key(mask, lmax_mask, n_iter_mask, precision=None):
    return (mask_fingerprint(mask), mask.numel(), str(mask.device), int(lmax_mask), int(n_iter_mask),
            precision or _precision())


def _precision():
    """Transform precision of the selected form: ``"fast"`` contracts in FP32 over an FP64
    recurrence (``sht._refine``), ``"fp64"`` keeps every pass in FP64."""
    return "mixed" if nmt_params.form in APPROXIMATE_FORMS else "fp64"


def clear_mask_cache():
    _MASK_ALMS.clear()


def _remember(key, alm):
    if len(_MASK_ALMS) >= 4:
        _MASK_ALMS.clear()
    _MASK_ALMS[key] = alm


# Fields whose transforms wait for the batched pass (fast form): each entry is the field,
# its spin-0 channels ((K, npix) masked map and optionally its mask, copied) or its (Q, U)
# maps, and the mask-alms cache key when the mask rides along.  The first access to any
# pending field's alms runs every pending transform, grouped by geometry and n_iter into
# the supported (spin-0 maps, spin-2 fields) batches.
_PENDING = []


def _batchable(nside, lmax):
    return nmt_params.form in APPROXIMATE_FORMS and nmt_params.batch_fields and lmax >= 2 and nside <= 4096


def flush_pending():
    """Run every deferred field transform (called on the first access to pending alms)."""
    if not _PENDING:
        return
    todo = list(_PENDING)
    _PENDING.clear()
    groups = {}
    for item in todo:
        f = item["field"]
        key = (f.minfo.nside, f.ainfo.lmax, f.n_iter, str(f.mask.device))
        groups.setdefault(key, []).append(item)
    for (nside, lmax, n_iter, _), items in groups.items():
        # channel lists: spin-0 (field, slot) with slot "alm" or "mask"; spin-2 fields
        c0, c2 = [], []
        for it in items:
            for k, what in enumerate(it["what0"]):
                c0.append((it, k, what))
            if it["maps2"] is not None:
                c2.append(it)
        while c0 or c2:
            if c2:
                n0 = min(2, len(c0))
                n2 = 1 if n0 or len(c2) == 1 else 2
            else:
                n0 = next(k for k in (8, 4, 2, 1) if k <= len(c0))
                n2 = 0
            take0, c0 = c0[:n0], c0[n0:]
            take2, c2 = c2[:n2], c2[n2:]
            if n0 == 1 and n2 == 0:  # a lone spin-0 map: the per-spin mixed transform
                it, k, what = take0[0]
                a = sht.map2alm_s0(it["maps0"][k:k + 1], lmax, n_iter=n_iter, precision="mixed")[0]
                _deliver(it, what, a)
                continue
            m0 = torch.stack([it["maps0"][k] for it, k, _ in take0]) if take0 else None
            m2 = torch.stack([it["maps2"] for it in take2]) if take2 else None
            a0, a2 = batched.map2alm(m0, m2, lmax, n_iter=n_iter)
            for i, (it, k, what) in enumerate(take0):
                _deliver(it, what, a0[i].clone() if n0 > 1 else a0[i])
            for i, it in enumerate(take2):
                it["field"]._alm = a2[i].clone() if n2 > 1 else a2[i]
    for it in todo:
        it["field"]._pending = False


def _deliver(item, what, alm):
    f = item["field"]
    if what == "alm":
        f._alm = alm[None]
    else:
        f._alm_mask = alm
        _remember(item["mask_key"], alm)


class NmtField:
    @property
    def alm(self):
        if self._pending:
            flush_pending()
        return self._alm

    @alm.setter
    def alm(self, value):
        self._alm = value

    @property
    def alm_mask(self):
        if self._pending:
            flush_pending()
        return self._alm_mask

    @alm_mask.setter
    def alm_mask(self, value):
        self._alm_mask = value

    def __init__(self, mask, maps, *, spin=None, templates=None, beam=None, purify_e=False, purify_b=False,
                 n_iter=None, n_iter_mask=None, tol_pinv=None, wcs=None, lmax=None, lmax_mask=None,
                 masked_on_input=False, lite=False, mask_22=None, mask_12=None):
        if mask_22 is not None or mask_12 is not None:
            raise NotImplementedError("anisotropic masks are not supported by this engine")
        self._pending = False
        self.lite = bool(lite)
        self.n_iter = nmt_params.n_iter_default if n_iter is None else int(n_iter)
        self.n_iter_mask = nmt_params.n_iter_mask_default if n_iter_mask is None else int(n_iter_mask)
        self.pure_e, self.pure_b = bool(purify_e), bool(purify_b)
        self.is_catalog = False
        self.anisotropic_mask = False
        self.alm = self.alm_mask = None
        self.maps = None
        self.n_temp = 0
        self.temp = self.alm_temp = self.alphas = self.iM = None
        self.Nw = 0.0
        self.Nf = 0.0
        self.alpha = None  # data-to-randoms ratio of clustering catalogues; None for maps

        mask_t = device_tensor(mask)
        if mask_t.ndim != 1:
            raise ValueError("Mask must be a one-dimensional HEALPix map")
        self.minfo = NmtMapInfo(wcs, mask_t.shape)
        self.mask = mask_t
        lmax = self.minfo.get_lmax() if lmax is None else int(lmax)
        lmax_mask = self.minfo.get_lmax() if lmax_mask is None else int(lmax_mask)
        if lmax <= 0 or lmax_mask <= 0:
            raise ValueError("`lmax` and `lmax_mask` must be positive")
        self.ainfo = NmtAlmInfo(lmax)
        self.ainfo_mask = NmtAlmInfo(lmax_mask)
        dev = mask_t.device
        if beam is None:
            self.beam = torch.ones(lmax + 1, dtype=torch.float64, device=dev)
        else:
            b = device_tensor(beam, device=dev)
            if b.ndim != 1 or len(b) <= lmax:
                raise ValueError(f"Input beam must have at least {lmax + 1} elements given the input map resolution")
            self.beam = b[: lmax + 1]

        if maps is None:
            if templates is not None:
                raise ValueError("Templates need the field's maps")
            if spin is None:
                raise ValueError("Please supply field spin")
            self.spin = int(spin)
            self.nmaps = 1 if self.spin == 0 else 2
            return

        maps_t = device_tensor(maps, device=dev)
        if maps_t.ndim == 1:
            maps_t = maps_t[None]
        if maps_t.ndim != 2 or len(maps_t) not in (1, 2):
            raise ValueError("Must supply 1 or 2 maps per field")
        if maps_t.shape[1] != self.minfo.npix:
            raise ValueError("Incompatible map!")
        if spin is None:
            spin = 0 if len(maps_t) == 1 else 2
        if (spin == 0) != (len(maps_t) == 1):
            raise ValueError("Spin-zero fields are associated with a single map")
        self.spin = int(spin)
        self.nmaps = 1 if self.spin == 0 else 2
        pure_any = self.pure_e or self.pure_b
        # Purified fields keep FP64 mask alms in every form: their coupling weights the mask
        # spectrum by up to (L+2)(L+1)L(L-1) ~ L^4, which turns FP32-level rounding of the mask
        # alms into ~1e-4 changes of the low-l pure blocks.
        self._mask_precision = "fp64" if pure_any else _precision()
        if pure_any and self.spin != 2:
            raise ValueError("Purification is only defined for spin-2 fields")
        if pure_any and self.ainfo != self.ainfo_mask:
            raise ValueError("Pure fields require lmax == lmax_mask")
        unmasked = None
        if pure_any:
            unmasked = maps_t.clone()
            if masked_on_input:
                good = self.mask > 0
                unmasked = torch.where(good[None, :], maps_t / torch.where(good, self.mask, 1.0)[None, :], maps_t)

        masked = maps_t if masked_on_input else maps_t * self.mask[None, :]
        temps = None
        if templates is not None:
            temps = device_tensor(templates, device=dev)
            if temps.ndim != 3 or tuple(temps.shape[1:]) != tuple(maps_t.shape):
                raise ValueError("Each template must have the same shape as the maps")
            if not masked_on_input:
                temps = temps * self.mask[None, None, :]
            if masked is maps_t:
                masked = masked.clone()  # the subtraction is in place; keep the caller's maps
            tol = nmt_params.tol_pinv_default if tol_pinv is None else float(tol_pinv)
            temps_unmasked = None
            if pure_any:
                temps_unmasked = device_tensor(templates, device=dev)
                if masked_on_input:
                    good = self.mask > 0
                    temps_unmasked = torch.where(good[None, None, :],
                                                 temps_unmasked / torch.where(good, self.mask, 1.0)[None, None, :],
                                                 temps_unmasked)
            self.alphas, self.iM = deproject(masked, temps, tol)
            self.n_temp = len(temps)
            if pure_any:
                alpha_d = torch.as_tensor(self.alphas, dtype=torch.float64, device=dev)
                _lib.call("gm_map_deproject", _lib.ptr(unmasked), unmasked.numel(), _lib.ptr(temps_unmasked),
                          len(temps), unmasked.numel(), _lib.ptr(alpha_d), _lib.stream_ptr(dev))
        key = _mask_key(self.mask, lmax_mask, self.n_iter_mask, self._mask_precision)
        self.alm_mask = _MASK_ALMS.get(key)
        if pure_any:
            amask = self.get_mask_alms().as_subclass(torch.Tensor)
            self.alm, masked = purify(self.mask, amask, unmasked, lmax, (self.pure_e, self.pure_b), self.n_iter,
                                      precision=_precision())
        elif temps is None and _batchable(self.minfo.nside, lmax):
            # deferred: the transform joins the next batched pass (with the mask when it
            # would be paired, as below)
            pair = (self.spin == 0 and not self.lite and self.ainfo == self.ainfo_mask
                    and self.n_iter == self.n_iter_mask and self._alm_mask is None
                    and not any(it["mask_key"] == key and "mask" in it["what0"] for it in _PENDING))
            own = masked.clone() if masked is maps_t else masked  # never the caller's tensor
            item = {"field": self, "mask_key": key, "maps0": None, "maps2": None, "what0": []}
            if self.spin == 0:
                item["maps0"] = torch.cat([masked, self.mask[None, :]]) if pair else own
                item["what0"] = ["alm", "mask"] if pair else ["alm"]
            else:
                item["maps2"] = own
            _PENDING.append(item)
            self._pending = True
        elif (self.spin == 0 and not self.lite and self.ainfo == self.ainfo_mask
                and self.n_iter == self.n_iter_mask and self.alm_mask is None):
            pair = sht.map2alm_s0(torch.cat([masked, self.mask[None, :]]), lmax, n_iter=self.n_iter,
                                  precision=_precision())
            self.alm = pair[0:1]
            self.alm_mask = pair[1]
            _remember(key, self.alm_mask)
        else:
            self.alm = sht.map2alm(masked, self.spin, lmax, n_iter=self.n_iter, precision=_precision())
        if not self.lite:
            self.maps = masked
            if temps is not None:
                self.temp = temps
                if pure_any:
                    amask = self.get_mask_alms().as_subclass(torch.Tensor)
                    self.alm_temp = torch.stack([
                        purify(self.mask, amask, t, lmax, (self.pure_e, self.pure_b), self.n_iter,
                               precision=_precision(), return_maps=False) for t in temps_unmasked])
                elif self.spin == 0:
                    self.alm_temp = sht.map2alm_s0(temps[:, 0, :], lmax, n_iter=self.n_iter,
                                                   precision=_precision())[:, None, :]
                else:
                    self.alm_temp = torch.stack([sht.map2alm(t, self.spin, lmax, n_iter=self.n_iter,
                                                             precision=_precision()) for t in temps])

    def is_compatible(self, other, strict=True):
        if strict and self.minfo != other.minfo:
            return False
        return self.ainfo_mask == other.ainfo_mask and self.ainfo == other.ainfo

    def get_mask(self):
        return wrap(self.mask)

    def get_maps(self):
        if self.maps is None:
            raise ValueError("Input maps unavailable for this field")
        return wrap(self.maps)

    def get_templates(self):
        if self.temp is None:
            raise ValueError("Input templates unavailable for this field")
        return wrap(self.temp)

    def get_alms(self):
        if self.alm is None:
            raise ValueError("Input maps unavailable for this field")
        return wrap(self.alm)

    def get_mask_alms(self):
        if self.alm_mask is None:
            alm = self.mask_alms_at(getattr(self, "_mask_precision", _precision()))
            if self.lite:
                return wrap(alm)
            self.alm_mask = alm
        return wrap(self.alm_mask)

    def mask_alms_at(self, precision):
        """Mask alms computed with the given transform precision ("fp64" or "mixed"),
        shared through the mask-alms cache."""
        key = _mask_key(self.mask, self.ainfo_mask.lmax, self.n_iter_mask, precision)
        alm = _MASK_ALMS.get(key)
        if alm is None:
            alm = sht.map2alm_s0(self.mask[None, :], self.ainfo_mask.lmax, n_iter=self.n_iter_mask,
                                 precision=precision)[0]
            _remember(key, alm)
        return alm
