"""Process-wide defaults, as ``pymaster.utils.nmt_params`` (``gmaster/_config.py:20-56``)."""


class NmtParams:
    """Default settings: Jacobi iterations for maps and masks, pseudo-inverse tolerance."""

    def __init__(self):
        self.sht_calculator = "gfx950"
        self.n_iter_default = 3
        self.n_iter_mask_default = 3
        self.tol_pinv_default = 1e-10
        self.form = "fp64"
        self.fast_toeplitz = (2000, 256, 256)
        # spot-check tolerance in correlation units: on random masks at Nside 4096 the wide
        # Toeplitz's bandpower error came out 0.3-1.2x its largest sampled error
        self.fast_tol = 1e-4
        # fast form: fields defer their transforms and run them batched (T, mask and Q/U
        # of one geometry through one recurrence) when their alms are first needed
        self.batch_fields = True


nmt_params = NmtParams()


def set_n_iter_default(n_iter, mask=False):
    """Set the default number of Jacobi iterations (``mask=True``: for mask transforms)."""
    if int(n_iter) < 0:
        raise ValueError("n_iter must be non-negative")
    if mask:
        nmt_params.n_iter_mask_default = int(n_iter)
    else:
        nmt_params.n_iter_default = int(n_iter)


def set_tol_pinv_default(tol_pinv):
    if tol_pinv is None or tol_pinv < 0 or tol_pinv > 1:
        raise ValueError("tol_pinv must be between 0 and 1")
    nmt_params.tol_pinv_default = float(tol_pinv)


def set_sht_calculator(calculator):
    """Accepted for pymaster compatibility; the only calculator is the gfx950 kernel set."""
    nmt_params.sht_calculator = str(calculator)


FORMS = ("fp64", "fast", "learned")
APPROXIMATE_FORMS = ("fast", "learned")  # forms with mixed-precision, batched transforms


def set_form(form, toeplitz=None, tol=None):
    """Select how workspaces build the coupling matrix.

    * ``"fp64"`` (default): the exact Wigner-3j coupling of every entry.
    * ``"fast"``: the Toeplitz approximation with a wide exact region (``toeplitz`` =
      ``(l_toeplitz, l_exact, dl_band)`` in physical multipoles, default (2000, 256, 256)),
      checked by exact spot checks of random non-exact entries: when their largest error in
      correlation units exceeds ``tol`` (default 1e-4, about the bandpower error it lets
      through), the exact coupling is used instead.
      Below lmax = 2 l_toeplitz the exact coupling is used (it is cheap there).
    * ``"learned"``: transforms as ``"fast"``; the coupling from the committed surrogate: the
      Toeplitz approximation with its narrow exact region plus the network's prediction of
      its residual, checked by the same exact spot checks (``tol``) with the exact fallback.
    """
    if form not in FORMS:
        raise ValueError(f"form must be one of {FORMS}")
    nmt_params.form = form
    if toeplitz is not None:
        lt, le, dl = (int(v) for v in toeplitz)
        if not (0 < le <= lt and dl >= 0):
            raise ValueError("need 0 < l_exact <= l_toeplitz and dl_band >= 0")
        nmt_params.fast_toeplitz = (lt, le, dl)
    if tol is not None:
        nmt_params.fast_tol = float(tol)


def set_batching(on):
    """Enable or disable deferred, batched field transforms in the fast form."""
    nmt_params.batch_fields = bool(on)


def get_default_params():
    return {
        "sht_calculator": nmt_params.sht_calculator,
        "n_iter": nmt_params.n_iter_default,
        "n_iter_mask": nmt_params.n_iter_mask_default,
        "tol_pinv": nmt_params.tol_pinv_default,
        "form": nmt_params.form,
        "batch_fields": nmt_params.batch_fields,
    }
