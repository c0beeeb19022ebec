"""Process-wide defaults, as ``pymaster.utils.nmt_params`` (``gmaster/_config.py:20-56``)."""


class NmtParams:
    """Default settings: Jacobi iterations for maps and masks, pseudo-inverse tolerance."""

    def __init__(self):
        self.sht_calculator = "gfx950"
        self.n_iter_default = 3
        self.n_iter_mask_default = 3
        self.tol_pinv_default = 1e-10


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


def get_default_params():
    return {
        "sht_calculator": nmt_params.sht_calculator,
        "n_iter": nmt_params.n_iter_default,
        "n_iter_mask": nmt_params.n_iter_mask_default,
        "tol_pinv": nmt_params.tol_pinv_default,
    }
