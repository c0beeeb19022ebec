"""gmaster_amd: a pymaster-compatible curved-sky MASTER engine on hand-written gfx950 kernels.

``import gmaster_amd as nmt`` (or ``import gmaster as nmt``) exposes the NaMaster API of
the curved-sky HEALPix path: ``NmtField``, ``NmtBin``, ``NmtWorkspace.from_fields``,
``compute_coupled_cell``, ``decouple_cell``, ``map2alm`` / ``alm2map`` of any spin,
contaminant deprojection (``templates=``, ``deprojection_bias``,
``uncorr_noise_deprojection_bias``, ``compute_full_master``).
Every transform, coupling matrix and spectrum runs in the HIP library
``libgmaster_amd.so`` (built by ``python -m gmaster_amd.build``); there is no CPU path.
"""

from . import _lib, batched, sht  # noqa: F401
from ._array import GmArray
from ._config import (get_default_params, nmt_params, set_batching, set_form, set_n_iter_default,
                      set_sht_calculator, set_tol_pinv_default)
from .bins import NmtBin
from .deproject import compute_full_master, deprojection_bias, moore_penrose_pinvh, uncorr_noise_deprojection_bias
from .field import NmtField
from .utils import (
    NmtAlmInfo,
    NmtMapInfo,
    alm2map,
    map2alm,
    map2alm_pair,
    mask_apodization,
    synfast_spherical,
)
from .workspaces import (NmtWorkspace, compute_coupled_cell, coupling_precision, get_general_coupling_matrix,
                         get_master_coefficients, set_coupling_precision)

__all__ = [
    "GmArray",
    "NmtAlmInfo",
    "NmtBin",
    "NmtField",
    "NmtMapInfo",
    "NmtWorkspace",
    "alm2map",
    "compute_coupled_cell",
    "compute_full_master",
    "deprojection_bias",
    "get_default_params",
    "coupling_precision",
    "get_general_coupling_matrix",
    "get_master_coefficients",
    "set_coupling_precision",
    "map2alm",
    "map2alm_pair",
    "mask_apodization",
    "moore_penrose_pinvh",
    "nmt_params",
    "set_batching",
    "set_form",
    "set_n_iter_default",
    "set_sht_calculator",
    "set_tol_pinv_default",
    "synfast_spherical",
    "uncorr_noise_deprojection_bias",
]
