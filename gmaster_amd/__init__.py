"""gmaster_amd: a pymaster-compatible curved-sky MASTER engine on hand-written gfx950 kernels."""

from . import _lib, sht  # noqa: F401
