"""Drop-in module name: ``import gmaster as nmt`` resolves to the gfx950 engine ``gmaster_amd``."""

from gmaster_amd import *  # noqa: F401,F403
from gmaster_amd import __all__  # noqa: F401
