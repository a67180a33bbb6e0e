"""MI355X (gfx950) two-site Trotter update for OpenSourceTEBD-style Vidal states."""
from . import models
from ._lib import TEBDError, LIB_PATH
from .engine import Context, parse_mode

__all__ = ["Context", "TEBDError", "models", "parse_mode", "LIB_PATH"]
