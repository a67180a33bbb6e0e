"""MI355X (gfx950) two-site Trotter update for OpenSourceTEBD-style Vidal states."""
from . import models, rank_model
from ._lib import TEBDError, LIB_PATH
from .engine import (Context, Ensemble, construct_propagators, imag_time_prop, parse_mode, read_state_files,
                     write_state_files)
from .measures import evaluate_measures

__all__ = ["Context", "Ensemble", "TEBDError", "construct_propagators", "imag_time_prop", "evaluate_measures",
           "read_state_files", "write_state_files", "models", "rank_model", "parse_mode", "LIB_PATH"]
