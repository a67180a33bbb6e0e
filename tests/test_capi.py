"""The C-ABI library loads and exports every entry point include/tebd_hip.h declares
(no device calls: argument errors are rejected before any HIP call)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from tebd_amd import _lib


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tebd_hip.h")).read()
    return sorted(set(re.findall(r"\b(tebd_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert "tebd_trotter_step" in names and len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("args", [
    (1, 2, 4, 2, b"O"),      # L < 2
    (8, 7, 4, 2, b"O"),      # unsupported local dimension
    (8, 2, 4, 3, b"O"),      # Trotter order other than 2 or 5
    (8, 2, 4, 2, b"X"),      # boundary other than open or periodic
    (2, 2, 4, 2, b"P"),      # periodic chain of fewer than 3 sites
    (8, 2, 0, 2, b"O"),      # chi_max < 1
    (8, 4, 2048, 2, b"O"),   # d * chi_max beyond the SVD work size
])
def test_create_rejects_bad_arguments_without_touching_the_gpu(args):
    lib = _lib.lib()
    h = ctypes.c_void_p()
    rc = lib.tebd_create(ctypes.byref(h), *args, 0)
    assert rc == _lib.TEBD_ERR_ARG
    assert not h.value


def test_null_context_calls_are_errors():
    lib = _lib.lib()
    t = ctypes.c_double()
    assert lib.tebd_trotter_step(None, ctypes.byref(t)) == _lib.TEBD_ERR_ARG
    assert lib.tebd_set_cutoff(None, 1e-12) == _lib.TEBD_ERR_ARG
    assert lib.tebd_last_error(None) == b"null context"


@pytest.mark.parametrize("args", [
    (0, 8, 2, 4, 2, b"O"),   # no chain
    (3, 1, 2, 4, 2, b"O"),   # chains of one site
    (3, 8, 7, 4, 2, b"O"),   # unsupported local dimension
    (3, 8, 2, 4, 2, b"P"),   # periodic boundaries only for a single chain
])
def test_create_ensemble_rejects_bad_arguments_without_touching_the_gpu(args):
    lib = _lib.lib()
    h = ctypes.c_void_p()
    rc = lib.tebd_create_ensemble(ctypes.byref(h), *args, 0)
    assert rc == _lib.TEBD_ERR_ARG
    assert not h.value


def test_python_contexts_expose_the_library_surface():
    """Every method the bench, the tests and the docs use is on the Python classes
    (checked without a device)."""
    from tebd_amd import Context, Ensemble
    for name in ("upload", "download", "set_gates", "trotter_step", "two_site_op", "canonical_form_all",
                 "canonical_form", "bond_dims", "stats", "set_profiling", "profile", "sync", "energy",
                 "one_site_expval", "chain_entropy", "expval_two_site", "set_rank_model", "rank_log",
                 "set_autotune", "autotune_info", "set_cutoff", "lambdas", "set_lanes", "set_graphs"):
        assert callable(getattr(Context, name, None)), name
    for name in ("upload_chains", "download_chains", "energy_chains", "one_site_expval_chains",
                 "chain_entropy_chains", "expval_two_site_chains"):
        assert callable(getattr(Ensemble, name, None)), name
    assert issubclass(Ensemble, Context)


def test_mode_flags_match_the_header():
    text = open(os.path.join(ROOT, "include", "tebd_hip.h")).read()
    flags = {m.group(1): eval(m.group(2).replace("u", "")) for m in
             re.finditer(r"TEBD_MODE_([A-Z_0-9]+)\s*=\s*([0-9u <]+)", text)}
    assert flags == {"FP64": _lib.MODE_FP64, "MIXED": _lib.MODE_MIXED, "BATCHED": _lib.MODE_BATCHED,
                     "LEARNED_RANK": _lib.MODE_LEARNED_RANK, "SPLIT": _lib.MODE_SPLIT}
    from tebd_amd.engine import parse_mode
    assert parse_mode("batched,learned,split") == _lib.MODE_BATCHED | _lib.MODE_LEARNED_RANK | _lib.MODE_SPLIT


def test_fp32_and_split_theta_are_exclusive_without_touching_the_gpu():
    lib = _lib.lib()
    h = ctypes.c_void_p()
    rc = lib.tebd_create(ctypes.byref(h), 8, 2, 4, 2, b"O", _lib.MODE_MIXED | _lib.MODE_SPLIT)
    assert rc == _lib.TEBD_ERR_ARG
    assert not h.value
