"""ctypes binding of libtebdhip.so (declarations: include/tebd_hip.h).

The HIP library is the only compute path: if it is not built, importing the
binding raises instead of falling back to anything else.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "libtebdhip.so")

TEBD_OK = 0
TEBD_ERR_ARG = -1
TEBD_ERR_HIP = -2
TEBD_ERR_SHAPE = -3
TEBD_ERR_NAN = -4
TEBD_ERR_NOCONV = -5
TEBD_ERR_STATE = -6
TEBD_ERR_UNSUPPORTED = -7

MODE_FP64 = 0
MODE_MIXED = 1 << 0
MODE_BATCHED = 1 << 1
MODE_LEARNED_RANK = 1 << 2
MODE_SPLIT = 1 << 3

MLP_INPUTS = 66
MLP_HIDDEN = 128
MLP_PARAMS = MLP_HIDDEN * MLP_INPUTS + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1

_lib = None


class TEBDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                              "(make -C opentebd-extended_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.tebd_create.argtypes = [ctypes.POINTER(P), I, I, I, I, ctypes.c_char, ctypes.c_uint]
        L.tebd_create_ensemble.argtypes = [ctypes.POINTER(P), I, I, I, I, I, ctypes.c_char, ctypes.c_uint]
        L.tebd_destroy.argtypes = [P]
        L.tebd_destroy.restype = None
        L.tebd_set_gates.argtypes = [P, P, I]
        L.tebd_upload.argtypes = [P, P, P, I]
        L.tebd_upload_packed.argtypes = [P, P, P, I]
        L.tebd_download.argtypes = [P, P, P, I]
        L.tebd_download_packed.argtypes = [P, P, P, I]
        L.tebd_download_lambdas.argtypes = [P, P, I]
        L.tebd_trotter_step.argtypes = [P, ctypes.POINTER(D)]
        L.tebd_two_site_op.argtypes = [P, I, P, ctypes.POINTER(D)]
        L.tebd_canonical_form_all.argtypes = [P]
        L.tebd_canonical_form.argtypes = [P, I]
        L.tebd_set_cutoff.argtypes = [P, D]
        L.tebd_get_bond_dims.argtypes = [P, P]
        L.tebd_get_stats.argtypes = [P, P, I]
        L.tebd_sync.argtypes = [P]
        L.tebd_memory_info.argtypes = [P, P, I]
        L.tebd_two_point_rdms.argtypes = [P, P, P, P]
        Cs = ctypes.c_char_p
        L.tebd_write_state_files.argtypes = [I, I, I, P, P, Cs, Cs, ctypes.c_char]
        L.tebd_read_state_files.argtypes = [I, I, I, P, P, Cs, Cs, ctypes.c_char]
        L.tebd_checkpoint.argtypes = [P, Cs, Cs, ctypes.c_char]
        L.tebd_restart.argtypes = [P, Cs, Cs, ctypes.c_char, I]
        L.tebd_construct_propagators.argtypes = [P, I, I, I, ctypes.c_char, D, D, P]
        L.tebd_imag_time_prop.argtypes = [P, P, I, D, I, I, I, D, D, P]
        L.tebd_set_autotune.argtypes = [P, I]
        L.tebd_set_lanes.argtypes = [P, I]
        L.tebd_set_graphs.argtypes = [P, I]
        L.tebd_autotune_info.argtypes = [P, P, I]
        L.tebd_set_profiling.argtypes = [P, I]
        L.tebd_get_profile.argtypes = [P, P, I]
        L.tebd_expval_two_site.argtypes = [P, P, P]
        L.tebd_rank_model_set.argtypes = [P, P, I]
        L.tebd_set_learned_params.argtypes = [P, D, D, D]
        L.tebd_rank_model_get.argtypes = [P, P, I]
        L.tebd_rank_log.argtypes = [P, I]
        L.tebd_rank_log_get.argtypes = [P, P, P, I]
        L.tebd_sketch_log_get.argtypes = [P, P, I]
        L.tebd_set_conserved.argtypes = [P, P, I]
        L.tebd_set_labels.argtypes = [P, P, P, I]
        L.tebd_get_labels.argtypes = [P, P, P, I]
        F = ctypes.c_float
        L.tebd_mlp_gradient.argtypes = [P, P, P, I, F, P, P]
        L.tebd_mlp_train.argtypes = [P, P, P, I, I, F, F, P]
        L.tebd_mlp_predict.argtypes = [P, P, I, P]
        L.tebd_last_error.argtypes = [P]
        L.tebd_last_error.restype = ctypes.c_char_p
        for name in ("tebd_create", "tebd_create_ensemble", "tebd_set_gates", "tebd_upload", "tebd_upload_packed", "tebd_download",
                     "tebd_download_packed", "tebd_download_lambdas", "tebd_trotter_step", "tebd_two_site_op", "tebd_canonical_form_all", "tebd_canonical_form",
                     "tebd_set_cutoff", "tebd_get_bond_dims", "tebd_get_stats", "tebd_sync", "tebd_memory_info", "tebd_two_point_rdms", "tebd_write_state_files", "tebd_read_state_files", "tebd_checkpoint", "tebd_restart",
                     "tebd_construct_propagators", "tebd_imag_time_prop", "tebd_set_autotune", "tebd_set_lanes", "tebd_set_graphs",
                     "tebd_autotune_info",
                     "tebd_set_profiling", "tebd_get_profile", "tebd_expval_two_site", "tebd_rank_model_set",
                     "tebd_set_learned_params", "tebd_rank_model_get", "tebd_rank_log", "tebd_rank_log_get", "tebd_sketch_log_get", "tebd_mlp_gradient",
                     "tebd_set_conserved", "tebd_set_labels", "tebd_get_labels",
                     "tebd_mlp_train", "tebd_mlp_predict"):
            getattr(L, name).restype = I
        _lib = L
    return _lib
