"""The learned-rank path's rank model: a 66 -> 128 -> 128 -> 1 FP32 perceptron whose
training, gradient and inference run in the library's own GPU kernels (tebd_mlp.hip).

The model predicts the rank k a bond update needs, as log2(k / chi_c): the rank it
keeps, or 2 kmax (kmax = min(chi_max, d chi_l, d chi_r)) when no sketch up to the cap
passes -- the update discards more than the sketch's tolerance 1e-10, or its sketch kept
kmax values and failed the probe check of the kept vectors.  Inputs: the last
(smallest) 63 of the bond's Schmidt values before the update (log10(lambda)/10, -3 in
front when fewer), log2(kmax / chi_c)/4, log2(chi_c)/12 and log10(||G - 1||_F)/4 of the
bond's gate G (how far the update moves theta).  The learned-rank path
sketches the bond with k + max(4, k/16) + 8 columns (rounded up to 32), or takes the
full SVD at once when k > sqrt(2) kmax or the width is more than half of theta's smaller
side.  Training data: Context.rank_log records (features, log2 ratio) of every bond
update.  The loss weights under-predictions by w_under: a sketch that is too narrow is
rejected and redone, one that is too wide only costs wider GEMMs.
"""
import ctypes

import numpy as np

from ._lib import MLP_HIDDEN, MLP_INPUTS, MLP_PARAMS, TEBDError, lib

_P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731


def _check(rc):
    if rc != 0:
        raise TEBDError(rc, "rank model kernel failed")


def init_params(seed=0):
    """He-initialised parameters (layout: W1[128][66], b1, W2[128][128], b2, W3[128], b3)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, np.sqrt(2.0 / MLP_INPUTS), (MLP_HIDDEN, MLP_INPUTS))
    w2 = rng.normal(0.0, np.sqrt(2.0 / MLP_HIDDEN), (MLP_HIDDEN, MLP_HIDDEN))
    w3 = rng.normal(0.0, np.sqrt(1.0 / MLP_HIDDEN), MLP_HIDDEN)
    p = np.concatenate([w1.ravel(), np.zeros(MLP_HIDDEN), w2.ravel(), np.zeros(MLP_HIDDEN), w3, [0.0]])
    assert p.size == MLP_PARAMS
    return p.astype(np.float32)


def _prep(X, y=None):
    X = np.ascontiguousarray(X, np.float32)
    if X.ndim != 2 or X.shape[1] != MLP_INPUTS:
        raise ValueError("features must be n x 66")
    if y is None:
        return X, None
    y = np.ascontiguousarray(y, np.float32)
    if y.shape != (X.shape[0],):
        raise ValueError("one target per sample")
    return X, y


def gradient(params, X, y, w_under=1.0):
    """(loss, dloss/dparams) of the weighted mean squared error, on the GPU."""
    X, y = _prep(X, y)
    p = np.ascontiguousarray(params, np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    _check(lib().tebd_mlp_gradient(_P(p), _P(X), _P(y), X.shape[0], ctypes.c_float(w_under), _P(g), _P(loss)))
    return float(loss[0]), g


def train(X, y, params=None, epochs=1500, lr=2e-3, w_under=1.0, seed=0):
    """Full-batch Adam on the GPU; returns (params, loss per epoch)."""
    X, y = _prep(X, y)
    p = np.array(init_params(seed) if params is None else params, np.float32)
    hist = np.zeros(epochs, np.float32)
    _check(lib().tebd_mlp_train(_P(p), _P(X), _P(y), X.shape[0], epochs, ctypes.c_float(lr), ctypes.c_float(w_under),
                                _P(hist)))
    return p, hist


def predict(params, X):
    X, _ = _prep(X)
    p = np.ascontiguousarray(params, np.float32)
    out = np.zeros(X.shape[0], np.float32)
    _check(lib().tebd_mlp_predict(_P(p), _P(X), X.shape[0], _P(out)))
    return out

