"""CPU reference of the rank model (test infrastructure): forward pass, weighted
squared-error loss and its gradient in float64, with the parameter layout of
tebd_mlp.hip (W1[128][66], b1[128], W2[128][128], b2[128], W3[128], b3).  The
gradient is pinned against central finite differences in test_rank_model_cpu.py."""
import numpy as np

NI, NH = 66, 128
OFF = np.cumsum([0, NH * NI, NH, NH * NH, NH, NH, 1])


def unpack(p):
    p = np.asarray(p, np.float64)
    W1 = p[OFF[0]:OFF[1]].reshape(NH, NI)
    b1 = p[OFF[1]:OFF[2]]
    W2 = p[OFF[2]:OFF[3]].reshape(NH, NH)
    b2 = p[OFF[3]:OFF[4]]
    W3 = p[OFF[4]:OFF[5]]
    b3 = p[OFF[5]]
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ W3 + b3, (X, h1, h2)


def loss_grad(p, X, y, w_under=4.0):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    out, (X, h1, h2) = forward(p, X)
    y = np.asarray(y, np.float64)
    n = X.shape[0]
    r = out - y
    w = np.where(r < 0.0, w_under, 1.0)
    loss = float(np.sum(w * r * r) / n)
    dout = 2.0 * w * r / n
    gW3 = h2.T @ dout
    gb3 = np.sum(dout)
    dh2 = np.outer(dout, W3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
