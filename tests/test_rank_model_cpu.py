"""The CPU reference of the rank model is pinned by finite differences (no GPU)."""
import numpy as np

from mlp_reference import NH, NI, forward, loss_grad


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0.0, 0.3, NH * NI + NH + NH * NH + NH + NH + 1)
    X = rng.normal(0.0, 1.0, (7, NI))
    y = rng.normal(0.0, 1.0, 7)
    loss, g = loss_grad(p, X, y, 4.0)
    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [p.size - 1]])  # include the output bias
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss_grad(p + e, X, y, 4.0)[0] - loss_grad(p - e, X, y, 4.0)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_reference_forward_is_the_layer_stack():
    rng = np.random.default_rng(4)
    p = rng.normal(0.0, 0.3, NH * NI + NH + NH * NH + NH + NH + 1)
    x = rng.normal(0.0, 1.0, NI)
    W1 = p[:NH * NI].reshape(NH, NI)
    b1 = p[NH * NI:NH * NI + NH]
    o = NH * NI + NH
    W2 = p[o:o + NH * NH].reshape(NH, NH)
    b2 = p[o + NH * NH:o + NH * NH + NH]
    W3 = p[o + NH * NH + NH:o + NH * NH + 2 * NH]
    b3 = p[-1]
    h1 = np.maximum(W1 @ x + b1, 0)
    h2 = np.maximum(W2 @ h1 + b2, 0)
    assert abs(forward(p, x[None])[0][0] - (W3 @ h2 + b3)) < 1e-12
