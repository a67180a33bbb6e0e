"""Rank model of the learned-rank path: GPU kernels against the CPU reference, and
what the trained model gains over the path without it.

The gradient and inference of the hand-written kernels (tebd_mlp.hip) must match the
float64 reference of tests/mlp_reference.py (itself pinned by finite differences) to
FP32 accuracy.  The ablation trains the model on the path's own record of accepted /
rejected sketches of one Fermi-Hubbard quench (U=4) and applies it to another (U=3):
it must redo fewer bonds by the full SVD after a wasted sketch than the path without
the model, with observables equal to the exact path's.
"""
import numpy as np
import pytest

from mlp_reference import forward, loss_grad
from tebd_amd import Context, models as M, rank_model as RM

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w_under", [1.0, 4.0])
def test_gpu_gradient_matches_reference(w_under):
    rng = np.random.default_rng(11)
    p = RM.init_params(2)
    X = rng.normal(0.0, 1.0, (257, 66)).astype(np.float32)
    y = rng.uniform(-0.5, 1.5, 257).astype(np.float32)
    loss, g = RM.gradient(p, X, y, w_under)
    loss_ref, g_ref = loss_grad(p, X, y, w_under)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    assert np.linalg.norm(g - g_ref) <= 1e-4 * np.linalg.norm(g_ref)


def test_gpu_inference_matches_reference():
    rng = np.random.default_rng(12)
    p = RM.init_params(3)
    X = rng.normal(0.0, 1.0, (100, 66)).astype(np.float32)
    assert np.max(np.abs(RM.predict(p, X) - forward(p, X)[0])) < 1e-4


def test_gpu_training_learns_a_threshold_rule():
    rng = np.random.default_rng(13)
    X = rng.normal(-1.0, 0.4, (2000, 66)).astype(np.float32)
    y = (X[:, 63] > -1.0).astype(np.float32)
    p, hist = RM.train(X, y, epochs=800)
    assert hist[-1] < 0.2 * hist[0]
    assert np.mean((RM.predict(p, X) >= 0.5) == (y > 0.5)) > 0.95


def _quench(U0, mode, model=None, log=False, L=64, chi=128, steps=40):
    H = M.fermi_hubbard(L, 1.0, U0)
    ctx = Context(L, 4, chi, mode=mode)
    ctx.upload(*M.product_state(M.neel(L, 4, 1, 2), chi))
    ctx.set_gates(M.construct_propagators(H, 0.01))
    if model is not None:
        ctx.set_rank_model(model)
    if log:
        ctx.rank_log(True)
    for _ in range(steps):
        ctx.trotter_step()
    _, _, ntot, _, _ = M.fermion_ops(2)
    obs = np.concatenate([[ctx.energy(H)], ctx.one_site_expval(ntot), ctx.chain_entropy()])
    st = ctx.stats()
    data = ctx.rank_log_data() if log else None
    ctx.close()
    return obs, st, data


def _auc(p, y):
    """Probability that a random accepted sketch scores above a random rejected one."""
    pos, neg = p[y > 0.5], p[y < 0.5]
    return float((pos[:, None] > neg[None, :]).mean() + 0.5 * (pos[:, None] == neg[None, :]).mean())


def test_trained_model_predicts_sketch_outcomes():
    """The network, trained on one quench (U=4), ranks the sketch outcomes of another
    (U=3) far better than a predictor without information (AUC 0.5), and running the
    learned-rank path with it keeps every observable within 1e-9 of the exact path.
    The count of sketches computed and then rejected, with and without the model, is
    printed (on these quenches about 9 % of the sketches are rejected, and the model,
    at its default threshold, changes that count by a few)."""
    print("[gpu] rank model ablation: FH L=64 chi=128, train U=4, apply U=3", flush=True)
    _, _, (X, y) = _quench(4.0, "batched,learned", log=True)
    assert 0.0 < y.mean() < 1.0
    params, _ = RM.train(X, y)
    exact, _, _ = _quench(3.0, "batched")
    plain, st0, (X3, y3) = _quench(3.0, "batched,learned", log=True)
    learned, st1, _ = _quench(3.0, "batched,learned", model=params)
    p = RM.predict(params, X3)
    auc = _auc(p, y3)
    ll = float(-np.mean(y3 * np.log(np.clip(p, 1e-7, 1.0)) + (1 - y3) * np.log(np.clip(1 - p, 1e-7, 1.0))))
    print(f"[gpu] held-out AUC {auc:.3f}, log-loss {ll:.4f}; rejected sketches without model "
          f"{st0['fallbacks']:.0f}, with model {st1['fallbacks']:.0f} ({st1['sketches_skipped']:.0f} skipped)",
          flush=True)
    assert auc > 0.7
    assert np.max(np.abs(plain - exact)) < 1e-9
    assert np.max(np.abs(learned - exact)) < 1e-9
