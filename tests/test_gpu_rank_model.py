"""Rank model of the learned-rank path: GPU kernels against the CPU reference, and
what the trained model gains over the path without it.

The gradient and inference of the hand-written kernels (tebd_mlp.hip) must match the
float64 reference of tests/mlp_reference.py (itself pinned by finite differences) to
FP32 accuracy.  The ablation trains the model on the path's own record of one
Fermi-Hubbard quench (U=4: every bond update's features and log2 of the rank it needs
over its dimension before) and applies it to another (U=3): the path sized by it must
redo at most half as many sketches by the full SVD, in less device time, with
observables within 1e-9 of the exact path's.
"""
import numpy as np
import pytest

from mlp_reference import forward, loss_grad
from tebd_amd import Context, models as M, rank_model as RM, train_rank_model as TR

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


def _quench(U0, mode, model=None, log=False, L=64, chi=128, steps=16, dt=0.05):
    H = M.fermi_hubbard(L, 1.0, U0)
    ctx = Context(L, 4, chi, mode=mode)
    ctx.upload(*M.product_state(M.neel(L, 4, 1, 2), chi))
    ctx.set_gates(M.construct_propagators(H, dt))
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


def test_trained_model_predicts_rank_and_sizes_the_sketch():
    """Fermi-Hubbard L=64, chi=128, dt=0.05, 16 steps from the Neel state: the bonds grow
    to chi and then discard weight the sketches cannot hold to their tolerances, so with
    the fixed width most sketches of the later steps are computed and then redone by the
    full SVD.  The model trained on U=4 must, on U=3, redo at most half as many, in less
    device time, with the observables of the exact path (1e-9).  (Eight steps more, with
    a total discarded weight of 5e-7, the exact path itself carries rounding-level
    differences in its state into ~1e-9 differences of the observables.)"""
    print("[gpu] rank model ablation: FH L=64 chi=128 dt=0.05 16 steps, train U=4, apply U=3", flush=True)
    _, _, (X, y) = _quench(4.0, "batched,learned", log=True)
    assert len(y) > 1000 and y.max() > 0.5
    params, _ = TR.fit(X, y)
    exact, stE, _ = _quench(3.0, "batched")
    plain, st0, (X3, y3) = _quench(3.0, "batched,learned", log=True)
    learned, st1, _ = _quench(3.0, "batched,learned", model=params)
    p = RM.predict(params, X3)
    kmax = np.exp2(X3[:, 63] * 4.0)  # kmax / chi_c
    beyond = y3 > np.log2(1.5 * kmax)
    kept = ~beyond
    print(f"[gpu] held-out: {beyond.mean():.3f} of the updates discard more than 1e-10; model flags "
          f"{np.mean(p[beyond] > np.log2(np.sqrt(2) * kmax[beyond])):.3f} of them and "
          f"{np.mean(p[kept] > np.log2(np.sqrt(2) * kmax[kept])):.3f} of the others; |log2 rank error| on the others "
          f"{np.mean(np.abs(p[kept] - y3[kept])):.3f}", flush=True)
    print(f"[gpu] sketches rejected and redone: fixed width {st0['fallbacks']:.0f} of "
          f"{st0['sketch_accepted'] + st0['fallbacks']:.0f}, model {st1['fallbacks']:.0f} of "
          f"{st1['sketch_accepted'] + st1['fallbacks']:.0f} ({st1['model_full_svd']:.0f} sent to the full SVD at once); "
          f"sketch columns {st0['sketch_cols']:.0f} vs {st1['sketch_cols']:.0f}; device ms: exact "
          f"{stE['device_ms']:.0f}, fixed {st0['device_ms']:.0f}, model {st1['device_ms']:.0f}", flush=True)
    assert st1["fallbacks"] <= 0.5 * st0["fallbacks"]
    assert st1["device_ms"] < st0["device_ms"]
    assert np.max(np.abs(plain - exact)) < 1e-9
    assert np.max(np.abs(learned - exact)) < 1e-9
