"""Train the learned-rank path's rank model on the GPU from Fermi-Hubbard quenches.

    python -m tebd_amd.train_rank_model OUT.npy [L chi]

Runs the learned-rank path (without a model, so every sketch has the fixed width) on
Neel-start Fermi-Hubbard quenches at U = 3 and U = 5, each at dt = 0.01 (100 steps:
growth, saturation at chi and the first steps with active truncation) and at dt = 0.05
(24 steps, whose saturated bonds discard more than the sketch's tolerance), records every
bond update (features, log2 of the needed rank over the bond dimension before), trains
the model on that record with the library's kernels (under-predictions weighted W_UNDER
times) and writes the FP32 parameters.
"""
import sys

import numpy as np

from . import Context, models as M, rank_model as RM

W_UNDER = 4.0


def record(L=64, chi=128, U0=4.0, steps=60, dt=0.01):
    """(features, targets) of every bond update of one quench."""
    ctx = Context(L, 4, chi, mode="batched,learned")
    ctx.upload(*M.product_state(M.neel(L, 4, 1, 2), chi))
    ctx.set_gates(M.construct_propagators(M.fermi_hubbard(L, 1.0, U0), dt))
    ctx.rank_log(True)
    for _ in range(steps):
        ctx.trotter_step()
    X, y = ctx.rank_log_data()
    ctx.close()
    return X, y


def fit(X, y, epochs=1500):
    return RM.train(X, y, epochs=epochs, w_under=W_UNDER)


def main(argv):
    out = argv[0]
    L, chi = (int(argv[1]), int(argv[2])) if len(argv) > 2 else (64, 128)
    parts = [record(L, chi, U0, steps, dt) for U0 in (3.0, 5.0) for dt, steps in ((0.01, 100), (0.05, 24))]
    X = np.concatenate([p[0] for p in parts])
    y = np.concatenate([p[1] for p in parts])
    params, hist = fit(X, y)
    np.save(out, params)
    err = RM.predict(params, X) - y
    print(f"{len(y)} bond updates recorded (log2 rank ratio mean {y.mean():.3f}, max {y.max():.3f}); "
          f"loss {hist[0]:.3e} -> {hist[-1]:.3e}; |pred - y| mean {np.abs(err).mean():.3f}, "
          f"under-predicted by > 0.1: {np.mean(err < -0.1):.3f}; written to {out}")


if __name__ == "__main__":
    main(sys.argv[1:])
