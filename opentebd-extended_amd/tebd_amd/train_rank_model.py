"""Train the learned-rank path's rank model on the GPU from a Fermi-Hubbard quench.

    python -m tebd_amd.train_rank_model OUT.npy [L chi U steps]

Runs the learned-rank path (without a model) on a Neel-start Fermi-Hubbard quench,
records every sketch it tries with its outcome, trains the model on that record with
the library's kernels and writes the FP32 parameters.
"""
import sys

import numpy as np

from . import Context, models as M, rank_model as RM


def record(L=64, chi=128, U0=4.0, steps=60, dt=0.01):
    ctx = Context(L, 4, chi, mode="batched,learned")
    ctx.upload(*M.product_state(M.neel(L, 4, 1, 2), chi))
    ctx.set_gates(M.construct_propagators(M.fermi_hubbard(L, 1.0, U0), dt))
    ctx.rank_log(True)
    for _ in range(steps):
        ctx.trotter_step()
    X, y = ctx.rank_log_data()
    ctx.close()
    return X, y


def main(argv):
    out = argv[0]
    L, chi, U0, steps = (int(argv[1]), int(argv[2]), float(argv[3]), int(argv[4])) if len(argv) > 4 else (64, 128, 4.0, 60)
    X, y = record(L, chi, U0, steps)
    params, hist = RM.train(X, y)
    np.save(out, params)
    print(f"{len(y)} sketches recorded ({y.mean():.3f} accepted), loss {hist[0]:.3e} -> {hist[-1]:.3e}, "
          f"written to {out}")


if __name__ == "__main__":
    main(sys.argv[1:])
