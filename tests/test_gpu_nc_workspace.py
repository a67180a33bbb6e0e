"""The number-conserving update gives the same state whichever workspace factorises a
block: the per-bond mode (one block at a time in the bond-sized slot) and the batched
mode (the blocks of a half-sweep in the small-block workspace, largest first) agree bit
for bit -- Gammas, lambdas, labels and truncation errors -- on a Bose-Hubbard chain
(L=16, d=3, chi=48, Mott start): after 6 steps below the cap, after 8 steps truncated at
chi=48 across number blocks."""
import numpy as np
import pytest

from tebd_amd import Context, models as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,steps,capped", [(43, 6, False), (41, 8, True)])
def test_nc_batched_equals_per_bond_bit_for_bit(seed, steps, capped):
    L, d, chi = 16, 3, 48
    print(f"[gpu] NC per-bond vs batched, BH L={L} chi={chi} seed={seed} steps={steps}", flush=True)
    J = np.random.default_rng(seed).uniform(0.8, 1.2, L - 1)
    U = M.construct_propagators(M.bose_hubbard(L, 2, J, 2.0, 0.0, 0.0), 0.05)
    coeffs = np.zeros((L, d), np.complex128)
    coeffs[:, 1] = 1.0
    cv = np.arange(d)
    out = {}
    for mode in ("fp64", "batched"):
        with Context(L, d, chi, mode=mode) as ctx:
            ctx.upload(*M.product_state(coeffs, chi))
            ctx.set_gates(U)
            ctx.set_conserved(cv, L, labels=M.product_labels(coeffs, chi, cv, L))
            tr = [ctx.trotter_step() for _ in range(steps)]
            out[mode] = (*ctx.download(chi), *ctx.labels(chi), np.array(tr))
    a, b = out["fp64"], out["batched"]
    diffs = [float(np.max(np.abs(x - y))) for x, y in zip(a, b)]
    print(f"[gpu] max differences (gamma, lambda, label left, label right, truncerr): {diffs}", flush=True)
    dims = (b[1] > 0).sum(axis=1)
    assert dims.max() == chi if capped else 27 < dims.max() < chi
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
