"""ImagTimeProp through the C ABI (tebd_imag_time_prop) against the reference's own run.

tests/golden/itp_stages_l10.npz is the reference's ImagTimeProp
(propagation_module.f90:636-760) with chi stages chiIn=4 -> chiMax=8 on a spin-1/2
Heisenberg chain with random fields: for each stage the judgements it printed (step j,
largest relative change of lambda_1, criterion) and the final state.  The device must
end every stage at the same j (both converge here, at j=150 and j=80) and reach the
same state: energy, <Sz_i>, entropy within 1e-9 (FP64 modes) or MIXED_TOL (FP32 theta).
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import golden_obs, max_obs_err, observables
from tebd_amd import Context, imag_time_prop, models as M

pytestmark = pytest.mark.gpu

TOL = 1e-9
MIXED_TOL = 1e-4


@pytest.mark.parametrize("mode", ["fp64", "batched", "batched,learned", "batched,learned,mixed"])
def test_imag_time_prop_stages_match_reference(mode):
    z = load_golden("itp_stages_l10")
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    print(f"[gpu] ImagTimeProp stages mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode) as ctx:
        ctx.upload(*M.product_state(z["coeffs"], chi))
        stages = imag_time_prop(ctx, z["H"], float(z["dtau"]), chi_in=int(z["chi_in"]),
                                max_steps=int(z["max_steps"]), steps_for_judge=int(z["judge"]),
                                conv1=float(z["conv1"]), conv2=float(z["conv2"]))
        dims = ctx.bond_dims()
        g, l = ctx.download(chi)
    print(f"[gpu] stages {stages} (reference {list(z['stage_steps'])})", flush=True)
    assert [s for s, _ in stages] == list(z["stage_steps"])
    assert [c for _, c in stages] == list(z["stage_converged"])
    err = max_obs_err(observables(g, l, z["H"], z["szop"], z["nop"]), golden_obs(z, -1))
    err["lambda"] = float(np.max(np.abs(l - z["lam"])))
    print(f"[gpu] errors {err}", flush=True)
    assert max(err.values()) < (MIXED_TOL if "mixed" in mode else TOL), err
    assert dims.max() <= chi


def test_imag_time_prop_single_stage_and_first_stage_cut():
    """chi_in == chi_max runs one stage; a stage below chi_max first cuts every bond to
    chi_in (entries past it dropped, as CopyGamLam does) and never exceeds it."""
    z = load_golden("itp_stages_l10")
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    with Context(L, d, chi, mode="batched") as ctx:
        ctx.upload(*M.product_state(z["coeffs"], chi))
        one = imag_time_prop(ctx, z["H"], 0.2, chi_in=chi, max_steps=30, steps_for_judge=10, conv1=1e-4,
                             conv2=1e-12)
        assert one == [(30, False)]
        assert ctx.bond_dims().max() == chi
        two = imag_time_prop(ctx, z["H"], 0.2, chi_in=3, max_steps=20, steps_for_judge=10, conv1=1e-12,
                             conv2=1e-12)
        assert len(two) == 2 and all(not c for _, c in two)
        # the chi_max stage ran last, from the cut state
        lam = ctx.lambdas()
        assert np.allclose(np.sum(lam ** 2, axis=1), 1.0, atol=1e-12)
