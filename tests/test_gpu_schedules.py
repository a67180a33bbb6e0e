"""The device's 5th-order (Forest-Ruth) and periodic-boundary schedules against the
reference's own trajectories (tests/golden/sched_*.npz), in every batching mode.

The comparison is the oracle's (test_oracle_schedules.py): truncation error, <Sz_i> and
entropy after every step by the reference's local formulas, the energy by TotalEnergy's
formula (open) or from the state vector with the edge bond (periodic, full rank), and the
final lambdas, all within 1e-9.
"""
import numpy as np
import pytest

from conftest import load_golden
from test_oracle_schedules import CASES, schedule_errors
from tebd_amd import Context, Ensemble, models as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["fp64", "batched", "batched,learned"])
@pytest.mark.parametrize("name", CASES)
def test_device_schedules_reproduce_reference(name, mode):
    z = load_golden(name)
    L, d, chi, order, bc = int(z["L"]), int(z["d"]), int(z["chi"]), int(z["order"]), str(z["bc"])
    print(f"[gpu] schedule {name} mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode, order=order, bc=bc) as ctx:
        ctx.upload(*M.product_state(z["coeffs"], chi))
        ctx.set_gates(z["U"])

        def steps():
            for _ in range(int(z["nsteps"])):
                tr = ctx.trotter_step()
                g, lam = ctx.download(chi)
                yield tr, g, lam

        worst = schedule_errors(z, steps)
    print(f"[gpu] {name} {mode}: {worst}", flush=True)
    assert max(worst.values()) < 1e-9, worst


def test_fifth_order_ensemble_matches_single_chains():
    """An ensemble of three chains under the 5th-order schedule: every chain equals the
    same chain run alone (per-chain gate tables, bond numbering restarted per chain)."""
    z = load_golden("sched_o5_obc_l10")
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    rng = np.random.default_rng(9)
    fields = [rng.uniform(-0.5, 0.5, L) for _ in range(3)]
    Us = [M.construct_propagators(M.heisenberg(L, 0.5, 1.0, 1.0, 0.7, field=f), 0.05, order=5) for f in fields]
    gam, lam = M.product_state(z["coeffs"], chi)
    with Ensemble(3, L, d, chi, mode="batched", order=5) as ens:
        ens.upload_chains(np.stack([gam] * 3), np.stack([lam] * 3))
        ens.set_gates(np.stack(Us))
        for _ in range(5):
            ens.trotter_step()
        lam_e = ens.lambdas_chains(chi)
    for r in range(3):
        with Context(L, d, chi, mode="batched", order=5) as ctx:
            ctx.upload(gam, lam)
            ctx.set_gates(Us[r])
            for _ in range(5):
                ctx.trotter_step()
            assert np.max(np.abs(ctx.lambdas(chi) - lam_e[r])) == 0.0
