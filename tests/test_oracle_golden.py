"""The FP64 oracle against vectors written by the reference itself (tests/golden)."""
import numpy as np
import pytest

from conftest import load_golden
from helpers import golden_obs, max_obs_err, observables
from oracle.oracle import OracleState

TRAJ = ["xxz_l32_chi32", "xxz_l16_chi8", "bh_l12_chi24", "fh_l10_chi32", "fh_l8_chi8", "s1_l8_chi12",
        "itp_l10_chi8"]


@pytest.mark.parametrize("name", TRAJ)
def test_oracle_trajectory_matches_reference(name):
    z = load_golden(name)
    st = OracleState.product(z["coeffs"], int(z["chi"]))
    cum = cum_ref = 0.0
    worst = {}
    for k in range(int(z["nsteps"]) + 1):
        if k:
            cum += st.trotter_step(z["U"])
            if int(z["canon"]):
                st.canonical_form_all()
            cum_ref += z["trunc"][k]
            assert abs(cum - cum_ref) < 1e-12
        err = max_obs_err(observables(st.gam, st.lam, z["H"], z["szop"], z["nop"]), golden_obs(z, k))
        for key, v in err.items():
            worst[key] = max(worst.get(key, 0.0), v)
    assert max(worst.values()) < 1e-11, worst
    np.testing.assert_allclose(st.lam, z["lam"], atol=1e-9)


def test_oracle_single_two_site_op():
    z = load_golden("two_site_op_l4")
    st = OracleState(z["gam_in"], z["lam_in"])
    link = int(z["link"])
    t = st.two_site_op(link, z["U"][link - 1])
    assert abs(t - z["trunc"][0]) < 1e-13
    np.testing.assert_allclose(st.lam, z["lam"], atol=1e-14)
    # Gamma is fixed only up to the SVD's column phases: compare moduli
    np.testing.assert_allclose(np.abs(st.gam), np.abs(z["gam"]), atol=1e-12)


def test_oracle_threads_do_not_change_results():
    z = load_golden("fh_l10_chi32")
    a = OracleState.product(z["coeffs"], int(z["chi"]))
    b = a.copy()
    for _ in range(3):
        ta = a.trotter_step(z["U"], nthreads=1)
        tb = b.trotter_step(z["U"], nthreads=4)
        assert ta == tb
    assert np.array_equal(a.lam, b.lam)
