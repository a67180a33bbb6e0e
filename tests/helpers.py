"""Shared helpers of the parity tests (the oracle is only ever the checker)."""
import numpy as np

from oracle.oracle import OracleState


def observables(gam, lam, H, szop, nop):
    st = OracleState(gam, lam)
    return dict(E=st.total_energy(H), sz=st.one_site_expval(szop), n=st.one_site_expval(nop),
                S=st.chain_entropy())


def max_obs_err(a, b):
    return {k: float(np.max(np.abs(np.asarray(a[k]) - np.asarray(b[k])))) for k in a}


def golden_obs(z, k):
    return dict(E=z["energy"][k], sz=z["sz"][k], n=z["n"][k], S=z["S"][k])


def hamiltonian_of(z):
    from tebd_amd import models as M
    model, L, c = int(z["model"]), int(z["L"]), z["consts"]
    if model == 1:
        return M.heisenberg(L, 0.5 * int(z["param"]), *c, field=z["field"])
    if model == 2:
        return M.bose_hubbard(L, int(z["param"]), z["field"], c[0], c[1], c[2])
    return M.fermi_hubbard(L, c[0], c[1], c[2], int(z["param"]))
