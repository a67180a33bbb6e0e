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


def mps_to_vector(gam, lam):
    """State vector psi[s_1, ..., s_L] of a Vidal state in the padded layout (small L):
    psi = Gamma_1 lambda_2 Gamma_2 ... lambda_L Gamma_L with the outer bonds of dimension 1."""
    L, chi, d, _ = gam.shape
    # site tensors A_s(a, i, b) = Gamma_s(a, i, b) lambda_{s+1}(b); gam is indexed [site, b, i, a]
    v = np.transpose(gam[0], (2, 1, 0))[0]  # (i, b)
    v = v * lam[1][None, :]
    for s in range(1, L):
        g = np.transpose(gam[s], (2, 1, 0))  # (a, i, b)
        v = np.tensordot(v, g, axes=([-1], [0]))
        if s < L - 1:
            v = v * lam[s + 1]
    return v[..., 0]


def vector_observables(psi, H, op, pbc=False):
    """Energy (bonds b, b+1, and with pbc the last entry of H on sites 1, L), <op_i> and
    the entanglement entropy (log_d) of every cut of a state vector."""
    L = psi.ndim
    d = psi.shape[0]

    def apply(G, a, b, p):
        g = G.reshape(d, d, d, d)
        q = np.tensordot(g, p, axes=([2, 3], [a, b]))
        return np.moveaxis(q, [0, 1], [a, b])

    nrm = np.vdot(psi, psi).real
    E = sum(np.vdot(psi, apply(H[b], b, b + 1, psi)).real for b in range(L - 1))
    if pbc:
        E += np.vdot(psi, apply(H[L - 1], 0, L - 1, psi)).real
    loc = np.array([np.vdot(psi, np.moveaxis(np.tensordot(op, psi, axes=([1], [s])), 0, s)).real
                    for s in range(L)]) / nrm
    S = []
    for s in range(1, L):
        sv = np.linalg.svd(psi.reshape(d ** s, -1), compute_uv=False)
        p = sv ** 2 / nrm
        p = p[p > 0]
        S.append(-np.sum(p * np.log(p)) / np.log(d))
    return E / nrm, loc, np.array(S)
