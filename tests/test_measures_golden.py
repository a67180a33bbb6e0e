"""EvaluateMeasures and the all-pairs correlators (observables_module.f90:836-883,
2286-2526) against the reference's own values (tests/golden/meas_*.npz, its
EvaluateMeasures after a trajectory: local and averaged <Sz> or <n>, the Sz x Sz (n x n)
correlation matrix, for fermions the fermi-phase correlator c+_up x c_up, single-site,
chain and two-site entropies, the Q measure, and the energy -- with the edge bond of a
periodic chain).

CPU: the oracle's numpy restatement of ThetaKernal / ThetaNext / TwoSiteRho, evaluated on
the reference's final state, reproduces every value; and at full rank the pair density
matrices equal those of the state vector (exact).
GPU (test_gpu_measures.py): the device's density matrices through tebd_amd.measures.
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import mps_to_vector
from oracle.oracle import two_point_rdms
from tebd_amd import models as M

CASES = ["meas_fh_l6", "meas_xxz_l8", "meas_xxz_pbc_l7"]


def case_ops(z):
    d, model = int(z["d"]), int(z["model"])
    if model == 3:
        a_up, _, ntot, sz, F = M.fermion_ops(2)
        return dict(loc=sz, corr=np.kron(sz, sz), fcorr=np.kron(a_up.T, a_up), phase=np.diag(F).real)
    sz = M.spin_ops((d - 1) / 2)[2]
    return dict(loc=sz, corr=np.kron(sz, sz), fcorr=None, phase=None)


def trace(op, rho):
    return np.einsum("ab,ba->", op, rho)


def measures_from_rdms(z, rho2, rho1, rho2f=None):
    L, d = int(z["L"]), int(z["d"])
    ops = case_ops(z)
    out = {"m_local": np.array([trace(ops["loc"], rho1[s]) for s in range(L)])}
    out["m_avg"] = np.mean(out["m_local"])
    c = np.zeros((L, L), np.complex128)
    for l1 in range(L):
        for l2 in range(l1 + 1, L):
            c[l1, l2] = trace(ops["corr"], rho2[l1, l2])
    out["m_corr"] = c
    if rho2f is not None:
        f = np.zeros((L, L), np.complex128)
        for l1 in range(L):
            for l2 in range(l1 + 1, L):
                f[l1, l2] = trace(ops["fcorr"], rho2f[l1, l2])
        out["m_fcorr"] = f
    H = z["H"]
    en = sum(trace(H[b], rho2[b, b + 1]).real for b in range(L - 1))
    if str(z["bc"]) == "P":
        en += trace(H[L - 1], rho2[0, L - 1]).real
    out["m_en"] = en
    ent = lambda r: (lambda w: -np.sum(w[w != 0] * np.log(np.abs(w[w != 0]))) / np.log(d))(  # noqa: E731
        np.linalg.eigvalsh(0.5 * (r + r.conj().T)))
    out["m_vn"] = np.array([ent(rho1[s]) for s in range(L)])
    tb = np.diag(2 * out["m_vn"])
    for l1 in range(L):
        for l2 in range(l1 + 1, L):
            tb[l1, l2] = tb[l2, l1] = ent(rho2[l1, l2])
    out["m_tbvn"] = tb
    out["m_qme"] = d / (d - 1) - sum(trace(rho1[s], rho1[s]).real for s in range(L)) * d / (L * (d - 1))
    return out


def compare(z, got):
    L = int(z["L"])
    iu = np.triu_indices(L, 1)
    err = {}
    for k, v in got.items():
        ref = z[k]
        if k in ("m_corr", "m_fcorr"):
            err[k] = float(np.max(np.abs(np.asarray(v)[iu] - ref[iu])))
            err[k + "_hermitian"] = float(np.max(np.abs(ref[iu] - np.conj(ref.T[iu]))))
        else:
            err[k] = float(np.max(np.abs(np.asarray(v) - ref)))
    return err


@pytest.mark.parametrize("name", CASES)
def test_oracle_rdms_reproduce_reference_measures(name):
    z = load_golden(name)
    ops = case_ops(z)
    gam, lam = z["gam"], z["lam"]
    rho2, rho1 = two_point_rdms(gam, lam)
    rho2f = two_point_rdms(gam, lam, phase=ops["phase"])[0] if ops["fcorr"] is not None else None
    err = compare(z, measures_from_rdms(z, rho2, rho1, rho2f))
    assert max(err.values()) < 1e-12, err


def test_oracle_rdms_equal_state_vector_at_full_rank():
    """A canonical state at full rank: every pair density matrix is the partial trace of
    the state vector onto sites (l1, l2)."""
    L, d, chi = 6, 2, 8
    rng = np.random.default_rng(2)
    H = M.heisenberg(L, 0.5, 1.0, 0.7, 1.2, field=rng.uniform(-1, 1, L))
    U = M.construct_propagators(H, 0.1)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    from oracle.oracle import OracleState

    st = OracleState.product(c, chi)
    for _ in range(4):
        st.trotter_step(U)
    psi = mps_to_vector(st.gam, st.lam)
    rho2, rho1 = two_point_rdms(st.gam, st.lam)
    for l1 in range(L):
        for l2 in range(l1 + 1, L):
            rest = [s for s in range(L) if s not in (l1, l2)]
            p = np.transpose(psi, [l1, l2] + rest).reshape(d * d, -1)
            exact = p @ p.conj().T
            assert np.max(np.abs(rho2[l1, l2] - exact)) < 1e-12
