"""The 5th-order (Forest-Ruth) and periodic-boundary Trotter schedules
(propagation_module.f90:116-614) and their propagator tables
(NEWConstructPropagators, TEBDtools_module.f90:1104-1193).

Pinned three ways:
  * the reference's own trajectories (tests/golden/sched_*.npz: its H, U, and the
    energy -- with the edge bond for periodic chains --, <Sz_i> and entropy after
    every step) against the oracle's schedules and the host gate builders;
  * exactness at full rank: each schedule is a product of exact exponentials, so its
    step must approximate exp(-i dt H) of the whole (periodic) Hamiltonian to the
    scheme's order -- the local error of one step falls as dt^3 (2nd order) or dt^5
    (Forest-Ruth), which a wrong factor, slot or order of sweeps breaks;
  * the periodic edge gate, applied after swapping site L next to site 1 and back,
    equals the gate applied directly to sites (1, L) of the state vector.
"""
import numpy as np
import pytest
import scipy.linalg

from conftest import load_golden
from helpers import mps_to_vector, vector_observables
from oracle.oracle import OracleState
from tebd_amd import models as M

CASES = ["sched_o5_obc_l10", "sched_o2_pbc_l8", "sched_o2_pbc_l9", "sched_o5_pbc_l8", "sched_o5_pbc_l9"]


@pytest.mark.parametrize("name", CASES)
def test_gate_tables_match_reference(name):
    z = load_golden(name)
    L, order, bc = int(z["L"]), int(z["order"]), str(z["bc"])
    H = M.heisenberg(L, 0.5, *z["consts"], field=z["field"], bc=bc)
    assert np.max(np.abs(H - z["H"])) <= 1e-14
    U = M.construct_propagators(H, complex(z["dt"]), order=order, bc=bc)
    assert U.shape == z["U"].shape
    assert np.max(np.abs(U - z["U"])) <= 1e-13


def schedule_errors(z, gam_lam_of_step):
    """Worst deviation from the reference's per-step outputs: truncation error, <Sz_i>
    and entropy by the reference's local formulas (OneSiteExpVal, ChainEntropy), the
    energy by TotalEnergy's formula (open chains) or, at full rank, from the state vector
    with the edge bond (periodic chains; a truncated periodic chain is left out of
    canonical form by the swaps, where TotalEnergy's edge term is not a plain
    expectation value), and the final lambdas."""
    L, chi, bc = int(z["L"]), int(z["chi"]), str(z["bc"])
    full_rank = chi >= 2 ** (L // 2)
    worst = {}
    for k, (tr, gam, lam) in enumerate(gam_lam_of_step(), start=1):
        st = OracleState(gam, lam)
        e = {"trunc": abs(tr - z["trunc"][k]), "sz": np.max(np.abs(st.one_site_expval(z["szop"]) - z["sz"][k])),
             "S": np.max(np.abs(st.chain_entropy() - z["S"][k]))}
        if bc == "O":
            e["E"] = abs(st.total_energy(z["H"]) - z["energy"][k])
        elif full_rank:
            E, _, _ = vector_observables(mps_to_vector(gam, lam), z["H"], z["szop"], pbc=True)
            e["E"] = abs(E - z["energy"][k])
        for key, v in e.items():
            worst[key] = max(worst.get(key, 0.0), float(v))
    worst["lambda"] = float(np.max(np.abs(lam - z["lam"])))
    return worst


@pytest.mark.parametrize("name", CASES)
def test_oracle_schedules_reproduce_reference_trajectories(name):
    z = load_golden(name)
    order, bc = int(z["order"]), str(z["bc"])
    st = OracleState.product(z["coeffs"], int(z["chi"]))

    def steps():
        for _ in range(int(z["nsteps"])):
            tr = st.trotter_step_sched(z["U"], order, bc)
            yield tr, st.gam.copy(), st.lam.copy()

    worst = schedule_errors(z, steps)
    assert max(worst.values()) < 1e-10, worst


def _dense(G, a, b, L, d):
    """d^L x d^L matrix of a two-site operator on sites a, b."""
    n = d ** L
    out = np.zeros((n, n), np.complex128)
    g = G.reshape(d, d, d, d)
    for col in range(n):
        e = np.zeros(n, np.complex128)
        e[col] = 1.0
        p = e.reshape((d,) * L)
        q = np.moveaxis(np.tensordot(g, p, axes=([2, 3], [a, b])), [0, 1], [a, b])
        out[:, col] = q.reshape(-1)
    return out


@pytest.mark.parametrize("order,bc,L", [(2, "O", 6), (5, "O", 6), (2, "P", 6), (2, "P", 7), (5, "P", 6),
                                        (5, "P", 7)])
def test_schedules_have_their_order_at_full_rank(order, bc, L):
    d, chi = 2, 2 ** (L // 2)
    rng = np.random.default_rng(7)
    H = M.heisenberg(L, 0.5, 1.0, 0.6, 1.2, field=rng.uniform(-1, 1, L), bc=bc)
    Hfull = sum(_dense(H[b], b, b + 1, L, d) for b in range(L - 1))
    if bc == "P":
        Hfull = Hfull + _dense(H[L - 1], 0, L - 1, L, d)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    psi0 = c[0]
    for s in range(1, L):
        psi0 = np.multiply.outer(psi0, c[s])
    errs = []
    for dt in (0.2, 0.1):
        st = OracleState.product(c, chi)
        assert st.trotter_step_sched(M.construct_propagators(H, dt, order=order, bc=bc), order, bc) < 1e-12
        psi = mps_to_vector(st.gam, st.lam).reshape(-1)
        exact = scipy.linalg.expm(-1j * dt * Hfull) @ psi0.reshape(-1)
        errs.append(np.linalg.norm(psi - exact))
    ratio = errs[0] / errs[1]
    expect = 8.0 if order == 2 else 32.0
    assert expect / 2 < ratio < expect * 2, (errs, ratio)


def test_edge_gate_through_swaps_equals_direct_gate():
    """One 2nd-order periodic step with every bulk gate the identity leaves only the edge
    gate, applied through the swaps: it must equal that gate on sites (L, 1) directly."""
    L, d = 7, 2
    rng = np.random.default_rng(5)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    st = OracleState.product(c, 2 ** (L // 2))
    G = scipy.linalg.expm(-1j * 0.3 * (lambda A: A + A.conj().T)(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))))
    U = np.array([np.eye(4, dtype=np.complex128)] * (L - 1) + [G])
    st.trotter_step_sched(U, 2, "P")
    psi = mps_to_vector(st.gam, st.lam)
    psi0 = c[0]
    for s in range(1, L):
        psi0 = np.multiply.outer(psi0, c[s])
    direct = _dense(G, 0, L - 1, L, d) @ psi0.reshape(-1)  # left index on site 1, right on site L
    assert np.max(np.abs(psi.reshape(-1) - direct)) < 1e-12


@pytest.mark.parametrize("name", CASES)
def test_library_gate_tables_match_reference(name):
    """tebd_construct_propagators (host-only) for every schedule against the reference's U."""
    from tebd_amd import construct_propagators

    z = load_golden(name)
    U = construct_propagators(z["H"], complex(z["dt"]), order=int(z["order"]), bc=str(z["bc"]))
    assert U.shape == z["U"].shape
    assert np.max(np.abs(U - z["U"])) <= 1e-13
