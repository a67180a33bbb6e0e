"""The host gate builders (tebd_amd.models) against the bond Hamiltonians and Trotter
gates the reference itself produced for every golden job.

The reference's HamiltonianHeisenberg (heisenberg_module.f90:143-174, spin 1/2 and
spin 1, site-dependent field split half and half between bonds),
HamiltonianBoseHubbardJVector (bose_hubbard_module.f90:275-345, bond-dependent
hopping), HamiltonianHubbard (fermi_hubbard_module.f90:289-356, Jordan-Wigner
operators of CreateFermiSops) and NEWConstructPropagators / matrix_exponential_c
(TEBDtools_module.f90:176-222,1079-1101; real time exp(-i dt H), imaginary time with
dt = -i dtau) were dumped as H(i), U(i) by the golden driver.  The builders are
checked element by element, and each gate against scipy's expm of the same H as a
library restatement of the exponential.
"""
import numpy as np
import pytest
import scipy.linalg

from conftest import load_golden
from helpers import hamiltonian_of
from tebd_amd import models as M

JOBS = ["xxz_l32_chi32", "xxz_l16_chi8", "bh_l12_chi24", "fh_l10_chi32", "fh_l8_chi8", "s1_l8_chi12",
        "itp_l10_chi8"]


@pytest.mark.parametrize("name", JOBS)
def test_bond_hamiltonians_match_the_reference(name):
    z = load_golden(name)
    H = np.asarray(hamiltonian_of(z))
    assert H.shape == z["H"].shape
    assert np.max(np.abs(H - z["H"])) <= 1e-14
    # every bond Hamiltonian is Hermitian
    assert np.max(np.abs(H - np.conj(np.transpose(H, (0, 2, 1))))) <= 1e-15


@pytest.mark.parametrize("name", JOBS)
def test_trotter_gates_match_the_reference(name):
    z = load_golden(name)
    dt = complex(z["dt"])
    U = np.asarray(M.construct_propagators(hamiltonian_of(z), dt))
    assert U.shape == z["U"].shape
    assert np.max(np.abs(U - z["U"])) <= 1e-13
    # 2nd-order OBC schedule: odd bonds (1-based) take dt/2, even bonds dt
    H = z["H"]
    for b in range(H.shape[0]):
        t = dt / 2 if b % 2 == 0 else dt
        ref = scipy.linalg.expm(-1j * t * H[b])
        assert np.max(np.abs(U[b] - ref)) <= 1e-12, b
    if dt.imag == 0.0:  # real time: unitary gates
        eye = np.eye(U.shape[1])
        for b in range(U.shape[0]):
            assert np.max(np.abs(U[b].conj().T @ U[b] - eye)) <= 1e-13


def test_two_site_op_fixture_gates():
    z = load_golden("two_site_op_l4")
    L = int(z["L"])
    H = M.heisenberg(L, 0.5, *z["consts"], field=z["field"])
    assert np.max(np.abs(np.asarray(H) - z["H"])) <= 1e-14
    U = M.construct_propagators(H, complex(z["dt"]))
    assert np.max(np.abs(np.asarray(U) - z["U"])) <= 1e-13


@pytest.mark.parametrize("name", JOBS)
def test_library_propagators_match_the_reference(name):
    """tebd_construct_propagators (host-only, no device call): the library's own
    Hermitian exponential against the reference's U(i)."""
    from tebd_amd import construct_propagators

    z = load_golden(name)
    U = construct_propagators(z["H"], complex(z["dt"]))
    assert np.max(np.abs(U - z["U"])) <= 1e-13


def test_library_propagators_on_degenerate_and_random_hamiltonians():
    """Closed forms: H = 0 gives the identity; a diagonal H gives exp(-i t diag); a
    random Hermitian H agrees with scipy's expm (odd bonds dt/2, even bonds dt)."""
    from tebd_amd import construct_propagators

    rng = np.random.default_rng(3)
    n = 16
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Hr = A + A.conj().T
    D = np.diag(rng.normal(size=n)).astype(np.complex128)
    Hs = np.stack([np.zeros((n, n), np.complex128), D, Hr, Hr])
    for dt in (0.07, -0.05j, 0.1 - 0.02j):
        U = construct_propagators(Hs, dt)
        for b in range(4):
            t = dt / 2 if b % 2 == 0 else dt
            assert np.max(np.abs(U[b] - scipy.linalg.expm(-1j * t * Hs[b]))) <= 1e-12 * max(1.0, np.abs(U[b]).max())
