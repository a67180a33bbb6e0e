"""Gates supplied as files, as the reference's Extensible wrapper loads them
(Propagator_*.nml, Case_Studies/Extensible_Wrapper/TEBDSolve.f90:172-206): the reference's
own data files (tests/golden/propagators/) read column-major into 16 x 16 Liouville-space
propagators (non-unitary), assembled into a gate table, and run through the two-site update.

CPU: the files are the propagators of one generator -- the 0.01 bulk propagator is the
square of the 0.005 one to the files' precision, which a wrong read (rows for columns,
real for imaginary) of either file keeps only if it is wrong the same way, so the reads are
also checked against the raw numbers; the table follows TEBDSolve's assignment.
GPU: the device propagates with these gates exactly as the FP64 oracle does.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.oracle import OracleState
from tebd_amd import Context, models as M

P = os.path.join(GOLDEN, "propagators")


def prop(name):
    return M.read_propagator(os.path.join(P, f"Propagator_{name}_REAL.nml"),
                             os.path.join(P, f"Propagator_{name}_IMAG.nml"))


def test_propagator_files_read_column_major():
    U = prop("0.01_bulk")
    raw = np.array(open(os.path.join(P, "Propagator_0.01_bulk_REAL.nml")).read().split(), float)
    assert U.shape == (16, 16)
    assert U[1, 0].real == raw[1] and U[0, 1].real == raw[16]  # Fortran's element order
    half = prop("0.005_bulk")
    assert np.max(np.abs(half @ half - U)) < 1e-12


def test_file_gate_table_follows_tebdsolve():
    left, odd, even, right = prop("0.005_left"), prop("0.005_bulk"), prop("0.01_bulk"), prop("0.005_right")
    U = M.file_gate_table(8, left, odd, even, right)
    assert np.array_equal(U[0], left) and np.array_equal(U[6], right)
    assert all(np.array_equal(U[i - 1], odd) for i in (3, 5))
    assert all(np.array_equal(U[i - 1], even) for i in (2, 4, 6))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "batched"])
def test_device_propagates_file_gates_like_the_oracle(mode):
    L, d, chi = 8, 4, 32
    U = M.file_gate_table(L, prop("0.005_left"), prop("0.005_bulk"), prop("0.01_bulk"), prop("0.005_right"))
    rng = np.random.default_rng(41)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    ref = OracleState.product(c, chi)
    op = (lambda A: A + A.T)(rng.normal(size=(d, d)))
    print(f"[gpu] file gates mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode) as ctx:
        ctx.upload(*M.product_state(c, chi))
        ctx.set_gates(U)
        for _ in range(10):
            t_dev = ctx.trotter_step()
            t_ref = ref.trotter_step(U)
            assert abs(t_dev - t_ref) < 1e-10
        g, lam = ctx.download(chi)
    dev = OracleState(g, lam)
    assert np.max(np.abs(lam - ref.lam)) < 1e-10
    assert np.max(np.abs(dev.one_site_expval(op) - ref.one_site_expval(op))) < 1e-9
    assert np.max(np.abs(dev.chain_entropy() - ref.chain_entropy())) < 1e-9
