"""Checkpoint files in the reference's formats (RecordLambdas / RecordGammas /
readGammaLambda, ToolsCore_v_2.0/io_module.f90:191-373), pinned by the files the
reference itself wrote (tests/golden/ckpt_l4/, made by the golden driver's mode 4):

* reading them gives the reference's state (kind 'B' bit for bit, kind 'S' to its 15
  digits);
* writing that state reproduces both of the reference's files byte for byte;
* missing or short files are errors, not a silent success (the reference's
  RestartParallel reports success after a mismatch).
GPU: a context checkpointed and restarted continues bit for bit like one that was not.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from tebd_amd import Context, TEBDError, models as M, read_state_files, write_state_files

CK = os.path.join(GOLDEN, "ckpt_l4")


def ref_state():
    with np.load(os.path.join(CK, "state.npz")) as z:
        return {k: z[k] for k in z.files}


def files(kind):
    return os.path.join(CK, f"ckpt_lambdas_{kind}.dat"), os.path.join(CK, f"ckpt_gammas_{kind}.dat")


def test_reading_reference_checkpoints():
    z = ref_state()
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    g, l = read_state_files(L, d, chi, *files("B"), kind="B")
    assert np.array_equal(g, z["gam"]) and np.array_equal(l, z["lam"])
    g, l = read_state_files(L, d, chi, *files("S"), kind="S")
    assert np.max(np.abs(g - z["gam"])) <= 1e-15 * np.max(np.abs(z["gam"])) * 10
    assert np.max(np.abs(l - z["lam"])) <= 1e-15


@pytest.mark.parametrize("kind", ["S", "B"])
def test_writing_reproduces_reference_files(tmp_path, kind):
    z = ref_state()
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    g, l = read_state_files(L, d, chi, *files("B"), kind="B")
    lp, gp = str(tmp_path / "l.dat"), str(tmp_path / "g.dat")
    write_state_files(g, l, lp, gp, kind=kind)
    rl, rg = files(kind)
    assert open(lp, "rb").read() == open(rl, "rb").read()
    assert open(gp, "rb").read() == open(rg, "rb").read()


def test_e30_formatting_edge_cases(tmp_path):
    """Fortran E30.15: rounding to 15 digits, negative values and zeros, and exponents
    beyond +-99 written without the E; each read back to within its 15 digits."""
    vals = np.array([[1.0, -0.5, 0.0, 9.999999999999999e-1, 1e-120, -2.5e150, 3.0e-5, 123456.789]])
    L, d, chi = 1, 1, 8
    gam = np.zeros((1, chi, 1, chi), np.complex128)
    lam = np.zeros((2, chi))
    lam[0] = vals[0]
    lp, gp = str(tmp_path / "l.dat"), str(tmp_path / "g.dat")
    write_state_files(gam, lam, lp, gp, kind="S")
    line = open(lp).readline()
    fields = [line[i:i + 30] for i in range(0, 240, 30)]
    assert fields[0] == "         0.100000000000000E+01"
    assert fields[1] == "        -0.500000000000000E+00"
    assert fields[2] == "         0.000000000000000E+00"
    assert fields[3] == "         0.100000000000000E+01"
    assert fields[4] == "         0.100000000000000-119"
    assert fields[5] == "        -0.250000000000000+151"
    g, l = read_state_files(L, d, chi, lp, gp, kind="S")
    np.testing.assert_allclose(l[0], vals[0], rtol=1e-15, atol=0)


def test_missing_or_short_files_are_errors(tmp_path):
    z = ref_state()
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    with pytest.raises(TEBDError):
        read_state_files(L, d, chi, str(tmp_path / "none.dat"), files("S")[1], kind="S")
    short = tmp_path / "short.dat"
    short.write_text("".join(open(files("S")[1]).readlines()[:-1]))
    with pytest.raises(TEBDError):
        read_state_files(L, d, chi, files("S")[0], str(short), kind="S")
    with pytest.raises(TEBDError):  # one site more than the files hold
        read_state_files(L + 1, d, chi, *files("S"), kind="S")


@pytest.mark.gpu
def test_checkpoint_restart_continues_bit_for_bit(tmp_path):
    z = ref_state()
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    H = M.heisenberg(L, 0.5, 1.0, 1.0, 0.5, field=np.random.default_rng(32).uniform(-0.5, 0.5, L))
    U = M.construct_propagators(H, 0.05)
    print("[gpu] checkpoint / restart", flush=True)
    with Context(L, d, chi, mode="batched") as a:
        a.upload(z["gam"], z["lam"])
        a.set_gates(U)
        a.trotter_step()
        a.checkpoint(tmp_path / "l.dat", tmp_path / "g.dat", kind="B")
        for _ in range(3):
            a.trotter_step()
        ga, la = a.download(chi)
    with Context(L, d, chi, mode="batched") as b:
        b.restart(tmp_path / "l.dat", tmp_path / "g.dat", chi, kind="B")
        b.set_gates(U)
        for _ in range(3):
            b.trotter_step()
        gb, lb = b.download(chi)
        with pytest.raises(TEBDError):
            b.restart(tmp_path / "l.dat", tmp_path / "g.dat", chi + 1, kind="B")
    assert np.array_equal(ga, gb) and np.array_equal(la, lb)
