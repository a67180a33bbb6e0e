"""Split-operand mode: theta's contraction over the middle bond on the int8 matrix cores
(v_mfma_i32_16x16x64_i8) from exact power-of-two slices of the FP64 operands, with
integer accumulation and FP64 combination; the gate, the lambda scaling and everything
after theta in FP64.

Tolerance: the exact (FP64) modes' 1e-9 against the reference's trajectories
(test_gpu_parity.py).  Five 7-bit slices per operand leave a truncation error of about
2^-35 of the row and column maxima per product term, far below what the reference's
own FP64 rounding and the 1e-9 bound resolve; the single-update test pins the theta
error directly against the FP64 path at 1e-10.
"""
import numpy as np
import pytest

from conftest import load_golden
from test_gpu_parity import RTP, run_device
from tebd_amd import Context, TEBDError

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-9


@pytest.mark.parametrize("mode", ["split", "batched,split", "batched,learned,split"])
@pytest.mark.parametrize("name", RTP)
def test_split_trajectory_matches_reference(name, mode):
    z = load_golden(name)
    print(f"[gpu] split trajectory {name} mode={mode}", flush=True)
    worst, dims = run_device(z, mode)
    print(f"[gpu] {name} {mode} errors {worst}", flush=True)
    assert max(worst.values()) < SPLIT_TOL, worst
    assert dims.max() <= int(z["chi"])


@pytest.mark.parametrize("mode", ["split", "batched,learned,split"])
def test_split_imaginary_time_matches_reference(mode):
    """Imaginary time with CanonicalFormAll every step against the reference."""
    z = load_golden("itp_l10_chi8")
    print(f"[gpu] split itp_l10_chi8 mode={mode}", flush=True)
    worst, dims = run_device(z, mode, check_every=5)
    print(f"[gpu] itp {mode} errors {worst}", flush=True)
    assert max(worst.values()) < SPLIT_TOL, worst


def two_site(z, mode, gate, scale=(1.0, 1.0)):
    """TwoSiteOp on bond `link` of the fixture's state, Gamma_link and Gamma_link+1 scaled."""
    L, d, chi, link = int(z["L"]), int(z["d"]), int(z["chi"]), int(z["link"])
    gam = np.array(z["gam_in"], dtype=np.complex128)
    gam[link - 1] *= scale[0]
    gam[link] *= scale[1]
    ctx = Context(L, d, chi, mode=mode, cutoff=0.0)
    ctx.upload(gam, z["lam_in"])
    ctx.two_site_op(link, gate)
    out = ctx.download(chi)
    ctx.close()
    return out


def block(G, lm, link):
    return np.einsum("a,bia,b,cjb,c->aijc", lm[link - 1], G[link - 1], lm[link], G[link], lm[link + 1])


def test_split_two_site_op_matches_fp64():
    """One update with a random non-unitary operator on a random state (entries of both
    signs and a graded spectrum): lambda and the gauge-invariant block lambda Gamma
    lambda Gamma lambda agree with the FP64 path to 1e-10, and the int8 path really ran
    (the results are not bit-identical)."""
    z = load_golden("two_site_op_l4")
    print("[gpu] split two_site_op_l4", flush=True)
    link = int(z["link"])
    g0, l0 = two_site(z, "fp64", z["U"][link - 1])
    g1, l1 = two_site(z, "split", z["U"][link - 1])
    r0, r1 = block(g0, l0, link), block(g1, l1, link)
    err_l = np.abs(l1 - l0).max()
    err_r = np.abs(r1 - r0).max() / np.abs(r0).max()
    print(f"[gpu] split two-site op: lambda err {err_l:.2e}, block err {err_r:.2e}", flush=True)
    assert err_l < 1e-10 and err_r < 1e-10
    assert err_l > 0.0 or err_r > 0.0


def test_split_two_site_op_with_a_wide_range_of_magnitudes():
    """The same update with Gamma_l scaled by 2^40 and Gamma_r by 2^-45 (and the other way
    round): the slices' power-of-two scales carry any magnitude, and the new lambda is
    normalised, so it is unchanged to rounding."""
    z = load_golden("two_site_op_l4")
    link = int(z["link"])
    _, l0 = two_site(z, "fp64", z["U"][link - 1])
    for sc in ((2.0 ** 40, 2.0 ** -45), (2.0 ** -45, 2.0 ** 40)):
        _, l1 = two_site(z, "split", z["U"][link - 1], sc)
        err = np.abs(l1 - l0).max()
        print(f"[gpu] split Gamma scales {sc}: lambda err {err:.2e}", flush=True)
        assert err < 1e-10


def test_split_and_mixed_are_exclusive():
    with pytest.raises(TEBDError):
        Context(4, 2, 4, mode="mixed,split")


def test_split_update_on_an_evolved_state_matches_fp64():
    """One update of bonds 8 and 16 of an XXZ state evolved 60 steps at chi=32 (Gamma
    entries up to ~1e11 against lambdas down to ~1e-14, the case the middle-bond balance
    is for): the split path agrees with the FP64 path to 1e-12."""
    from tebd_amd import models as M
    z = load_golden("xxz_l32_chi32")
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    ctx = Context(L, d, chi, mode="batched")
    ctx.upload(*M.product_state(z["coeffs"], chi))
    ctx.set_gates(z["U"])
    for _ in range(60):
        ctx.trotter_step()
    gam, lam = ctx.download(chi)
    ctx.close()
    for link in (8, 16):
        out = {}
        for mode in ("fp64", "split"):
            c = Context(L, d, chi, mode=mode, cutoff=0.0)
            c.upload(gam, lam)
            c.two_site_op(link, z["U"][link - 1])
            out[mode] = c.download(chi)
            c.close()
        (g0, l0), (g1, l1) = out["fp64"], out["split"]
        r0, r1 = block(g0, l0, link), block(g1, l1, link)
        err_l = np.abs(l1 - l0).max()
        err_r = np.abs(r1 - r0).max() / np.abs(r0).max()
        print(f"[gpu] split evolved state bond {link}: lambda err {err_l:.2e}, block err {err_r:.2e}", flush=True)
        assert err_l < 1e-12 and err_r < 1e-12


@pytest.mark.parametrize("d", [5, 6, 8])
def test_split_local_dimensions_match_oracle(d):
    """The tile shapes of d = 5 (16 x 16 (alpha, beta) blocks, 80 rows), 6 and 8 (8 x 8
    blocks): four Trotter steps with random two-site unitaries from a random product state
    at chi = 16, against the oracle's state vector."""
    from scipy.linalg import expm
    from helpers import mps_to_vector
    from oracle.oracle import OracleState
    from tebd_amd import models as M
    L, chi = 5, 16
    rng = np.random.default_rng(40 + d)
    U = []
    for _ in range(L - 1):
        X = rng.standard_normal((d * d, d * d)) + 1j * rng.standard_normal((d * d, d * d))
        U.append(expm(-0.3j * (X + X.conj().T)))
    U = np.array(U)
    coeffs = rng.standard_normal((L, d)) + 1j * rng.standard_normal((L, d))
    coeffs /= np.linalg.norm(coeffs, axis=1, keepdims=True)
    ref = OracleState.product(coeffs, chi)
    ctx = Context(L, d, chi, mode="batched,split")
    ctx.upload(*M.product_state(coeffs, chi))
    ctx.set_gates(U)
    for _ in range(4):
        ctx.trotter_step()
        ref.trotter_step(U)
    g, lam = ctx.download(chi)
    ctx.close()
    err_l = np.abs(lam - ref.lam).max()
    err_v = np.abs(mps_to_vector(g, lam) - mps_to_vector(ref.gam, ref.lam)).max()
    print(f"[gpu] split d={d}: lambda err {err_l:.2e}, state err {err_v:.2e}", flush=True)
    assert err_l < 1e-10 and err_v < 1e-10
