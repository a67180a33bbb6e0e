"""Device path vs the reference's own outputs (tests/golden) and vs the FP64 oracle.

Tolerance: every observable (energy, <Sz_i>, <n_i>, entanglement entropy in
log_d, cumulative truncation error) within 1e-9 of the reference.  The device
keeps the reference's truncation rule (normalised |s| > 1e-15, at most chi).
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import golden_obs, max_obs_err, observables
from oracle.oracle import OracleState
from tebd_amd import Context, models as M

pytestmark = pytest.mark.gpu

TOL = 1e-9
RTP = ["xxz_l32_chi32", "xxz_l16_chi8", "bh_l12_chi24", "fh_l10_chi32", "fh_l8_chi8", "s1_l8_chi12"]


def run_device(z, mode, check_every=10):
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    ctx = Context(L, d, chi, mode=mode)
    gam, lam = M.product_state(z["coeffs"], chi)
    ctx.upload(gam, lam)
    ctx.set_gates(z["U"])
    worst, cum, cum_ref = {}, 0.0, 0.0
    n = int(z["nsteps"])
    canon = int(z["canon"])
    for k in range(1, n + 1):
        cum += ctx.trotter_step()
        if canon:
            ctx.canonical_form_all()
        cum_ref += z["trunc"][k]
        if k % check_every == 0 or k == n:
            g, l = ctx.download(chi)
            err = max_obs_err(observables(g, l, z["H"], z["szop"], z["nop"]), golden_obs(z, k))
            err["trunc"] = abs(cum - cum_ref)
            for key, v in err.items():
                worst[key] = max(worst.get(key, 0.0), v)
    dims = ctx.bond_dims()
    ctx.close()
    return worst, dims


@pytest.mark.parametrize("mode", ["fp64", "batched", "learned", "batched,learned"])
@pytest.mark.parametrize("name", RTP)
def test_trajectory_matches_reference(name, mode):
    z = load_golden(name)
    print(f"[gpu] trajectory {name} mode={mode}", flush=True)
    worst, dims = run_device(z, mode)
    assert max(worst.values()) < TOL, worst
    assert dims.max() <= int(z["chi"]) and dims[0] == 1 and dims[-1] == 1


@pytest.mark.parametrize("mode", ["fp64", "batched", "learned", "batched,learned"])
def test_imaginary_time_with_canonical_form_matches_reference(mode):
    """TrotterStep with exp(-dtau H) gates + CanonicalFormAll every step
    (propagation_module.f90:741-743) against the reference's trajectory."""
    z = load_golden("itp_l10_chi8")
    print(f"[gpu] itp_l10_chi8 mode={mode}", flush=True)
    worst, dims = run_device(z, mode, check_every=5)
    assert max(worst.values()) < TOL, worst


def test_canonical_form_restores_orthonormality():
    """After random non-unitary updates, CanonicalFormAll leaves A = lambda_l Gamma
    left-orthonormal and B = Gamma lambda_r right-orthonormal on every bond, and
    keeps the state (two-site density matrices) unchanged up to normalisation."""
    print("[gpu] canonical form invariants", flush=True)
    rng = np.random.default_rng(5)
    L, d, chi = 8, 2, 8
    G = np.eye(4) + 0.4 * (rng.normal(size=(L - 1, 4, 4)) + 1j * rng.normal(size=(L - 1, 4, 4)))
    ctx = Context(L, d, chi)
    ctx.upload(*M.product_state(rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d)), chi))
    ctx.set_gates(G)
    for _ in range(3):
        ctx.trotter_step()
    ref = OracleState(*ctx.download(chi))
    ref.canonical_form_all()
    ctx.canonical_form_all()
    g, l = ctx.download(chi)
    dims = ctx.bond_dims()
    for s in range(L):
        cl, cr = dims[s], dims[s + 1]
        A = g[s, :cr, :, :cl] * l[s, None, None, :cl]          # [b, i, a] * lambda_l(a)
        Ml = np.einsum("bia,cia->bc", A.conj(), A)
        B = g[s, :cr, :, :cl] * l[s + 1, :cr, None, None]      # Gamma(a,i,b) lambda_r(b)
        Mr = np.einsum("bia,bic->ac", B, B.conj())
        keep_l = l[s + 1, :cr] > 1e-6
        keep_r = l[s, :cl] > 1e-6
        np.testing.assert_allclose(Ml[np.ix_(keep_l, keep_l)], np.eye(keep_l.sum()), atol=1e-10)
        np.testing.assert_allclose(Mr[np.ix_(keep_r, keep_r)], np.eye(keep_r.sum()), atol=1e-10)
    np.testing.assert_allclose(l, ref.lam, atol=1e-10)
    np.testing.assert_allclose(OracleState(g, l).two_site_rho(), ref.two_site_rho(), atol=1e-10)


def test_single_two_site_op_on_random_state():
    z = load_golden("two_site_op_l4")
    print("[gpu] two_site_op_l4", flush=True)
    L, d, chi, link = int(z["L"]), int(z["d"]), int(z["chi"]), int(z["link"])
    ctx = Context(L, d, chi, cutoff=0.0)
    ctx.upload(z["gam_in"], z["lam_in"])
    t = ctx.two_site_op(link, z["U"][link - 1])
    g, l = ctx.download(chi)
    assert abs(t - z["trunc"][0]) < 1e-13
    np.testing.assert_allclose(l, z["lam"], atol=1e-13)
    # gauge-invariant: lambda_l Gamma_l lambda_c Gamma_r lambda_r on the updated bond
    rec = lambda G, lm: np.einsum("a,bia,b,cjb,c->aijc", lm[link - 1], G[link - 1], lm[link], G[link], lm[link + 1])
    np.testing.assert_allclose(rec(g, l), rec(z["gam"], z["lam"]), atol=1e-13)


@pytest.mark.parametrize("d,chi", [(2, 1), (3, 7), (5, 12)])
def test_random_gates_against_oracle(d, chi):
    """Non-unitary random gates, uneven bonds, truncation active: device vs oracle."""
    print(f"[gpu] random gates d={d} chi={chi}", flush=True)
    rng = np.random.default_rng(100 + d)
    L = 7
    G = rng.normal(size=(L - 1, d * d, d * d)) + 1j * rng.normal(size=(L - 1, d * d, d * d))
    G = G / np.linalg.norm(G, axis=(1, 2), keepdims=True) * d + np.eye(d * d)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    ctx = Context(L, d, chi, cutoff=0.0)
    ctx.upload(*M.product_state(c, chi))
    ctx.set_gates(G)
    ref = OracleState.product(c / np.linalg.norm(c, axis=1, keepdims=True), chi)
    # the reference normalises theta inside every update: an unnormalised product
    # state and the normalised one give the same Vidal state after one update
    for _ in range(3):
        t_dev = ctx.trotter_step()
        t_ref = ref.trotter_step(G)
        assert abs(t_dev - t_ref) < 1e-10
    g, l = ctx.download(chi)
    H = M.heisenberg(L, 0.5 * (d - 1)) if d <= 3 else np.zeros((L - 1, d * d, d * d), complex)
    np.testing.assert_allclose(l, ref.lam, atol=1e-10)
    rho_dev = OracleState(g, l).two_site_rho()
    np.testing.assert_allclose(rho_dev, ref.two_site_rho(), atol=1e-10)


def test_nan_input_is_reported():
    print("[gpu] nan input", flush=True)
    from tebd_amd import TEBDError
    L, d = 4, 2
    ctx = Context(L, d, 4)
    c = np.ones((L, d), complex)
    c[1, 0] = np.nan
    ctx.upload(*M.product_state(c, 4))
    ctx.set_gates(M.construct_propagators(M.heisenberg(L), 0.1))
    with pytest.raises(TEBDError) as e:
        ctx.trotter_step()
    assert e.value.code == -4


def test_deterministic():
    print("[gpu] determinism", flush=True)
    z = load_golden("fh_l10_chi32")
    outs = []
    for _ in range(2):
        ctx = Context(int(z["L"]), int(z["d"]), int(z["chi"]))
        ctx.upload(*M.product_state(z["coeffs"], int(z["chi"])))
        ctx.set_gates(z["U"])
        for _ in range(5):
            ctx.trotter_step()
        outs.append(ctx.download())
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
