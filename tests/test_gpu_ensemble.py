"""Ensembles (tebd_create_ensemble): several independent chains -- disorder
realisations, parameter sweeps -- updated by the same kernel launches.

Each chain of an ensemble must evolve exactly as the same chain in a context of its
own: the bonds between chains are never touched and every kernel works per bond.
Tolerance 1e-12 on every observable for the exact path (the two runs take the same
floating-point operations per bond).  On the learned-rank path the sketch columns are
drawn per bond position, so they are the same in both runs, but which bonds share a
batch decides whether a small full-SVD bond is QR-preconditioned, so the two runs may
differ at rounding level; that case is held to 1e-10 and runs without a bond-dimension
cap: a cap that cuts through a degenerate multiplet of Schmidt values (frequent in
the Hubbard chain's spin-symmetric spectra) makes the kept subspace, and observables
at the order of the discarded weight, depend on rounding.
"""
import numpy as np
import pytest

from tebd_amd import Context, Ensemble, models as M

pytestmark = pytest.mark.gpu


def run_single(L, d, chi, mode, U, coeffs, steps, canon=False):
    ctx = Context(L, d, chi, mode=mode)
    ctx.upload(*M.product_state(coeffs, chi))
    ctx.set_gates(U)
    trunc = 0.0
    for _ in range(steps):
        trunc += ctx.trotter_step()
        if canon:
            ctx.canonical_form_all()
    out = ctx.download(chi), ctx.bond_dims(), trunc
    ctx.close()
    return out


def ensemble_vs_single(L, d, chi, mode, Us, coeffs, steps, H, ops, canon=False):
    n = len(Us)
    ens = Ensemble(n, L, d, chi, mode=mode)
    states = [M.product_state(c, chi) for c in coeffs]
    ens.upload_chains(np.stack([s[0] for s in states]), np.stack([s[1] for s in states]))
    ens.set_gates(np.stack(Us))
    trunc = 0.0
    for _ in range(steps):
        trunc += ens.trotter_step()
        if canon:
            ens.canonical_form_all()
    gams, lams = ens.download_chains(chi)
    dims = ens.bond_dims()
    E_dev = ens.energy_chains(np.stack(H))
    one_dev = ens.one_site_expval_chains(ops)
    ens.close()
    worst = 0.0
    tr_single = 0.0
    for r in range(n):
        (g, l), dr, tr = run_single(L, d, chi, mode, Us[r], coeffs[r], steps, canon)
        tr_single += tr
        np.testing.assert_array_equal(dims[r * L:r * L + L + 1], dr)
        worst = max(worst, float(np.abs(lams[r] - l).max()))
        ctx = Context(L, d, chi, mode=mode)
        ctx.upload(g, l)
        worst = max(worst, abs(ctx.energy(H[r]) - E_dev[r]), float(np.abs(ctx.one_site_expval(ops) - one_dev[r]).max()))
        ctx.close()
    worst = max(worst, abs(trunc - tr_single))
    return worst, dims


def test_disorder_ensemble_real_time_matches_single_chains():
    """Three XXZ chains with their own random fields (h_i ~ U[-1, 1]), truncation active."""
    print("[gpu] ensemble xxz L=12 chi=16 x3 batched", flush=True)
    L, d, chi, n = 12, 2, 16, 3
    rng = np.random.default_rng(7)
    H = [M.heisenberg(L, 0.5, 1.0, 1.0, 1.0, field=rng.uniform(-1.0, 1.0, L)) for _ in range(n)]
    Us = [M.construct_propagators(h, 0.05) for h in H]
    coeffs = [M.neel(L, d, 0, 1)] * n
    sz = M.spin_ops(0.5)[2]
    worst, dims = ensemble_vs_single(L, d, chi, "batched", Us, coeffs, 30, H, sz)
    print(f"[gpu] ensemble xxz worst {worst:.2e}", flush=True)
    assert worst < 1e-12
    assert dims.max() == chi
    assert all(dims[r * L] == 1 for r in range(n + 1))


def test_parameter_ensemble_learned_rank_matches_single_chains():
    """Three Fermi-Hubbard chains (U = 2, 4, 6) on the learned-rank path (bond dimensions
    up to 4^4 = 256 = no cap, so bonds of 4 chi_l > 2 x sketch width are sketched)."""
    print("[gpu] ensemble fh L=8 chi=256 x3 batched,learned", flush=True)
    L, d, chi = 8, 4, 256
    H = [M.fermi_hubbard(L, 1.0, u) for u in (2.0, 4.0, 6.0)]
    Us = [M.construct_propagators(h, 0.02) for h in H]
    coeffs = [M.neel(L, d, 1, 2)] * 3
    nop = M.fermion_ops(2)[2]
    worst, _ = ensemble_vs_single(L, d, chi, "batched,learned", Us, coeffs, 20, H, nop)
    print(f"[gpu] ensemble fh learned worst {worst:.2e}", flush=True)
    assert worst < 1e-10


def test_imaginary_time_ensemble_with_canonical_form_matches_single_chains():
    """Four disorder realisations of the Heisenberg chain (h_i ~ U[-1, 1]) in imaginary
    time with CanonicalFormAll after every step, as ImagTimeProp does
    (propagation_module.f90:741-743)."""
    print("[gpu] ensemble itp heisenberg L=10 chi=8 x4 batched", flush=True)
    L, d, chi, n = 10, 2, 8, 4
    rng = np.random.default_rng(11)
    H = [M.heisenberg(L, 0.5, 1.0, 1.0, 1.0, field=rng.uniform(-1.0, 1.0, L)) for _ in range(n)]
    Us = [M.construct_propagators(h, -1j * 0.05) for h in H]
    coeffs = [M.neel(L, d, 0, 1)] * n
    sz = M.spin_ops(0.5)[2]
    worst, _ = ensemble_vs_single(L, d, chi, "batched", Us, coeffs, 20, H, sz, canon=True)
    print(f"[gpu] ensemble itp worst {worst:.2e}", flush=True)
    assert worst < 1e-12


def test_ensemble_refuses_bonds_between_chains():
    ens = Ensemble(2, 4, 2, 4, mode="batched")
    gam, lam = M.product_state(M.neel(8, 2, 0, 1), 4)
    ens.upload(gam, lam)
    ens.set_gates(np.stack([M.construct_propagators(M.heisenberg(4), 0.1)] * 2))
    with pytest.raises(Exception):
        ens.two_site_op(4, np.eye(4))
    ens.two_site_op(3, np.eye(4))
    ens.close()
