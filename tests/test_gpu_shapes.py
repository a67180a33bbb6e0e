"""Bond updates at the configurations' real bond-matrix sizes against the FP64 oracle.

Shapes (theta is (d chi_l) x (d chi_r) on a saturated bond):
  * Fermi-Hubbard d=4, chi=256: 1024 x 1024 (the metric's shape), a full Trotter step
    of a 20-site chain whose bonds 4..16 are at chi=256, in every batching mode,
    with and without lanes, exact and learned-rank;
  * spin-1 Heisenberg d=3, chi=512: 1536 x 1536 (the learned-rank configuration),
    three bond updates around the centre of a 14-site chain;
  * spin-1/2 Heisenberg d=2, chi=1024: 2048 x 2048 (the imaginary-time ensemble's
    bond), three bond updates around the centre of a 22-site chain.
Bond matrices with more than 1024 rows take the pivoted QR's global-memory panel.

The warm states come from random product states (seeded) evolved by the exact batched
path until every bond whose cap is chi has reached it; the Schmidt spectra are then
graded as in a real run.  The oracle (reference algorithm, padded chi, LAPACK ZGESVD;
its checker mode forms theta by ZGEMM and lets ZGESVD block) applies the same
updates.  Compared: every lambda within LAMBDA_TOL and the observables of both states
(bond energies, <n_i> or <Sz_i>, entanglement entropy in log_d, evaluated by the FP64
device kernels on both) within OBS_TOL.
"""
import os

import numpy as np
import pytest

from oracle.oracle import OracleState, checker_speed
from tebd_amd import Context, models as M

pytestmark = pytest.mark.gpu

LAMBDA_TOL = 1e-10
OBS_TOL = 1e-9
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def random_product(L, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    return c / np.linalg.norm(c, axis=1, keepdims=True)


def caps(L, d, chi):
    return np.array([min(d ** b, d ** (L - b), chi) for b in range(L + 1)])


def warm_state(H, L, d, chi, dt, seed, max_steps=60):
    U = M.construct_propagators(H, dt)
    cap = caps(L, d, chi)
    with Context(L, d, chi, mode="batched") as ctx:
        ctx.set_lanes(1)
        ctx.upload(*M.product_state(random_product(L, d, seed), chi))
        ctx.set_gates(U)
        for k in range(max_steps):
            ctx.trotter_step()
            dims = ctx.bond_dims()
            if np.all(dims[cap == chi] == chi):
                break
        assert np.all(ctx.bond_dims()[cap == chi] == chi), ctx.bond_dims()
        gam, lam = ctx.download(chi)
    return gam, lam, U


def local_ops(d):
    if d == 4:
        _, _, ntot, sz, _ = M.fermion_ops(2)
        return {"n": ntot, "sz": sz}
    return {"sz": M.spin_ops((d - 1) / 2)[2]}


def device_observables(gam, lam, H):
    L, chi, d, _ = gam.shape
    with Context(L, d, chi, mode="fp64") as ev:
        ev.upload(gam, lam)
        out = {"Eb": ev.expval_two_site(H).real, "S": ev.chain_entropy()}
        for k, op in local_ops(d).items():
            out[k] = ev.one_site_expval(op)
    return out


def compare(dev, ref, H):
    errs = {"lambda": float(np.max(np.abs(dev[1] - ref[1])))}
    a, b = device_observables(*dev, H), device_observables(*ref, H)
    for k in a:
        errs[k] = float(np.max(np.abs(a[k] - b[k])))
    return errs


# ---- Fermi-Hubbard d=4 chi=256: a full Trotter step at 1024 x 1024 --------------------

@pytest.fixture(scope="module")
def fh_case():
    L, d, chi = 20, 4, 256
    H = M.fermi_hubbard(L, 1.0, 4.0)
    print("[gpu] warm FH L=20 chi=256 state", flush=True)
    gam, lam, U = warm_state(H, L, d, chi, 0.05, seed=31)
    checker_speed(THREADS)
    try:
        ref = OracleState(gam.copy(), lam.copy())
        tr = ref.trotter_step(U)
    finally:
        checker_speed(0)
    return dict(H=H, U=U, state=(gam, lam), ref=(ref.gam, ref.lam), trunc=tr, L=L, d=d, chi=chi)


@pytest.mark.parametrize("mode,lanes", [("fp64", 1), ("batched", 1), ("batched", 3), ("batched,learned", 1),
                                        ("batched,learned", 3), ("batched,learned,split", 3)])
def test_fh_trotter_step_at_1024_matches_oracle(fh_case, mode, lanes):
    c = fh_case
    print(f"[gpu] FH L=20 chi=256 step mode={mode} lanes={lanes}", flush=True)
    with Context(c["L"], c["d"], c["chi"], mode=mode) as ctx:
        ctx.set_lanes(lanes)
        ctx.upload(*c["state"])
        ctx.set_gates(c["U"])
        tr = ctx.trotter_step()
        st = ctx.stats()
        dev = ctx.download(c["chi"])
    errs = compare(dev, c["ref"], c["H"])
    errs["trunc"] = abs(tr - c["trunc"])
    print(f"[gpu] errors {errs} stats sketches={st['sketch_accepted']} fallbacks={st['fallbacks']} "
          f"lane_splits={st['lane_splits']}", flush=True)
    assert errs["lambda"] < LAMBDA_TOL, errs
    assert max(v for k, v in errs.items() if k != "lambda") < OBS_TOL, errs
    if "learned" in mode:
        assert st["sketch_accepted"] > 0
    if lanes > 1:
        assert st["lane_splits"] > 0


# ---- spin-1 d=3 chi=512 (1536) and spin-1/2 d=2 chi=1024 (2048): bond updates -------

BIG = {
    "spin1_chi512": dict(L=14, d=3, chi=512, spin=1.0, dt=0.1, seed=4),
    "spinhalf_chi1024": dict(L=22, d=2, chi=1024, spin=0.5, dt=0.1, seed=100),
}


@pytest.fixture(scope="module", params=sorted(BIG))
def big_case(request):
    p = dict(BIG[request.param])
    L, d, chi = p["L"], p["d"], p["chi"]
    rng = np.random.default_rng(p["seed"])
    H = M.heisenberg(L, p["spin"], 1.0, 1.0, 1.0, field=rng.uniform(-1.0, 1.0, L))
    print(f"[gpu] warm {request.param} state", flush=True)
    gam, lam, U = warm_state(H, L, d, chi, p["dt"], seed=p["seed"])
    bonds = (L // 2 - 1, L // 2, L // 2 + 1)  # 1-based; the middle one is (d chi) x (d chi)
    checker_speed(THREADS)
    try:
        ref = OracleState(gam.copy(), lam.copy())
        tr = [ref.two_site_op(b, U[b - 1]) for b in bonds]
    finally:
        checker_speed(0)
    p.update(name=request.param, H=H, U=U, state=(gam, lam), ref=(ref.gam, ref.lam), trunc=tr, bonds=bonds)
    return p


@pytest.mark.parametrize("mode", ["fp64", "batched", "batched,learned", "batched,learned,split"])
def test_large_bond_updates_match_oracle(big_case, mode):
    c = big_case
    L, d, chi = c["L"], c["d"], c["chi"]
    print(f"[gpu] {c['name']} bond updates {c['bonds']} mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode) as ctx:
        ctx.upload(*c["state"])
        tr = [ctx.two_site_op(b, c["U"][b - 1]) for b in c["bonds"]]
        st = ctx.stats()
        dims = ctx.bond_dims()
        dev = ctx.download(chi)
    errs = compare(dev, c["ref"], c["H"])
    errs["trunc"] = float(np.max(np.abs(np.array(tr) - np.array(c["trunc"]))))
    print(f"[gpu] errors {errs} sketches={st['sketch_accepted']} fallbacks={st['fallbacks']} "
          f"dims={dims[c['bonds'][0]:c['bonds'][-1] + 1]}", flush=True)
    assert dims[c["bonds"][1]] == chi
    assert errs["lambda"] < LAMBDA_TOL, errs
    assert max(v for k, v in errs.items() if k != "lambda") < OBS_TOL, errs
    if "learned" in mode and c["name"] == "spin1_chi512":
        assert st["sketch_accepted"] > 0


# ---- the imaginary-time ensemble at its full size ------------------------------------

def test_imaginary_time_ensemble_allocates_and_steps_at_full_size():
    """16 disorder realisations of the spin-1/2 Heisenberg chain, L=256, chi_max=1024,
    learned-rank batched mode: the context fits one MI355X (workspace slots budgeted,
    half-sweeps in batches of that many bonds), and one imaginary-time step with
    CanonicalFormAll from random product states keeps every chain normalised and
    lowers every chain's energy."""
    nchain, L, d, chi, dtau = 16, 256, 2, 1024, 0.05
    print("[gpu] ensemble 16 x L=256 chi=1024 allocation and one ITP step", flush=True)
    Hs = [M.heisenberg(L, 0.5, 1.0, 1.0, 1.0, field=np.random.default_rng(100 + r).uniform(-1.0, 1.0, L))
          for r in range(nchain)]
    from tebd_amd import Ensemble

    with Ensemble(nchain, L, d, chi, mode="batched,learned") as ens:
        info = ens.memory_info()
        print(f"[gpu] memory {info}", flush=True)
        assert info["total_bytes"] < 250e9
        assert 1 <= info["slots"] <= nchain * L // 2
        gams = np.zeros((nchain, L, 1, d, 1), np.complex128)
        lams = np.zeros((nchain, L + 1, 1))
        lams[:, :, 0] = 1.0
        for r in range(nchain):
            gams[r, :, 0, :, 0] = random_product(L, d, 100 + r)
        ens.upload_chains(gams, lams)
        ens.set_gates(np.stack([M.construct_propagators(H, -1j * dtau) for H in Hs]))
        e0 = ens.energy_chains(np.stack(Hs))
        ens.trotter_step()
        ens.canonical_form_all()
        e1 = ens.energy_chains(np.stack(Hs))
        lam = ens.lambdas_chains(int(ens.bond_dims().max()))
    norms = np.sum(lam[:, 1:-1] ** 2, axis=2)
    assert np.max(np.abs(norms - 1.0)) < 1e-12
    assert np.all(e1 < e0), (e0, e1)
