"""The number-conserving (U(1) block-sparse) update on the device against the reference
and the oracle (TwoSiteOpNC / TrotterStepNC, local_operations_module.f90:748-1964,
propagation_module.f90:890-1241).

  * the reference's own NC trajectories (tests/golden/nc_*.npz): Bose-Hubbard d=3 at L=8
    chi=16 (truncated across number blocks) and L=6 chi=32, Fermi-Hubbard d=4 with the
    model's numbers at L=6 chi=64 -- every step's energy, <n_i>, <Sz_i>, entropy and
    truncation error within 1e-10, the final state vector, lambdas and labels (the Hubbard
    chain's spin-degenerate values as number-sector spectra), in the per-bond and the
    batched mode;
  * a larger chain against the oracle: Bose-Hubbard L=16 d=3 chi=48, 8 steps, bond
    matrices up to 144 x 144 in up to 13 blocks -- observables within 1e-9;
  * one two-site update through tebd_two_site_op, and the refusals;
  * the other schedules (TrotterStep5thOrderNC, the periodic TrotterStep2ndOrderPBCNC /
    TrotterStep5thOrderPBCNC with SwappingNC) against the reference's own runs
    (tests/golden/nc_sched_*.npz) within 1e-10, labels exactly;
  * CanonicalFormAllNC against the oracle, and ImagTimePropNC against the reference's own
    runs (tests/golden/itp_nc_*.npz): Bose-Hubbard L=8 with chi stages 4 -> 12 (both
    converge, at j=60 and j=100) and Fermi-Hubbard L=6 at full rank (100 steps) -- the same
    judgement steps, the final observables within 1e-9, every link's sector spectra and
    labels.
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import mps_to_vector
from oracle.oracle import OracleState, product_labels
from tebd_amd import Context, Ensemble, TEBDError, models as M
from test_oracle_nc import conserv_of, sector_spectra_err

pytestmark = pytest.mark.gpu

CASES = ["nc_bh_l8", "nc_bh_l6_full", "nc_fh_l6"]


def device_nc(z, mode):
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    c, tot = conserv_of(z), int(z["tot_num"])
    ctx = Context(L, d, chi, mode=mode)
    ctx.upload(*M.product_state(z["coeffs"], chi))
    ctx.set_gates(z["U"])
    ctx.set_conserved(c, tot, labels=M.product_labels(z["coeffs"], chi, c, tot))
    return ctx


@pytest.mark.parametrize("mode", ["fp64", "batched"])
@pytest.mark.parametrize("name", CASES)
def test_device_nc_reproduces_reference(name, mode):
    print(f"[gpu] NC {name} mode={mode}", flush=True)
    z = load_golden(name)
    worst = {}
    with device_nc(z, mode) as ctx:
        for k in range(1, int(z["nsteps"]) + 1):
            tr = ctx.trotter_step()
            g, l = ctx.download(int(z["chi"]))
            st = OracleState(g, l)
            e = {"trunc": abs(tr - z["trunc"][k]), "E": abs(st.total_energy(z["H"]) - z["energy"][k]),
                 "n": np.max(np.abs(st.one_site_expval(z["nop"]) - z["n"][k])),
                 "sz": np.max(np.abs(st.one_site_expval(z["szop"]) - z["sz"][k])),
                 "S": np.max(np.abs(st.chain_entropy() - z["S"][k]))}
            for key, v in e.items():
                worst[key] = max(worst.get(key, 0.0), float(v))
        LL, LR = ctx.labels(int(z["chi"]))
    worst["lambda"] = float(np.max(np.abs(l - z["lam"])))
    worst["psi"] = float(np.max(np.abs(mps_to_vector(g, l) - mps_to_vector(z["gam"], z["lam"]))))
    worst["sectors"] = sector_spectra_err(LL, l, z["labl"], z["lam"])
    print(f"[gpu] {worst}", flush=True)
    assert max(worst.values()) < 1e-10, worst
    if name.startswith("nc_bh"):
        assert np.array_equal(LL, z["labl"]) and np.array_equal(LR, z["labr"])
    assert np.array_equal(np.sort(LL, axis=1), np.sort(z["labl"], axis=1))


def test_device_nc_matches_oracle_on_a_larger_chain():
    L, d, chi, steps = 16, 3, 48, 8
    print(f"[gpu] NC BH L={L} chi={chi} vs oracle", flush=True)
    J = np.random.default_rng(41).uniform(0.8, 1.2, L - 1)
    H = M.bose_hubbard(L, 2, J, 2.0, 0.0, 0.0)
    U = M.construct_propagators(H, 0.05)
    coeffs = np.zeros((L, d), np.complex128)
    coeffs[:, 1] = 1.0
    cv = np.arange(d)
    LL0, LR0 = product_labels(coeffs, chi, cv, L)
    ref = OracleState.product(coeffs, chi)
    ref.set_conserved(LL0, LR0, cv, L)
    nop = np.diag(np.arange(d).astype(float))
    with Context(L, d, chi, mode="batched") as ctx:
        ctx.upload(*M.product_state(coeffs, chi))
        ctx.set_gates(U)
        ctx.set_conserved(cv, L, labels=M.product_labels(coeffs, chi, cv, L))
        worst = 0.0
        for _ in range(steps):
            t_dev = ctx.trotter_step()
            t_ref = ref.trotter_step_nc(U)
            g, l = ctx.download(chi)
            st = OracleState(g, l)
            err = max(abs(t_dev - t_ref), abs(st.total_energy(H) - ref.total_energy(H)),
                      np.max(np.abs(st.one_site_expval(nop) - ref.one_site_expval(nop))),
                      np.max(np.abs(st.chain_entropy() - ref.chain_entropy())))
            worst = max(worst, err)
        LL, _ = ctx.labels(chi)
    print(f"[gpu] worst {worst:.3e}, kept {ctx_dims(l)}", flush=True)
    assert worst < 1e-9
    assert sector_spectra_err(LL, l, ref.LL, ref.lam) < 1e-9
    assert ctx_dims(l).max() == chi  # the cap was reached: truncation across blocks


def ctx_dims(lam):
    return (lam > 0).sum(axis=1)


def test_device_two_site_op_nc_and_refusals():
    z = load_golden("nc_bh_l8")
    L, chi = int(z["L"]), int(z["chi"])
    c, tot = conserv_of(z), int(z["tot_num"])
    ref = OracleState.product(z["coeffs"], chi)
    ref.set_conserved(*product_labels(z["coeffs"], chi, c, tot), c, tot)
    with device_nc(z, "batched") as ctx:
        for link in (4, 3, 5):
            t1 = ctx.two_site_op(link, z["U"][link - 1])
            t2 = ref.two_site_op_nc(link, z["U"][link - 1])
            assert abs(t1 - t2) < 1e-12
        g, l = ctx.download(chi)
        assert np.max(np.abs(l - ref.lam)) < 1e-12
        assert np.max(np.abs(mps_to_vector(g, l) - mps_to_vector(ref.gam, ref.lam))) < 1e-12
        LL, LR = ctx.labels(chi)
        assert np.array_equal(LL, ref.LL) and np.array_equal(LR, ref.LR)
        with pytest.raises(TEBDError):
            ctx.canonical_form(3)
    with Ensemble(2, L, 3, chi, mode="batched") as ctx:
        with pytest.raises(TEBDError):
            ctx.set_conserved(c, tot)


@pytest.mark.parametrize("mode", ["fp64", "batched"])
def test_device_canonical_form_all_nc_matches_oracle(mode):
    """CanonicalFormAllNC on a truncated, non-canonical NC state (three imaginary-time NC
    steps of Bose-Hubbard L=8 at chi=12): lambdas, state vector and labels as the
    oracle's."""
    from test_oracle_nc import itp_nc_start, propagators
    z = load_golden("itp_nc_bh_l8")
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    c, tot = conserv_of(z), int(z["tot_num"])
    gates = propagators(z["H"], -0.3j)
    ref = itp_nc_start(z, chi)
    print(f"[gpu] CanonicalFormAllNC mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode) as ctx:
        ctx.upload(*M.product_state(z["coeffs"], chi))
        ctx.set_conserved(c, tot, labels=M.product_labels(z["coeffs"], chi, c, tot))
        ctx.set_gates(gates)
        for _ in range(3):
            ctx.trotter_step()
            ref.trotter_step_nc(gates)
        ctx.canonical_form_all()
        ref.canonical_form_all_nc()
        g, l = ctx.download(chi)
        LL, LR = ctx.labels(chi)
    err_l = float(np.max(np.abs(l - ref.lam)))
    err_v = float(np.max(np.abs(mps_to_vector(g, l) - mps_to_vector(ref.gam, ref.lam))))
    print(f"[gpu] lambda err {err_l:.2e}, state err {err_v:.2e}", flush=True)
    assert err_l < 1e-12 and err_v < 1e-12
    assert np.array_equal(LL, ref.LL) and np.array_equal(LR, ref.LR)


@pytest.mark.parametrize("mode", ["fp64", "batched"])
@pytest.mark.parametrize("name", ["itp_nc_bh_l8", "itp_nc_fh_l6"])
def test_device_imag_time_prop_nc_reproduces_reference(name, mode):
    from tebd_amd import imag_time_prop
    z = load_golden(name)
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    c, tot = conserv_of(z), int(z["tot_num"])
    print(f"[gpu] ImagTimePropNC {name} mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode) as ctx:
        ctx.upload(*M.product_state(z["coeffs"], chi))
        ctx.set_conserved(c, tot, labels=M.product_labels(z["coeffs"], chi, c, tot))
        stages = imag_time_prop(ctx, z["H"], float(z["dtau"]), chi_in=int(z["chi_in"]),
                                max_steps=int(z["max_steps"]), steps_for_judge=int(z["judge"]),
                                conv1=float(z["conv1"]), conv2=float(z["conv2"]))
        g, l = ctx.download(chi)
        LL, LR = ctx.labels(chi)
    print(f"[gpu] stages {stages} (reference {list(z['stage_steps'])})", flush=True)
    assert [s for s, _ in stages] == list(z["stage_steps"])
    assert [b for _, b in stages] == list(z["stage_converged"])
    st = OracleState(g, l)
    err = {"E": abs(st.total_energy(z["H"]) - z["energy"][-1]),
           "n": float(np.max(np.abs(st.one_site_expval(z["nop"]) - z["n"][-1]))),
           "sz": float(np.max(np.abs(st.one_site_expval(z["szop"]) - z["sz"][-1]))),
           "S": float(np.max(np.abs(st.chain_entropy() - z["S"][-1]))),
           "sectors": sector_spectra_err(LL, l, z["labl"], z["lam"])}
    print(f"[gpu] errors {err}", flush=True)
    assert max(err.values()) < 1e-9, err
    assert np.array_equal(np.sort(LL, axis=1), np.sort(z["labl"], axis=1))
    assert np.array_equal(np.sort(LR, axis=1), np.sort(z["labr"], axis=1))


@pytest.mark.parametrize("mode", ["fp64", "batched"])
@pytest.mark.parametrize("name", ["nc_sched_o5_obc_l8", "nc_sched_o2_pbc_l8", "nc_sched_o2_pbc_l7",
                                  "nc_sched_o5_pbc_l8", "nc_sched_o5_pbc_l7"])
def test_device_nc_schedules_reproduce_reference(name, mode):
    from test_oracle_nc import nc_sched_errors
    z = load_golden(name)
    L, d, chi, order, bc = int(z["L"]), int(z["d"]), int(z["chi"]), int(z["order"]), str(z["bc"])
    c, tot = conserv_of(z), int(z["tot_num"])
    print(f"[gpu] NC schedule {name} mode={mode}", flush=True)
    with Context(L, d, chi, mode=mode, order=order, bc=bc) as ctx:
        ctx.upload(*M.product_state(z["coeffs"], chi))
        ctx.set_gates(z["U"])
        ctx.set_conserved(c, tot, labels=M.product_labels(z["coeffs"], chi, c, tot))

        def steps():
            for _ in range(int(z["nsteps"])):
                tr = ctx.trotter_step()
                g, lam = ctx.download(chi)
                yield tr, g, lam

        worst = nc_sched_errors(z, steps)
        LL, LR = ctx.labels(chi)
    print(f"[gpu] {name} {mode}: {worst}", flush=True)
    assert max(worst.values()) < 1e-10, worst
    assert np.array_equal(LL, z["labl"]) and np.array_equal(LR, z["labr"])
