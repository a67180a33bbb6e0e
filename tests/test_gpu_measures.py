"""Device all-pairs density matrices and EvaluateMeasures (tebd_two_point_rdms,
tebd_amd.measures) against the reference's EvaluateMeasures (tests/golden/meas_*.npz)
and against the oracle's restatement on a larger state.  Tolerance 1e-10."""
import numpy as np
import pytest

from conftest import load_golden
from oracle.oracle import OracleState, two_point_rdms
from test_measures_golden import CASES, case_ops
from tebd_amd import Context, models as M
from tebd_amd.measures import evaluate_measures

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CASES)
def test_device_measures_match_reference(name):
    z = load_golden(name)
    L, d, chi, bc = int(z["L"]), int(z["d"]), int(z["chi"]), str(z["bc"])
    ops = case_ops(z)
    print(f"[gpu] measures {name}", flush=True)
    with Context(L, d, chi, mode="fp64", bc=bc) as ctx:
        ctx.upload(z["gam"], z["lam"])
        m = evaluate_measures(ctx, H=z["H"], local=[ops["loc"]], avg=[ops["loc"]], corr=[ops["corr"]],
                              fermicorr=[ops["fcorr"]] if ops["fcorr"] is not None else (), ent=2,
                              phase=ops["phase"], bc=bc)
    iu = np.triu_indices(L, 1)
    err = {"local": np.max(np.abs(m["local"][0] - z["m_local"])), "avg": abs(m["avg"][0] - z["m_avg"]),
           "corr": np.max(np.abs(m["corr"][0][iu] - z["m_corr"][iu])), "en": abs(m["en"] - z["m_en"]),
           "vN": np.max(np.abs(m["vN"] - z["m_vn"])), "qme": abs(m["qme"] - z["m_qme"]),
           "chain": np.max(np.abs(m["chain"] - z["m_chain"])), "tbvN": np.max(np.abs(m["tbvN"] - z["m_tbvn"]))}
    if ops["fcorr"] is not None:
        err["fcorr"] = np.max(np.abs(m["fermicorr"][0][iu] - z["m_fcorr"][iu]))
    print(f"[gpu] {name}: {err}", flush=True)
    assert max(err.values()) < 1e-10, err


def test_device_rdms_match_oracle_on_a_larger_state():
    """Fermi-Hubbard L=12, chi=64 (a state evolved by the device): every pair's density
    matrix, with and without the fermi phase, against the oracle's restatement."""
    L, d, chi = 12, 4, 64
    H = M.fermi_hubbard(L, 1.0, 4.0)
    rng = np.random.default_rng(8)
    c = rng.normal(size=(L, d)) + 1j * rng.normal(size=(L, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    F = np.diag(M.fermion_ops(2)[4]).real
    print("[gpu] rdms FH L=12 chi=64 vs oracle", flush=True)
    with Context(L, d, chi, mode="batched") as ctx:
        ctx.upload(*M.product_state(c, chi))
        ctx.set_gates(M.construct_propagators(H, 0.1))
        for _ in range(6):
            ctx.trotter_step()
        g, lam = ctx.download(chi)
        r2, r1 = ctx.two_point_rdms()
        r2f, _ = ctx.two_point_rdms(phase=F)
    o2, o1 = two_point_rdms(g, lam)
    o2f, _ = two_point_rdms(g, lam, phase=F)
    iu = np.triu_indices(L, 1)
    err = (np.max(np.abs(r2[iu] - o2[iu])), np.max(np.abs(r1 - o1)), np.max(np.abs(r2f[iu] - o2f[iu])))
    print(f"[gpu] rdm errors {err}", flush=True)
    assert max(err) < 1e-11, err
