"""Device observables (tebd_expval_two_site) against the reference's own outputs.

TotalEnergy, OneSiteExpVal (<Sz_i> or <n_i>) and ChainEntropy evaluated on the
device-resident state along a trajectory must match the reference's values from
the golden fixtures (observables_module.f90:134-252,1432-1466,1583-1602) to 1e-9,
and the oracle's evaluation of the downloaded state to 1e-12.
"""
import numpy as np
import pytest

from conftest import load_golden
from helpers import golden_obs, observables
from tebd_amd import Context, models as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["xxz_l32_chi32", "fh_l10_chi32", "bh_l12_chi24"])
def test_device_observables_match_reference(name):
    z = load_golden(name)
    print(f"[gpu] device observables {name}", flush=True)
    L, d, chi = int(z["L"]), int(z["d"]), int(z["chi"])
    ctx = Context(L, d, chi, mode="batched")
    ctx.upload(*M.product_state(z["coeffs"], chi))
    ctx.set_gates(z["U"])
    n = int(z["nsteps"])
    for k in range(1, n + 1):
        ctx.trotter_step()
        if k % 20 == 0 or k == n:
            dev = dict(E=ctx.energy(z["H"]), sz=ctx.one_site_expval(z["szop"]), n=ctx.one_site_expval(z["nop"]),
                       S=ctx.chain_entropy())
            ref = golden_obs(z, k)
            for key in dev:
                err = float(np.max(np.abs(np.asarray(dev[key]) - np.asarray(ref[key]))))
                assert err < 1e-9, (k, key, err)
            g, l = ctx.download(chi)
            orc = observables(g, l, z["H"], z["szop"], z["nop"])
            for key in dev:
                err = float(np.max(np.abs(np.asarray(dev[key]) - np.asarray(orc[key]))))
                assert err < 1e-12, (k, key, err)
    ctx.close()


def test_expval_rejects_wrong_shape():
    ctx = Context(4, 2, 4, mode="fp64")
    ctx.upload(*M.product_state(M.neel(4, 2, 0, 1), 4))
    with pytest.raises(ValueError):
        ctx.expval_two_site(np.zeros((2, 4, 4)))
    ctx.close()
