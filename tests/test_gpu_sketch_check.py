"""The learned-rank path's acceptance of a sketch, on a theta whose weight outside any
sketch is known.

A random Vidal state with chi_l = chi_r = 64 and chi_c = 16 (d = 2) gives a 128 x 128
theta of rank 16.  The operator I + eps X (X a fixed random d^2 x d^2 matrix) adds
rank up to 48 more, every new direction of size ~eps.  The fixed sketch width is
round_up(16 + 4 + 8, 32) = 32 columns (16 pre-gate Schmidt vectors, 16 Gaussian), so at
least 16 directions of weight ~eps^2 ||theta||^2 lie outside any sketch.  The probe check
(10 x the 16-probe estimate of sum_j s_j^2 ||E v_j||^2 <= 1e-20 ||theta||^4) must
  * accept at eps = 1e-13 (outside weight ~1e-26): the result agrees with the exact path
    within 1e-12 in every kept value and in the two sites' state;
  * reject at eps = 1e-7 (outside weight ~1e-14, far below the discarded-weight limit
    1e-10, so only the probe check can catch it): the bond is redone by the full SVD and
    equals the exact path's result.
"""
import numpy as np
import pytest

from tebd_amd import Context

pytestmark = pytest.mark.gpu

L, D, CHI = 4, 2, 64
DIMS = [1, 64, 16, 64, 1]


def random_state(seed):
    rng = np.random.default_rng(seed)
    gam = np.zeros((L, CHI, D, CHI), np.complex128)
    lam = np.zeros((L + 1, CHI))
    for s in range(L):
        cl, cr = DIMS[s], DIMS[s + 1]
        gam[s, :cr, :, :cl] = (rng.normal(size=(cr, D, cl)) + 1j * rng.normal(size=(cr, D, cl))) / np.sqrt(cl * D)
    for b in range(L + 1):
        v = np.sort(rng.uniform(0.05, 1.0, DIMS[b]))[::-1]
        lam[b, :DIMS[b]] = v / np.linalg.norm(v)
    return gam, lam


def two_site_state(gam, lam):
    """theta of sites 1, 2 with its outer lambdas: (a, i, j, b) -> lambda_l Gamma lambda_c Gamma lambda_r."""
    cl, cc, cr = DIMS[1], DIMS[2], DIMS[3]
    g1 = gam[1, :cc, :, :cl].transpose(2, 1, 0)  # (a, i, c)
    g2 = gam[2, :cr, :, :cc].transpose(2, 1, 0)  # (c, j, b)
    return np.einsum("a,aic,c,cjb,b->aijb", lam[1, :cl], g1, lam[2, :cc], g2, lam[3, :cr])


def run(mode, op, gam, lam):
    with Context(L, D, CHI, mode=mode) as ctx:
        ctx.upload(gam, lam)
        ctx.rank_log(True)
        ctx.two_site_op(2, op)
        st = ctx.stats()
        g, l = ctx.download(CHI)
        dims = ctx.bond_dims()
        chk = ctx.sketch_log_data() if "learned" in mode else None
    return g, l, dims, st, chk


def gated_singular_values(gam, lam, op):
    """Singular values of the gated theta (rows (i, a), columns (j, b)), by numpy."""
    th = two_site_state(gam, lam)
    thp = np.einsum("ijkl,akld->aijd", op.reshape(D, D, D, D), th)
    M = thp.transpose(1, 0, 2, 3).reshape(D * DIMS[1], D * DIMS[3])
    return np.linalg.svd(M, compute_uv=False)


@pytest.mark.parametrize("eps,accept", [(1e-13, True), (1e-7, False)])
def test_probe_check_accepts_and_rejects_by_outside_weight(eps, accept):
    print(f"[gpu] sketch probe check eps={eps:g}", flush=True)
    gam, lam = random_state(5)
    rng = np.random.default_rng(9)
    X = rng.normal(size=(D * D, D * D)) + 1j * rng.normal(size=(D * D, D * D))
    op = np.eye(D * D) + eps * X
    g0, l0, dims0, _, _ = run("batched", op, gam, lam)
    g1, l1, dims1, st, chk = run("batched,learned", op, gam, lam)
    print(f"[gpu] exact kept {dims0[2]}, learned kept {dims1[2]}; stats accepted {st['sketch_accepted']:.0f}, "
          f"redone {st['fallbacks']:.0f}, for discarded weight {st['rejected_discarded']:.0f}; check {chk}", flush=True)
    assert st["rejected_discarded"] == 0
    # the probe estimate of the weight outside the 32-column sketch against what any
    # sketch of that width must miss (the 32 smallest of the 64 nonzero sigma^2) and what
    # it can miss at most (all but the 16 largest); ||theta||^2 as numpy finds it
    sv = gated_singular_values(gam, lam, op)
    assert chk["k"][0] == 32 and chk["kmax"][0] == 64
    assert abs(chk["total"][0] / np.sum(sv ** 2) - 1.0) < 1e-12
    lo, hi = np.sum(sv[32:64] ** 2), np.sum(sv[16:] ** 2)
    print(f"[gpu] missed estimate {chk['missed'][0]:.3e}, bounds [{lo:.3e}, {hi:.3e}]", flush=True)
    assert lo / 10 <= chk["missed"][0] <= 10 * hi
    # lambda_l Gamma lambda_c Gamma lambda_r of the two sites is the normalised,
    # truncated theta itself: free of the new bond's gauge
    psi0, psi1 = two_site_state(g0, l0), two_site_state(g1, l1)
    if accept:
        assert st["sketch_accepted"] == 1 and st["fallbacks"] == 0
        k = min(dims0[2], dims1[2])
        assert np.max(np.abs(l1[2, :k] - l0[2, :k])) < 1e-12
        # the sketch keeps at most its own width; the values it leaves out are below 1e-12
        assert np.all(l0[2, k:] < 1e-12)
        assert np.max(np.abs(psi1 - psi0)) < 1e-11
    else:
        assert st["sketch_accepted"] == 0 and st["fallbacks"] == 1
        assert dims1[2] == dims0[2]
        assert np.max(np.abs(l1 - l0)) < 1e-13
        assert np.max(np.abs(psi1 - psi0)) < 1e-12
