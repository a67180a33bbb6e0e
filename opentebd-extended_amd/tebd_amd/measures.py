"""EvaluateMeasures (observables_module.f90:2286-2526) on the device's reduced density
matrices.

The reference fills a `measure` record (TEBDtools_module.f90:59-88, AllocateMeasures
observables_module.f90:2165-2231) with local values <O_i>, site averages, correlation
matrices <O2_{l1,l2}> with and without a fermi phase, the single-site von Neumann
entropies (log_d), the Meyer-Wallach Q measure, the chain (bond) entropies, two-site
entropies, and the energy (with the edge bond of a periodic chain).  Here the pair and
site density matrices come from the device (Context.two_point_rdms, all pairs in one
sweep of ThetaKernal / ThetaNext / TwoSiteRho per right site); the traces and the small
eigenvalue problems are evaluated on the host.
"""
import numpy as np


def _trace(op, rho):
    """TraceMatmul(Op, rho) = sum_ab Op(a,b) rho(b,a)."""
    return np.einsum("ab,...ba->...", op, rho)


def _entropy(rho, d):
    w = np.linalg.eigvalsh(0.5 * (rho + rho.conj().T))
    w = w[w != 0.0]
    return float(-np.sum(w * np.log(np.abs(w))) / np.log(d))


def evaluate_measures(ctx, H=None, local=(), avg=(), corr=(), fermicorr=(), ent=0, phase=None, bc="O"):
    """The reference's EvaluateMeasures on the context's state.

    local, avg: d x d operators; corr, fermicorr: d^2 x d^2 operators (left site slow, as
    the gates); phase: the fermi phase operator (or its diagonal) carried by fermicorr;
    ent: 0 none, 1 single-site and chain entropies and the Q measure, 2 also two-site
    entropies.  H: bond Hamiltonians (L-1, or L with bc='P') for the energy.  Returns a
    dict with the record's fields: 'local' (n, L), 'avg' (n,), 'corr' / 'fermicorr'
    (n, L, L; off-diagonal, the diagonal left at zero as the reference leaves it unset),
    'en', and with ent >= 1 'vN' (L), 'qme', 'chain' (L+1), with ent == 2 'tbvN' (L, L)."""
    d, L = ctx.d, ctx.L
    rho2, rho1 = ctx.two_point_rdms()
    out = {}
    out["local"] = np.array([[_trace(np.asarray(o), rho1[s]) for s in range(L)] for o in local])
    out["avg"] = np.array([np.mean([_trace(np.asarray(o), rho1[s]) for s in range(L)]) for o in avg])

    def corr_of(ops, r2):
        m = np.zeros((len(ops), L, L), np.complex128)
        for n, o in enumerate(ops):
            for l1 in range(L):
                for l2 in range(l1 + 1, L):
                    m[n, l1, l2] = _trace(np.asarray(o), r2[l1, l2])
                    m[n, l2, l1] = np.conj(m[n, l1, l2])
        return m

    out["corr"] = corr_of(corr, rho2)
    if len(fermicorr):
        if phase is None:
            raise ValueError("fermicorr needs the fermi phase operator")
        rho2f, _ = ctx.two_point_rdms(phase=phase)
        out["fermicorr"] = corr_of(fermicorr, rho2f)
    if H is not None:
        H = np.asarray(H)
        en = sum(_trace(H[b], rho2[b, b + 1]).real for b in range(L - 1))
        if bc == "P":
            en += _trace(H[L - 1], rho2[0, L - 1]).real
        out["en"] = float(en)
    if ent >= 1:
        out["vN"] = np.array([_entropy(rho1[s], d) for s in range(L)])
        out["qme"] = float(d / (d - 1.0) - sum(_trace(rho1[s], rho1[s]).real for s in range(L)) * d / (L * (d - 1.0)))
        lam = ctx.lambdas()
        ch = []
        for b in range(L + 1):
            v = lam[b]
            v = v[v != 0.0]
            ch.append(float(-np.sum(v * v * np.log(np.abs(v * v))) / np.log(d)))
        out["chain"] = np.array(ch)
    if ent == 2:
        tb = np.zeros((L, L))
        for s in range(L):
            tb[s, s] = 2.0 * out["vN"][s]
        for l1 in range(L):
            for l2 in range(l1 + 1, L):
                tb[l1, l2] = tb[l2, l1] = _entropy(rho2[l1, l2], d)
        out["tbvN"] = tb
    return out
