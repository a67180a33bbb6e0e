"""ctypes front-end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  tests/, the package's smoke() and bench.py's
cpu_baseline leg use it as the checker or as the timed CPU baseline; the
shipped library never imports it.

State convention (the reference's, TEBDtools_module.f90:486-521): uniform chi,
``gam`` is a complex128 array of shape (L, chi, d, chi) indexed [site, b, i, a]
so that each site's memory is the Fortran column-major Gamma(a, i, b);
``lam`` is float64 of shape (L+1, chi); gates are complex128 (nb, d^2, d^2)
matrices indexed [bond, row, col] with row (a-1)d+b (left site slow).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        # liboracle links MKL's single dynamic library with the GNU OpenMP layer; every
        # LAPACK call runs on one thread (as the reference links mkl_sequential) unless
        # checker_speed() asks for more.
        os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        P = ctypes.c_void_p
        I = ctypes.c_int
        L.orc_two_site_op.argtypes = [I, I, I, P, P, I, P, P]
        L.orc_trotter_step.argtypes = [I, I, I, P, P, P, P, I]
        L.orc_canonical_form_all.argtypes = [I, I, I, P, P]
        L.orc_canonical_form.argtypes = [I, I, I, P, P, I]
        L.orc_total_energy.argtypes = [I, I, I, P, P, P]
        L.orc_total_energy.restype = ctypes.c_double
        L.orc_two_site_rho_all.argtypes = [I, I, I, P, P, P]
        L.orc_one_site_expval.argtypes = [I, I, I, P, P, P, P]
        L.orc_chain_entropy.argtypes = [I, I, I, P, P]
        L.orc_set_checker_speed.argtypes = [I]
        L.orc_trotter_step_sched.argtypes = [I, I, I, P, P, P, I, I, ctypes.c_char, P]
        L.orc_two_site_op_nc.argtypes = [I, I, I, P, P, P, P, I, P, P, I, I, P]
        L.orc_trotter_step_nc.argtypes = [I, I, I, P, P, P, P, P, P, I, I, P]
        L.orc_canonical_form_all_nc.argtypes = [I, I, I, P, P, P, P, P, I, I]
        L.orc_trotter_step_nc_sched.argtypes = [I, I, I, P, P, P, P, P, I, I, ctypes.c_char, P, I, I, P, I]
        _LIB = L
    return _LIB


def checker_speed(threads):
    """threads > 0: theta by ZGEMM and ZGESVD with its optimal workspace on that many
    threads (a checker for large bonds); 0: the reference-faithful mode (default)."""
    lib().orc_set_checker_speed(int(threads))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _gate_stack(gates, d):
    # (nb, d2, d2) row/col indexed -> Fortran column-major per bond
    g = np.asarray(gates, dtype=np.complex128)
    return np.ascontiguousarray(np.transpose(g, (0, 2, 1)))


class OracleState:
    """Vidal-form MPS held in the reference's padded layout."""

    def __init__(self, gam, lam):
        self.gam = np.ascontiguousarray(gam, dtype=np.complex128)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64)
        self.L, self.chi, self.d, _ = self.gam.shape

    @classmethod
    def product(cls, coeffs, chi):
        """ProductStateMPD (Hamiltonian_tools_module.f90:89-121)."""
        c = np.asarray(coeffs, dtype=np.complex128)
        L, d = c.shape
        gam = np.zeros((L, chi, d, chi), np.complex128)
        gam[:, 0, :, 0] = c
        lam = np.zeros((L + 1, chi))
        lam[:, 0] = 1.0
        return cls(gam, lam)

    def copy(self):
        return OracleState(self.gam.copy(), self.lam.copy())

    def _args(self):
        return self.L, self.d, self.chi, _p(self.gam), _p(self.lam)

    def two_site_op(self, link, gate):
        g = np.ascontiguousarray(np.asarray(gate, np.complex128).T)
        t = ctypes.c_double()
        info = lib().orc_two_site_op(*self._args(), int(link), _p(g), ctypes.byref(t))
        if info != 0:
            raise RuntimeError(f"zgesvd info={info}")
        return t.value

    def trotter_step(self, gates, nthreads=1):
        g = _gate_stack(gates, self.d)
        t = ctypes.c_double()
        info = lib().orc_trotter_step(*self._args(), _p(g), ctypes.byref(t), int(nthreads))
        if info != 0:
            raise RuntimeError(f"zgesvd info={info}")
        return t.value

    def trotter_step_sched(self, gates, order=2, bc="O"):
        """TrotterStep of any of the reference's schedules: order 2 or 5, open ('O') or
        periodic ('P') boundaries; gates = the matching propagator table."""
        g = _gate_stack(gates, self.d)
        t = ctypes.c_double()
        info = lib().orc_trotter_step_sched(*self._args(), _p(g), g.shape[0], int(order), bc.encode()[:1],
                                            ctypes.byref(t))
        if info != 0:
            raise RuntimeError(f"zgesvd info={info} (or a gate table of the wrong size)")
        return t.value

    # -- number conservation (TwoSiteOpNC) --------------------------------------
    def set_conserved(self, labels_left, labels_right, conserv, tot_num):
        """Labels (L+1, chi) of the Schmidt states (number left / right of each link,
        10000 = unused), the number of each local state and the total number."""
        self.LL = np.ascontiguousarray(labels_left, np.int32)
        self.LR = np.ascontiguousarray(labels_right, np.int32)
        self.conserv = np.ascontiguousarray(conserv, np.int32)
        self.tot = int(tot_num)
        self.maxf = int(self.conserv.max())

    def _nc(self):
        return _p(self.LL), _p(self.LR)

    def two_site_op_nc(self, link, gate):
        g = np.ascontiguousarray(np.asarray(gate, np.complex128).T)
        t = ctypes.c_double()
        info = lib().orc_two_site_op_nc(*self._args(), *self._nc(), int(link), _p(g), _p(self.conserv), self.tot,
                                        self.maxf, ctypes.byref(t))
        if info != 0:
            raise RuntimeError(f"zgesvd info={info} (or no block)")
        return t.value

    def trotter_step_nc(self, gates):
        g = _gate_stack(gates, self.d)
        t = ctypes.c_double()
        info = lib().orc_trotter_step_nc(*self._args(), *self._nc(), _p(g), _p(self.conserv), self.tot, self.maxf,
                                         ctypes.byref(t))
        if info != 0:
            raise RuntimeError(f"zgesvd info={info} (or no block)")
        return t.value

    def trotter_step_nc_sched(self, gates, order=2, bc="O", swap_renorm=False):
        """TrotterStepNC of any of the reference's schedules (order 2 or 5, 'O' or 'P').
        swap_renorm=True plants a deviation (theta normalised in the swaps, which the
        reference's SwappingNC does not do) for tests of the fixtures' resolution."""
        g = _gate_stack(gates, self.d)
        t = ctypes.c_double()
        info = lib().orc_trotter_step_nc_sched(*self._args(), *self._nc(), _p(g), g.shape[0], int(order),
                                               bc.encode()[:1], _p(self.conserv), self.tot, self.maxf,
                                               ctypes.byref(t), int(bool(swap_renorm)))
        if info != 0:
            raise RuntimeError(f"zgesvd info={info} (or no block, or a gate table of the wrong size)")
        return t.value

    def canonical_form_all_nc(self):
        """CanonicalFormAllNC (propagation_module.f90:1943-1983)."""
        info = lib().orc_canonical_form_all_nc(*self._args(), *self._nc(), _p(self.conserv), self.tot, self.maxf)
        if info != 0:
            raise RuntimeError(f"zgesvd info={info} (or no block)")

    def with_chi(self, chi):
        """The state in arrays of bond capacity chi, entries past the smaller capacity
        dropped or zero (CopyGamLam, TEBDtools_module.f90; labels as CopyLabel :623-642)."""
        m = min(chi, self.chi)
        gam = np.zeros((self.L, chi, self.d, chi), np.complex128)
        gam[:, :m, :, :m] = self.gam[:, :m, :, :m]
        lam = np.zeros((self.L + 1, chi))
        lam[:, :m] = self.lam[:, :m]
        out = OracleState(gam, lam)
        if hasattr(self, "LL"):
            LL = np.full((self.L + 1, chi), 10000, np.int32)
            LR = np.full((self.L + 1, chi), 10000, np.int32)
            LL[:, :m] = self.LL[:, :m]
            LR[:, :m] = self.LR[:, :m]
            out.set_conserved(LL, LR, self.conserv, self.tot)
        return out

    def canonical_form_all(self):
        info = lib().orc_canonical_form_all(*self._args())
        if info != 0:
            raise RuntimeError(f"zgesvd info={info}")

    def total_energy(self, H):
        h = _gate_stack(H, self.d)
        return lib().orc_total_energy(*self._args(), _p(h))

    def two_site_rho(self):
        d2 = self.d * self.d
        out = np.zeros((self.L - 1, d2, d2), np.complex128)
        lib().orc_two_site_rho_all(*self._args(), _p(out))
        # stored column-major per bond -> [bond, row, col]
        return np.transpose(out, (0, 2, 1))

    def one_site_expval(self, op):
        o = np.ascontiguousarray(np.asarray(op, np.float64).T)
        out = np.zeros(self.L)
        lib().orc_one_site_expval(*self._args(), _p(o), _p(out))
        return out

    def chain_entropy(self):
        out = np.zeros(self.L - 1)
        lib().orc_chain_entropy(self.L, self.d, self.chi, _p(self.lam), _p(out))
        return out


def imag_time_prop_nc(state, gates, chi_in, chi_max, max_steps, steps_for_judge, conv1, conv2):
    """ImagTimePropNC (propagation_module.f90:1986-2121) from `state` (labels set, any
    capacity) with the imaginary-time gates (ConstructPropagators with dt = -i dtITP):
    stages chi = chi_in, chi_in + (chi_max - chi_in) (one stage when they are equal); each
    stage copies the state into capacity chi (CopyGamLam, CopyLabel), judges every
    steps_for_judge steps whether the largest relative change of lambda_1 over all links is
    below its criterion (conv2 on the chi_max stage or when chi_in == chi_max, else conv1),
    and otherwise takes TrotterStepNC and CanonicalFormAllNC.  Returns ([(steps,
    converged)] per stage, the final state)."""
    inc = 1 if chi_in == chi_max else chi_max - chi_in
    stages = []
    for chi in range(chi_in, chi_max + 1, inc):
        conv = conv2 if (chi_in == chi_max or chi == chi_max) else conv1
        state = state.with_chi(chi)
        before = state.lam[:, 0].copy()
        done = None
        for j in range(1, max_steps + 1):
            if j % steps_for_judge == 0:
                after = state.lam[:, 0].copy()
                gap = float(np.max(np.abs((before - after) / before)))
                if gap < conv:
                    done = j
                    break
                before = after
            state.trotter_step_nc(gates)
            state.canonical_form_all_nc()
        stages.append((done if done is not None else max_steps, done is not None))
    return stages, state


def product_labels(coeffs, chi, conserv, tot_num):
    """ProductStateLabels (Hamiltonian_tools_module.f90:124-178) for a product state of
    number eigenstates: the number left of each link from the first nonzero coefficient
    of each site, the right labels tot_num - left, 10000 past the first Schmidt state."""
    c = np.asarray(coeffs)
    L = c.shape[0]
    LL = np.full((L + 1, chi), 10000, np.int32)
    LR = np.full((L + 1, chi), 10000, np.int32)
    LL[0, 0] = 0
    for s in range(L):
        j = int(np.flatnonzero(np.abs(c[s]) != 0.0)[0])
        LL[s + 1, 0] = LL[s, 0] + int(conserv[j])
    LR[:, 0] = tot_num - LL[:, 0]
    return LL, LR


def two_point_rdms(gam, lam, phase=None):
    """All-pairs two-site and one-site reduced density matrices, restating the
    reference's ThetaKernal / ThetaNext / TwoSiteRho recursion (observables_module.f90:
    704-832, as TwoSiteExpVal and EvaluateMeasures drive it, :836-883,2286-2526) with numpy
    on the padded state (gam [site, b, i, a] = Gamma(a, i, b), lam [bond, a]).  phase: the
    diagonal of the one-site operator applied to every ket Gamma of sites l1 .. l2-1, or
    None.  Returns rho2 [l1, l2, i d + j, l d + k] for l1 < l2 and rho1 [s, i, j]."""
    gam = np.asarray(gam)
    L, chi, d, _ = gam.shape
    G = np.transpose(gam, (0, 3, 2, 1))  # [site, a, i, b]
    ph = np.ones(d) if phase is None else np.asarray(phase, float)
    rho2 = np.zeros((L, L, d * d, d * d), np.complex128)
    rho1 = np.zeros((L, d, d), np.complex128)
    for l2 in range(L - 1, -1, -1):
        l1v, l2v = lam[l2], lam[l2 + 1]
        # ThetaKernal: Theta(a,i,j,b) = sum_e l1(a) G(a,i,e) l2(e)^2 conj(G(b,j,e)) l1(b)
        th = np.einsum("a,aie,e,bje,b->aijb", l1v, G[l2], l2v ** 2, G[l2].conj(), l1v, optimize=True)
        rho1[l2] = np.einsum("aija->ij", th)
        for l1 in range(l2 - 1, -1, -1):
            lv = lam[l1]
            Gp = G[l1] * ph[None, :, None]
            # ThetaTemp(b,j,k,l,a) = sum_e Theta(b,j,k,e) conj(G(a,l,e)) lam(a)
            tt = np.einsum("bjke,ale,a->bjkla", th, G[l1].conj(), lv, optimize=True)
            # rho2((i)d+j,(l)d+k) = sum_{a,b} lam(a) Gp(a,i,b) ThetaTemp(b,j,k,l,a)
            r = np.einsum("a,aib,bjkla->ijlk", lv, Gp, tt, optimize=True)
            rho2[l1, l2] = r.reshape(d * d, d * d)
            # ThetaNext: Theta(a,i,j,b) = sum_{k,e} lam(a) Gp(a,k,e) ThetaTemp(e,i,j,k,b)
            th = np.einsum("a,ake,eijkb->aijb", lv, Gp, tt, optimize=True)
    return rho2, rho1
