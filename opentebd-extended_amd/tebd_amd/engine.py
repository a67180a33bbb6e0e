"""Device-resident Vidal state and the two-site update, through libtebdhip.so.

Mirrors the reference's call surface (VidalCore_v_2.0/propagation_module.f90,
local_operations_module.f90) on a context object:

=========================  ==================================================
reference                  here
=========================  ==================================================
AllocateGamLam + SVDInit   ``Context(L, d, chi_max)``  (TEBDtools_module.f90:486,
                           local_operations_module.f90:84)
ConstructPropagators       ``models.construct_propagators`` + ``set_gates``
TrotterStep                ``trotter_step()`` (propagation_module.f90:42-114)
TwoSiteOp                  ``two_site_op(link, op)`` (local_operations_module.f90:463-494)
CanonicalFormAll           ``canonical_form_all()`` (propagation_module.f90:617-634)
SVDFinish + Deallocate     ``close()``
=========================  ==================================================

States cross the boundary in the reference's padded layout: ``gam`` complex128
(L, chi, d, chi) indexed [site, b, i, a] (each site's memory is the Fortran
Gamma(a, i, b)), ``lam`` float64 (L+1, chi).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import TEBDError


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_MODES = {"fp64": _lib.MODE_FP64, "mixed": _lib.MODE_MIXED, "batched": _lib.MODE_BATCHED,
          "learned": _lib.MODE_LEARNED_RANK, "split": _lib.MODE_SPLIT}


def parse_mode(mode):
    if isinstance(mode, int):
        return mode
    m = 0
    for part in str(mode).replace("+", ",").split(","):
        part = part.strip()
        if part:
            m |= _MODES[part]
    return m


# profiled launch classes of tebd_get_profile, in order, with the kernels each covers
PROFILE_KINDS = {
    "theta_gate": "k_theta_gate / k_theta_gate_mixed / k_split_* + k_theta_split",
    "jacobi_round": "k_jacobi_round",
    "truncate_gamma": "k_truncate, k_write_gammas, k_qrp_vectors_init, k_qrp_reflect<1,2>, k_build_sketch",
    "qr_panel": "k_qrp_panel_reg",
    "sketch_gemm": "k_zgemm",
    "jacobi_rotate": "k_jacobi_rotate",
    "qr_trailing": "k_qrp_trailing_reg (k_qrp_reflect<0> above 1024 rows)",
}


class Context:
    """One chain on one GPU (one HIP stream)."""

    def __init__(self, L, d, chi_max, mode="fp64", order=2, bc="O", cutoff=None):
        self._lib = _lib.lib()
        self.L, self.d, self.chi_max = int(L), int(d), int(chi_max)
        self.mode = parse_mode(mode)
        h = ctypes.c_void_p()
        rc = self._lib.tebd_create(ctypes.byref(h), self.L, self.d, self.chi_max, int(order),
                                   bc.encode()[:1], self.mode)
        if rc != 0:
            raise TEBDError(rc, "tebd_create failed")
        self._h = h
        if cutoff is not None:
            self.set_cutoff(cutoff)

    def _check(self, rc):
        if rc != 0:
            raise TEBDError(rc, self._lib.tebd_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tebd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- set-up -------------------------------------------------------------
    def set_cutoff(self, cutoff):
        self._check(self._lib.tebd_set_cutoff(self._h, float(cutoff)))

    def set_gates(self, U):
        """U: the propagator table of the context's schedule, (n, d^2, d^2) complex indexed
        [entry, row, col]; n = L-1 for 2nd order with open boundaries (one gate per bond)."""
        u = np.ascontiguousarray(np.transpose(np.asarray(U, np.complex128), (0, 2, 1)))
        self._check(self._lib.tebd_set_gates(self._h, _p(u), u.shape[0]))

    def upload(self, gam, lam):
        gam = np.ascontiguousarray(gam, np.complex128)
        lam = np.ascontiguousarray(lam, np.float64)
        chi = gam.shape[1]
        if gam.shape != (self.L, chi, self.d, chi) or lam.shape != (self.L + 1, chi):
            raise ValueError("state shape does not match the context")
        self._check(self._lib.tebd_upload_packed(self._h, _p(gam), _p(lam), chi))

    def download(self, chi_out=None):
        chi = int(chi_out or self.chi_max)
        gam = np.zeros((self.L, chi, self.d, chi), np.complex128)
        lam = np.zeros((self.L + 1, chi))
        self._check(self._lib.tebd_download_packed(self._h, _p(gam), _p(lam), chi))
        return gam, lam

    def checkpoint(self, lambda_path, gamma_path, kind="B"):
        """The state to checkpoint files in the reference's formats (RecordLambdas /
        RecordGammas, io_module.f90:191-260) at chi_max; kind 'B' restarts bit for bit."""
        self._check(self._lib.tebd_checkpoint(self._h, str(lambda_path).encode(), str(gamma_path).encode(),
                                              kind.encode()[:1]))

    def restart(self, lambda_path, gamma_path, chi, kind="B"):
        """The state from checkpoint files recorded at chi (readGammaLambda, io_module.f90:284-373)."""
        self._check(self._lib.tebd_restart(self._h, str(lambda_path).encode(), str(gamma_path).encode(),
                                           kind.encode()[:1], int(chi)))

    def lambdas(self, chi_out=None):
        chi = int(chi_out or self.chi_max)
        lam = np.zeros((self.L + 1, chi))
        self._check(self._lib.tebd_download_lambdas(self._h, _p(lam), chi))
        return lam

    def chain_entropy(self):
        """ChainEntropy for links 2..L in log base d (observables_module.f90:1583-1602)."""
        lam = self.lambdas()
        out = np.zeros(self.L - 1)
        for link in range(2, self.L + 1):
            v = lam[link - 1]
            v = v[v != 0.0]
            out[link - 2] = -np.sum(v * v * np.log(np.abs(v * v))) / np.log(self.d)
        return out

    # -- device observables ---------------------------------------------------
    def expval_two_site(self, ops):
        """<psi|O_b|psi> for every bond b = 1..L-1 (ops: (L-1, d^2, d^2) complex, row (a-1)d+b),
        formed on the device from the resident canonical state (TwoSiteRho,
        observables_module.f90:704-832)."""
        ops = np.asarray(ops, np.complex128)
        if ops.shape != (self.L - 1, self.d ** 2, self.d ** 2):
            raise ValueError("need one d^2 x d^2 operator per bond")
        buf = np.ascontiguousarray(np.transpose(ops, (0, 2, 1)))
        out = np.zeros(2 * (self.L - 1))
        self._check(self._lib.tebd_expval_two_site(self._h, _p(buf), _p(out)))
        return out[0::2] + 1j * out[1::2]

    def two_point_rdms(self, phase=None):
        """Two-site reduced density matrices of every pair l1 < l2 within each chain and the
        one-site ones (TwoSiteRho / FormSingleSiteRho, observables_module.f90:95-117,704-832),
        from the device.  phase: diagonal of the ket operator carried over sites l1..l2-1 (the
        fermi phase), or None.  Returns (rho2, rho1): rho2[(chain,) l1, l2] is a
        (d^2, d^2) matrix indexed [i d + j, l d + k] (ket i / bra l on l1, ket j / bra k on l2),
        zero for l1 >= l2; rho1[site] is (d, d)."""
        d = self.d
        nchain = getattr(self, "nchain", 1)
        Lc = self.L // nchain
        npair = Lc * (Lc - 1) // 2
        buf = np.zeros((nchain, npair, d * d, d * d), np.complex128)
        r1 = np.zeros((self.L, d, d), np.complex128)
        ph = None if phase is None else np.ascontiguousarray(np.real(np.diag(phase) if np.ndim(phase) == 2 else phase),
                                                             np.float64)
        self._check(self._lib.tebd_two_point_rdms(self._h, None if ph is None else _p(ph), _p(buf), _p(r1)))
        rho2 = np.zeros((nchain, Lc, Lc, d * d, d * d), np.complex128)
        p = 0
        for l1 in range(Lc):
            for l2 in range(l1 + 1, Lc):
                rho2[:, l1, l2] = np.transpose(buf[:, p], (0, 2, 1))  # column-major -> [row, col]
                p += 1
        rho1 = np.transpose(r1, (0, 2, 1))
        if nchain == 1:
            rho2 = rho2[0]
        return rho2, rho1

    def energy(self, H):
        """TotalEnergy: sum over bonds of Tr(H_b rho_b) (observables_module.f90:1432-1466)."""
        return float(np.sum(self.expval_two_site(H).real))

    def one_site_expval(self, op):
        """<op_i> for every site (OneSiteExpVal, observables_module.f90:134-252): op x 1 on
        bond i, and 1 x op on the last bond for the last site."""
        op = np.asarray(op, np.complex128)
        one = np.eye(self.d)
        left = np.repeat(np.kron(op, one)[None], self.L - 1, axis=0)
        vals = self.expval_two_site(left).real
        right = np.repeat(np.kron(one, op)[None], self.L - 1, axis=0)
        last = self.expval_two_site(right).real[-1]
        return np.concatenate([vals, [last]])

    # -- rank model of the learned-rank path -----------------------------------
    def set_learned_params(self, tol_missed=1e-20, tol_trunc=1e-10, rank_slack=1.0 / 16.0):
        """Acceptance tolerances of the sketched SVD (probe check, discarded weight) and the
        relative margin added to the rank model's predicted rank."""
        self._check(self._lib.tebd_set_learned_params(self._h, tol_missed, tol_trunc, rank_slack))

    def set_rank_model(self, params):
        """Size each bond's sketch from this model's predicted rank (None: the fixed width)."""
        if params is None:
            self._check(self._lib.tebd_rank_model_set(self._h, None, 0))
            return
        p = np.ascontiguousarray(params, np.float32)
        self._check(self._lib.tebd_rank_model_set(self._h, _p(p), p.size))

    def rank_log(self, on=True):
        """Record (features, log2(kept rank / chi_c before)) of every bond update of the
        learned-rank path while on."""
        self._check(self._lib.tebd_rank_log(self._h, int(bool(on))))

    def rank_log_data(self):
        n = self._lib.tebd_rank_log_get(self._h, None, None, 0)
        if n < 0:
            self._check(n)
        X = np.zeros((n, 66), np.float32)
        y = np.zeros(n, np.float32)
        if n:
            m = self._lib.tebd_rank_log_get(self._h, _p(X), _p(y), n)
            if m < 0:
                self._check(m)
        return X, y

    # -- number conservation (TwoSiteOpNC) ----------------------------------------
    def set_conserved(self, conserv=None, tot_num=0, labels=None):
        """Keep the number of every Schmidt state (the reference's number-conserving update):
        conserv[i] the number of local state i (None: i), tot_num the total; labels =
        (LabelLeft, LabelRight), (L+1) x chi ints with 10000 for unused states, e.g.
        models.product_labels.  set_conserved(None, -1) switches it off."""
        c = None if conserv is None else np.ascontiguousarray(conserv, np.int32)
        if c is not None and c.shape != (self.d,):
            raise ValueError("one number per local state")
        self._check(self._lib.tebd_set_conserved(self._h, None if c is None else _p(c), int(tot_num)))
        if labels is not None:
            self.set_labels(*labels)

    def set_labels(self, label_left, label_right):
        LL = np.ascontiguousarray(label_left, np.int32)
        LR = np.ascontiguousarray(label_right, np.int32)
        if LL.shape != LR.shape or LL.ndim != 2 or LL.shape[0] != self.L + 1:
            raise ValueError("labels: two (L+1) x chi arrays")
        self._check(self._lib.tebd_set_labels(self._h, _p(LL), _p(LR), LL.shape[1]))

    def labels(self, chi_out=None):
        chi = int(chi_out or self.chi_max)
        LL = np.zeros((self.L + 1, chi), np.int32)
        LR = np.zeros((self.L + 1, chi), np.int32)
        self._check(self._lib.tebd_get_labels(self._h, _p(LL), _p(LR), chi))
        return LL, LR

    def sketch_log_data(self):
        """Per probe-checked sketch of the record: dict of arrays k, kmax, total (||theta||^2),
        discarded (fraction), smallest (kept s^2), eta (estimate of sum_j ||E v_j||^2),
        missed (estimate of ||E||_F^2), accepted."""
        n = self._lib.tebd_sketch_log_get(self._h, None, 0)
        if n < 0:
            self._check(n)
        out = np.zeros((n, 8))
        if n:
            m = self._lib.tebd_sketch_log_get(self._h, _p(out), n)
            if m < 0:
                self._check(m)
        names = ("k", "kmax", "total", "discarded", "smallest", "eta", "missed", "accepted")
        return {nm: out[:, i] for i, nm in enumerate(names)}

    # -- the hot path ---------------------------------------------------------
    def trotter_step(self):
        t = ctypes.c_double()
        self._check(self._lib.tebd_trotter_step(self._h, ctypes.byref(t)))
        return t.value

    def two_site_op(self, link, op):
        o = np.ascontiguousarray(np.asarray(op, np.complex128).T)
        t = ctypes.c_double()
        self._check(self._lib.tebd_two_site_op(self._h, int(link), _p(o), ctypes.byref(t)))
        return t.value

    def canonical_form_all(self):
        self._check(self._lib.tebd_canonical_form_all(self._h))


    def canonical_form(self, link):
        self._check(self._lib.tebd_canonical_form(self._h, int(link)))

    # -- introspection --------------------------------------------------------
    def bond_dims(self):
        out = np.zeros(self.L + 1, np.int32)
        self._check(self._lib.tebd_get_bond_dims(self._h, _p(out)))
        return out

    def stats(self):
        out = np.zeros(46)
        self._check(self._lib.tebd_get_stats(self._h, _p(out), 46))
        return {"bond_updates": out[0], "jacobi_sweeps": out[1], "fallbacks": out[2], "device_ms": out[3],
                "sketch_cols": out[4], "sketch_accepted": out[5], "sketch_rank": out[6], "model_predictions": out[7],
                "pair_grams": out[8], "pair_rotations": out[9], "pair_visits": out[10],
                "model_full_svd": out[11], "rejected_discarded": out[12], "lane_splits": out[13],
                # pairs rotated in sweep s + 1 of a Jacobi solve, and solves that ran s + 1 sweeps
                "sweep_rotations": out[14:30].copy(), "solve_sweeps": out[30:46].copy()}

    def set_profiling(self, on=True):
        self._check(self._lib.tebd_set_profiling(self._h, int(bool(on))))

    def profile(self):
        """{kernel: (launches, milliseconds, algorithmic flops, algorithmic HBM bytes)} since
        set_profiling(True)."""
        out = np.zeros(4 * len(PROFILE_KINDS))
        self._check(self._lib.tebd_get_profile(self._h, _p(out), out.size))
        names = tuple(PROFILE_KINDS)
        return {n: tuple(out[4 * i:4 * i + 4]) for i, n in enumerate(names)}

    def sync(self):
        self._check(self._lib.tebd_sync(self._h))

    def memory_info(self):
        """{slots, slot_bytes, state_bytes, total_bytes} of the context's device memory."""
        out = np.zeros(4)
        self._check(self._lib.tebd_memory_info(self._h, _p(out), 4))
        return {"slots": int(out[0]), "slot_bytes": out[1], "state_bytes": out[2], "total_bytes": out[3]}

    def set_graphs(self, on=True):
        """Replay each Jacobi sweep as a captured hipGraph (while per-launch profiling is off)."""
        self._check(self._lib.tebd_set_graphs(self._h, int(bool(on))))

    def set_lanes(self, n=3):
        """Split each batched half-sweep over n (1 to 4) streams driven by n host threads."""
        self._check(self._lib.tebd_set_lanes(self._h, int(n)))

    def set_autotune(self, on=True):
        """Autotuning of launch shapes: True on, False off, -1 on and forget the choices."""
        self._check(self._lib.tebd_set_autotune(self._h, int(on)))

    def autotune_info(self):
        out = np.zeros(8)
        self._check(self._lib.tebd_autotune_info(self._h, _p(out), 8))
        return {"shape_classes": int(out[0]), "timing_launches": int(out[1]),
                "chosen": {"valu_64x64": int(out[2]), "mfma_64x64": int(out[3]), "mfma_128x64": int(out[4]),
                           "mfma_64x96": int(out[5]), "mfma_64x64_k32": int(out[6]),
                           "mfma_128x96_w8": int(out[7])}}


class Ensemble(Context):
    """`nchain` independent chains of `L` sites (e.g. disorder realisations) in one
    context: every bond of one parity of every chain is updated by the same kernel
    launches (tebd_create_ensemble).  The Context methods act on the joined chain of
    nchain * L sites; the *_chains methods take and return per-chain arrays."""

    def __init__(self, nchain, L, d, chi_max, mode="batched", order=2, bc="O", cutoff=None):
        self._lib = _lib.lib()
        self.nchain, self.chain_len = int(nchain), int(L)
        self.L, self.d, self.chi_max = self.nchain * self.chain_len, int(d), int(chi_max)
        self.mode = parse_mode(mode)
        h = ctypes.c_void_p()
        rc = self._lib.tebd_create_ensemble(ctypes.byref(h), self.nchain, self.chain_len, self.d, self.chi_max,
                                            int(order), bc.encode()[:1], self.mode)
        if rc != 0:
            raise TEBDError(rc, "tebd_create_ensemble failed")
        self._h = h
        if cutoff is not None:
            self.set_cutoff(cutoff)

    def set_gates(self, U):
        """U: (nchain, n, d^2, d^2) complex, every chain's propagator table (n = L-1 for 2nd
        order with open boundaries)."""
        U = np.asarray(U, np.complex128)
        n = self.nchain * U.shape[1]
        u = np.ascontiguousarray(np.transpose(U.reshape(n, self.d ** 2, self.d ** 2), (0, 2, 1)))
        self._check(self._lib.tebd_set_gates(self._h, _p(u), n))

    def upload_chains(self, gams, lams):
        """gams: (nchain, L, chi, d, chi), lams: (nchain, L+1, chi); the outer lambdas of
        every chain are (1, 0, ...)."""
        gams = np.asarray(gams, np.complex128)
        lams = np.asarray(lams, np.float64)
        n, Lc = self.nchain, self.chain_len
        if gams.shape[:2] != (n, Lc) or lams.shape[:2] != (n, Lc + 1):
            raise ValueError("need nchain states of L sites")
        lam = np.concatenate([lams[:, :Lc].reshape(n * Lc, -1), lams[-1:, Lc]], axis=0)
        self.upload(gams.reshape((n * Lc,) + gams.shape[2:]), lam)

    def _split_lam(self, lam):
        Lc = self.chain_len
        return np.stack([lam[r * Lc:r * Lc + Lc + 1] for r in range(self.nchain)])

    def download_chains(self, chi_out=None):
        gam, lam = self.download(chi_out)
        return gam.reshape((self.nchain, self.chain_len) + gam.shape[1:]), self._split_lam(lam)

    def lambdas_chains(self, chi_out=None):
        return self._split_lam(self.lambdas(chi_out))

    def expval_two_site_chains(self, ops):
        """ops: (nchain, L-1, d^2, d^2) -> (nchain, L-1) complex expectation values."""
        ops = np.asarray(ops, np.complex128)
        Lc, d2 = self.chain_len, self.d ** 2
        full = np.zeros((self.L - 1, d2, d2), np.complex128)
        for r in range(self.nchain):
            full[r * Lc:r * Lc + Lc - 1] = ops[r]
        vals = self.expval_two_site(full)
        return np.stack([vals[r * Lc:r * Lc + Lc - 1] for r in range(self.nchain)])

    def energy_chains(self, H):
        """TotalEnergy of every chain; H: (nchain, L-1, d^2, d^2)."""
        return self.expval_two_site_chains(H).real.sum(axis=1)

    def one_site_expval_chains(self, op):
        op = np.asarray(op, np.complex128)
        one = np.eye(self.d)
        n, Lc = self.nchain, self.chain_len
        left = np.broadcast_to(np.kron(op, one), (n, Lc - 1, self.d ** 2, self.d ** 2))
        right = np.broadcast_to(np.kron(one, op), (n, Lc - 1, self.d ** 2, self.d ** 2))
        vl = self.expval_two_site_chains(left).real
        vr = self.expval_two_site_chains(right).real
        return np.concatenate([vl, vr[:, -1:]], axis=1)

    def chain_entropy_chains(self):
        lam = self.lambdas_chains()
        out = np.zeros((self.nchain, self.chain_len - 1))
        for r in range(self.nchain):
            for link in range(2, self.chain_len + 1):
                v = lam[r, link - 1]
                v = v[v != 0.0]
                out[r, link - 2] = -np.sum(v * v * np.log(np.abs(v * v))) / np.log(self.d)
        return out


def read_state_files(L, d, chi, lambda_path, gamma_path, kind="S"):
    """A Vidal state (gam [site, b, i, a], lam) from checkpoint files in the reference's
    formats, recorded at chi (readGammaLambda, io_module.f90:284-373).  Host-only."""
    gam = np.zeros((L, chi, d, chi), np.complex128)
    lam = np.zeros((L + 1, chi))
    rc = _lib.lib().tebd_read_state_files(L, d, chi, _p(gam), _p(lam), str(lambda_path).encode(),
                                          str(gamma_path).encode(), kind.encode()[:1])
    if rc != 0:
        raise TEBDError(rc, "cannot read the state files")
    return gam, lam


def write_state_files(gam, lam, lambda_path, gamma_path, kind="S"):
    """Checkpoint files of a Vidal state in the reference's formats (RecordLambdas /
    RecordGammas, io_module.f90:191-260).  Host-only."""
    gam = np.ascontiguousarray(gam, np.complex128)
    lam = np.ascontiguousarray(lam, np.float64)
    L, chi, d, _ = gam.shape
    rc = _lib.lib().tebd_write_state_files(L, d, chi, _p(gam), _p(lam), str(lambda_path).encode(),
                                           str(gamma_path).encode(), kind.encode()[:1])
    if rc != 0:
        raise TEBDError(rc, "cannot write the state files")


def construct_propagators(H, dt, order=2, bc="O"):
    """ConstructPropagators (TEBDtools_module.f90:1079-1193) by the library's host routine:
    the propagator table of the Trotter schedule (order 2 or 5, open 'O' or periodic 'P'
    boundaries), entry k = exp(-i f_k dt H(bond_k)); dt complex (imaginary time:
    dt = -i dtau).  H: (L-1, d^2, d^2), or (L, d^2, d^2) with periodic boundaries (the last
    one the edge) -> U: (table size, d^2, d^2)."""
    H = np.asarray(H, np.complex128)
    nb, n = H.shape[0], H.shape[1]
    d = int(round(np.sqrt(n)))
    L = nb if bc == "P" else nb + 1
    size = {2: (L if bc == "P" else L - 1), 5: (2 * (L - 1) if bc != "P" else (2 * L if L % 2 == 0 else 3 * L))}
    h = np.ascontiguousarray(np.transpose(H, (0, 2, 1)))
    u = np.zeros((size[int(order)], n, n), np.complex128)
    dt = complex(dt)
    rc = _lib.lib().tebd_construct_propagators(_p(h), L, d, int(order), bc.encode()[:1], dt.real, dt.imag, _p(u))
    if rc != u.shape[0]:
        raise TEBDError(rc, "tebd_construct_propagators failed")
    return np.transpose(u, (0, 2, 1))


def imag_time_prop(ctx, H, dtau, chi_in=None, max_steps=2000, steps_for_judge=100, conv1=1e-5, conv2=1e-6):
    """ImagTimeProp (propagation_module.f90:636-760) on the device: gates exp(-dtau H),
    chi stages chi_in -> ctx.chi_max (each starting with every bond cut to that chi), and
    in each stage TrotterStep + CanonicalFormAll until, at a multiple of steps_for_judge,
    the largest relative change of any bond's largest lambda is below conv1 (stages below
    chi_max) or conv2 (the chi_max stage).  The defaults are the reference's
    system_parameters (maxITPsteps, stepsForJudge, convCriterion1/2).  H: (L-1, d^2, d^2),
    or (nchain, L-1, d^2, d^2) for an Ensemble.  Returns [(steps, converged), ...] per stage."""
    H = np.asarray(H, np.complex128)
    n = H.shape[-1]
    h = np.ascontiguousarray(np.transpose(H.reshape(-1, n, n), (0, 2, 1)))
    steps = np.zeros(4, np.int32)
    chi_in = ctx.chi_max if chi_in is None else int(chi_in)
    ctx._check(ctx._lib.tebd_imag_time_prop(ctx._h, _p(h), h.shape[0], float(dtau), chi_in, int(max_steps),
                                            int(steps_for_judge), float(conv1), float(conv2), _p(steps)))
    nst = 1 if chi_in == ctx.chi_max else 2
    return [(int(steps[2 * k]), bool(steps[2 * k + 1])) for k in range(nst)]
