/* C ABI of libtebdhip.so: the two-site Trotter update of OpenSourceTEBD v2.0
 * (vvolkl/OpenTEBD-extended) on one MI355X (gfx950).
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved (re, im)
 * doubles, i.e. Fortran COMPLEX(KIND=8); matrices are column-major, exactly
 * as the reference's derived types hold them:
 *   Gammas(i)%t(chi,d,chi)  complex(8)  (VidalCore_v_2.0/TEBDtools_module.f90:45,486-521)
 *   Lambdas(i)%v(chi)       real(8)     (TEBDtools_module.f90:31-33,486-521)
 *   Udt(i)%m(d*d,d*d)       complex(8), row (a-1)d+b, left site slow
 *                                       (TEBDtools_module.f90:37-39,937-998)
 * Every call returns 0 (TEBD_OK) or a negative code; tebd_last_error() gives
 * the message.  The reference has no error channel (ZGESVD info is ignored,
 * VidalCore_v_2.0/local_operations_module.f90:335-336; bad schedules STOP,
 * propagation_module.f90:64,76); a Fortran shim maps non-zero codes to STOP.
 *
 * One context = one chain = one HIP stream; calls on a context are
 * serialised, contexts are independent (no global state).  State lives on
 * the device between calls; host arrays are authoritative only after
 * tebd_download().
 */
#ifndef TEBD_HIP_H
#define TEBD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tebd_ctx tebd_ctx;

enum tebd_status {
  TEBD_OK = 0,
  TEBD_ERR_ARG = -1,         /* bad argument / unsupported order or boundary */
  TEBD_ERR_HIP = -2,         /* HIP runtime error (allocation, launch, copy) */
  TEBD_ERR_SHAPE = -3,       /* dimensions do not fit chi_max / d / L */
  TEBD_ERR_NAN = -4,         /* theta is zero, NaN or Inf */
  TEBD_ERR_NOCONV = -5,      /* SVD sweeps did not converge */
  TEBD_ERR_STATE = -6,       /* call out of order (no gates, no state) */
  TEBD_ERR_UNSUPPORTED = -7  /* feature not available in this build */
};

/* Mode bits: they change precision and launch policy of the same kernels. */
enum tebd_mode {
  TEBD_MODE_FP64 = 0u,              /* plain FP64 drop-in: one bond at a time, full SVD */
  TEBD_MODE_MIXED = 1u << 0,        /* FP32 theta/gate on MFMA, FP64 Gram/SVD/lambda */
  TEBD_MODE_BATCHED = 1u << 1,      /* all bonds of one parity per launch */
  TEBD_MODE_LEARNED_RANK = 1u << 2, /* randomized SVD sized by the rank model, full-SVD fallback */
  TEBD_MODE_SPLIT = 1u << 3         /* theta on the int8 matrix cores from exact power-of-two
                                       slices of the FP64 operands (integer accumulation, FP64
                                       combination); not with TEBD_MODE_MIXED */
};

/* Allocates the device state for a chain of L sites, local dimension d and
 * bond dimension cap chi_max, with the Trotter schedule of order 2 or 5 (Forest-Ruth) and
 * open ('O') or periodic ('P', L >= 3, not in an ensemble) boundaries
 * (trotterOrder / BoundaryCond, propagation_module.f90:42-614).  Replaces
 * AllocateGamLam (TEBDtools_module.f90:486) + SVDInit (local_operations_module.f90:84-122). */
int tebd_create(tebd_ctx** ctx, int L, int d, int chi_max, int order, char bc, unsigned mode);

/* An ensemble: nchain independent chains of L sites each (e.g. disorder
 * realisations) in one context, so that every bond of one parity of every chain
 * is updated by the same kernel launches.  No counterpart in the reference, which
 * runs one chain per program (realisations are separate runs of a main, e.g.
 * Case_Studies/Bose_Hubbard_Wrapper/BoseHubbard_Quench.f90).  The context then
 * behaves as one chain of nchain * L sites whose nchain - 1 bonds between chains
 * (bond r*L, r = 1..nchain-1, 1-based left site) are never updated: upload and
 * download take nchain * L Gammas and nchain * L + 1 lambdas, and the lambda on each
 * such bond must be (1, 0, ...); tebd_set_gates takes nchain * (L - 1) gates, chain
 * after chain; the Trotter schedule restarts its odd/even bond numbering in every
 * chain; tebd_two_site_op and tebd_canonical_form refuse the bonds between chains;
 * tebd_expval_two_site still takes and returns nchain * L - 1 entries (those of
 * the bonds between chains evaluate the operator on the product of two chains). */
int tebd_create_ensemble(tebd_ctx** ctx, int nchain, int L, int d, int chi_max, int order, char bc, unsigned mode);

/* Frees everything; replaces SVDFinish (local_operations_module.f90:124-153)
 * and DeallocateGamLam. */
void tebd_destroy(tebd_ctx* ctx);

/* The Udt table built by ConstructPropagators (TEBDtools_module.f90:1079-1193) for the
 * context's schedule: nb = its size (L-1 for 2nd order with open boundaries; see
 * tebd_construct_propagators), matrices of (d*d) x (d*d) complex, column-major, packed
 * back to back; an ensemble takes the tables of its chains one after another. */
int tebd_set_gates(tebd_ctx* ctx, const double* U, int nb);

/* Uploads a Vidal state held in the reference's padded layout: gamma[s] points
 * at site s's chi_in x d x chi_in complex array, lambda[l] (l = 0..L) at bond
 * l's chi_in reals.  Bond dimensions are taken from the non-zero lambdas. */
int tebd_upload(tebd_ctx* ctx, const double* const* gamma, const double* const* lambda, int chi_in);
/* Same, with all sites (and all lambdas) contiguous. */
int tebd_upload_packed(tebd_ctx* ctx, const double* gamma, const double* lambda, int chi_in);

/* Downloads the state into the padded layout (zero-filled past each bond's
 * current dimension); chi_out must be >= every bond dimension. */
int tebd_download(tebd_ctx* ctx, double* const* gamma, double* const* lambda, int chi_out);
int tebd_download_packed(tebd_ctx* ctx, double* gamma, double* lambda, int chi_out);
/* Only the lambdas, (L+1) x chi_out reals, zero-filled: what ChainEntropy
 * (observables_module.f90:1583-1602) and the imaginary-time convergence test
 * (propagation_module.f90:704-738) read. */
int tebd_download_lambdas(tebd_ctx* ctx, double* lambda, int chi_out);

/* Checkpoint files of a Vidal state in the reference's text formats (RecordLambdas and
 * RecordGammas, ToolsCore_v_2.0/io_module.f90:191-260, read as readGammaLambda does,
 * :284-373): the lambda file holds L+1 records of chi values, the gamma file for every site,
 * alpha and beta a record of the d real parts of Gamma(alpha, i, beta), then the same of the
 * imaginary parts.  kind 'S' formats values as Fortran's E30.15 (15 significant digits),
 * kind 'B' as B80 (the 64 bits; a restart from it is bit for bit).  gamma / lambda are
 * host arrays in the padded layout of tebd_upload_packed.  Host-only (no device call). */
int tebd_write_state_files(int L, int d, int chi, const double* gamma, const double* lambda, const char* lambda_path,
                           const char* gamma_path, char kind);
int tebd_read_state_files(int L, int d, int chi, double* gamma, double* lambda, const char* lambda_path,
                          const char* gamma_path, char kind);
/* Checkpoint of the context's state to those files at its chi_max (the size the reference's
 * arrays are recorded at), and restart from files recorded at chi (<= chi_max), e.g. by the
 * reference itself.  Every failure -- files missing or short, chi too large -- returns an
 * error (the reference's RestartParallel returns success after a mismatch,
 * Extensions_Core_v_2.0/SitesParallel_Extension_v_1.0/SitesParallelSetup_Module.f90:722-733). */
int tebd_checkpoint(tebd_ctx* ctx, const char* lambda_path, const char* gamma_path, char kind);
int tebd_restart(tebd_ctx* ctx, const char* lambda_path, const char* gamma_path, char kind, int chi);

/* One Trotter step of the context's schedule; *truncerr = sum of the bond truncation
 * errors.  2nd order, open boundaries: odd bonds (gate dt/2), even bonds (dt), odd bonds
 * (dt/2).  The other schedules follow TrotterStep2ndOrderPBC, TrotterStep5thOrder and
 * TrotterStep5thOrderPBC: the periodic edge gate is applied on bond 1 after swapping site L
 * next to site 1 (two-site updates with the SWAP gate on bonds L-1 .. 2, the reference's
 * Swapping, local_operations_module.f90:665-695) and the swaps are undone afterwards; the
 * swaps' truncation errors count.  Replaces TrotterStep (propagation_module.f90:42-614). */
int tebd_trotter_step(tebd_ctx* ctx, double* truncerr);

/* One two-site update with an arbitrary (d*d)x(d*d) operator op2 on bond
 * `link` (1-based, sites link and link+1).  Replaces TwoSiteOp_c
 * (local_operations_module.f90:463-494). */
int tebd_two_site_op(tebd_ctx* ctx, int link, const double* op2, double* truncerr);

/* Re-orthogonalises every bipartition (left-to-right, then right-to-left
 * sweep).  Replaces CanonicalFormAll (propagation_module.f90:617-634).  In an
 * ensemble the same bond of every chain is processed as one batch (batched Jacobi
 * SVDs and GEMMs), and the chains are spread over the lanes (tebd_set_lanes).  With
 * number conservation on it replaces CanonicalFormAllNC (propagation_module.f90:1943-1983):
 * the number-conserving update with the identity on bonds 1 .. L-1, then L-1 .. 1. */
int tebd_canonical_form_all(tebd_ctx* ctx);

/* Re-orthogonalises one bipartition, bond `link` (1-based).  Replaces
 * CanonicalForm (local_operations_module.f90:496-620). */
int tebd_canonical_form(tebd_ctx* ctx, int link);

/* <psi| O_b |psi> for every bond b = 1..L-1, with O_b the d^2 x d^2 operator ops[b-1]
 * (column-major, row (a-1)d+b as the gates) on sites b, b+1 of the canonical state;
 * out[2(b-1)], out[2(b-1)+1] = real and imaginary part.  Replaces the host-side
 * TwoSiteRho / TotalEnergy / OneSiteExpVal evaluation
 * (observables_module.f90:95-162,704-832,1432-1466): Tr(O rho) with
 * rho = theta theta^dagger is formed on the device from the resident state. */
int tebd_expval_two_site(tebd_ctx* ctx, const double* ops, double* out);

/* The propagator table of one chain of L sites for the Trotter schedule (order, bc), from
 * its bond Hamiltonians H ((d*d) x (d*d) complex, column-major, back to back: L-1 for open
 * boundaries, L for periodic ones, H[L-1] coupling sites L and 1).  Entry k of the table
 * U is exp(-i f_k dt H(bond_k)), dt = dt_re + i dt_im (imaginary time: dt = -i dtau), laid
 * out as the reference's NEWConstructPropagators fills it (TEBDtools_module.f90:1079-1193):
 * 2nd order, open: bond i gets f = 1/2 (odd i) or 1 (even i); 2nd order, periodic: also
 * the edge with f = 1, and even bonds f = 1/2 when L is odd; 5th order (Forest-Ruth,
 * theta = 1 / (2 - 2^(1/3)) with the reference's single-precision exponent): two entries
 * per bond, and for periodic chains two (L even) or three (L odd, with four per even
 * bond) for the edge.  Returns the table size (L-1, L, 2(L-1), 2L or 3L) or a negative
 * status.  Host-only (no device call).  Replaces ConstructPropagators and
 * matrix_exponential_c (TEBDtools_module.f90:97-99,176-222), which diagonalise H with
 * ZHEEV; here by complex Jacobi rotations. */
int tebd_construct_propagators(const double* H, int L, int d, int order, char bc, double dt_re, double dt_im,
                               double* U);

/* Imaginary-time ground-state search on the resident state.  Replaces ImagTimeProp and
 * ImagTimePropSpin (propagation_module.f90:636-876): gates exp(-dtau H) built from the
 * nh = L-1 bond Hamiltonians H (nchain * (L-1), chain after chain, in an ensemble);
 * chi stages chi_in and then the context's chi_max (one stage when they are equal),
 * each starting with every bond cut to the stage's chi (entries past chi dropped, as
 * CopyGamLam does) and truncating at that chi; in each stage, for j = 1..max_steps:
 * every steps_for_judge-th j the largest relative change of lambda_l(1) over all bonds
 * since the last judgement is compared with conv1 (stages below chi_max) or conv2 (the
 * chi_max stage, or the only stage) and the stage ends when it is smaller; otherwise one
 * TrotterStep and CanonicalFormAll.  steps[0..3] = {j at which stage 1 ended (max_steps
 * if it did not converge), 1 if it converged, the same for stage 2}.  The gates of the
 * context are replaced by the imaginary-time ones.  The reference's printed totals
 * (TotalNumber, TotalEnergy, TotalOneSite at each judgement) are not formed.  With number
 * conservation on it replaces ImagTimePropNC (propagation_module.f90:1986-2121): the same
 * stages with TrotterStepNC and CanonicalFormAllNC, the labels past a stage's chi reset
 * to 10000 as CopyLabel does. */
int tebd_imag_time_prop(tebd_ctx* ctx, const double* H, int nh, double dtau, int chi_in, int max_steps,
                        int steps_for_judge, double conv1, double conv2, int* steps);

/* Two-site reduced density matrices of every pair of sites l1 < l2 of each chain, and the
 * one-site ones, formed on the device from the resident canonical state the way
 * TwoSiteExpVal and EvaluateMeasures form them (observables_module.f90:836-883,2286-2526):
 * for each right site l2 the environment ThetaKernal(l2) is carried leftwards one site at
 * a time (ThetaNext) and each step's TwoSiteRho is one pair.  With phase (d reals, the
 * diagonal of a one-site operator such as the fermi phase (-1)^n) every ket Gamma of the
 * sites l1 .. l2-1 carries it, as the reference's phaseStat / fermicorr do; NULL for none.
 * rho2 (host): per chain Lc(Lc-1)/2 matrices of d^2 x d^2 complex, column-major, pair
 * (l1, l2) (0-based within the chain) at index l1 Lc - l1(l1+1)/2 + (l2 - l1 - 1); row
 * i d + j, column l d + k for ket i, bra l on site l1 and ket j, bra k on site l2 (the
 * reference's rho2((i-1)d+j,(l-1)d+k)).  rho1 (host, or NULL): L matrices of d x d, element
 * i + d j (FormSingleSiteRho, :95-117).  Tr(O rho) gives any correlator or local value. */
int tebd_two_point_rdms(tebd_ctx* ctx, const double* phase, double* rho2, double* rho1);

/* Singular-value cutoff on the normalised weight s^2/||theta||^2.  Default
 * 1e-30, the reference's rule (keep |s| > 1e-15 after normalisation,
 * FormLambda1, local_operations_module.f90:360-375).  Larger cutoffs shrink
 * the bond dimensions but perturb later observables at order sqrt(cutoff). */
int tebd_set_cutoff(tebd_ctx* ctx, double cutoff);

/* Current bond dimensions chi[l], l = 0..L (chi[0] = chi[L] = 1). */
int tebd_get_bond_dims(tebd_ctx* ctx, int* chi);

/* Counters since creation (n <= 46): [0] bond updates, [1] Jacobi sweeps,
 * [2] sketches of the learned-rank path redone by the full SVD, [3] device
 * milliseconds in trotter steps, [4] sketch columns requested, [5] sketches
 * accepted, [6] sum of the sketches' numerical ranks, [7] rank-model predictions,
 * [8] pair Gram matrices formed, [9] pairs rotated, [10] pair visits of the sweeps
 * run, [11] bonds the rank model sent to the full SVD at once (predicted rank
 * beyond the cap, or too wide to sketch; the fixed width would have sketched
 * them), [12] sketches
 * rejected for their discarded weight (the rest of [2] by the probe check),
 * [13] half-sweeps split over lanes (tebd_set_lanes), [14 + s] (s = 0..15) pairs
 * rotated in sweep s + 1 of a Jacobi solve (the last entry: sweeps 16 and later), and
 * [30 + s] Jacobi solves that ran s + 1 sweeps (the last entry: 16 or more; a solve
 * ends with a sweep that rotates nothing). */
int tebd_get_stats(tebd_ctx* ctx, double* stats, int n);

/* Per-kernel timing with HIP events recorded on the context's stream (no
 * counterpart in the reference, whose only timing is CPU_TIME around whole
 * mains, Case_Studies/Heisenberg_Wrapper/XXDynamics.f90:89,196-198).
 * tebd_set_profiling(ctx, 1) clears and enables the counters;
 * tebd_get_profile fills out[4k .. 4k+3] = {launches, milliseconds, algorithmic
 * flops, algorithmic HBM bytes} of launch class k (n <= 28):
 *   0 theta + gate (k_theta_gate, or k_theta_gate_mixed in TEBD_MODE_MIXED)
 *   1 Jacobi round: pair Grams and 32x32 eigenproblems (k_jacobi_round)
 *   2 truncation, Gamma rebuild, singular-vector and sketch formation
 *   3 pivoted-QR panel factorisation (k_qrp_panel_reg / k_qrp_panel)
 *   4 batched complex GEMMs of the learned-rank path (k_zgemm)
 *   5 Jacobi rotation of W and V (k_jacobi_rotate)
 *   6 pivoted-QR trailing update (k_qrp_trailing_reg; k_qrp_reflect<0> above 1024 rows)
 * Each launch of classes 1, 3, 4, 5, 6 is one kernel, so its event duration
 * is that kernel's duration. */
int tebd_set_profiling(tebd_ctx* ctx, int on);
int tebd_get_profile(tebd_ctx* ctx, double* out, int n);

/* hipGraph replay of the one-sided Jacobi sweeps (no counterpart in the reference):
 * with on = 1 (the default) each sweep's launches (sweep start, and the Gram /
 * eigenproblem and rotation kernels of its 2P-1 rounds) are captured once per launch
 * shape into an executable graph and replayed, instead of being launched one by one;
 * the convergence test between sweeps stays on the host.  Per-launch profiling
 * (tebd_set_profiling) launches eagerly.  Results are identical either way. */
int tebd_set_graphs(tebd_ctx* ctx, int on);

/* Host concurrency of a batched context (no counterpart in the reference): with
 * n lanes (1 to 4, default 3) the bonds of each half-sweep are dealt out to n parts in
 * turn (bond i of the batch to part n - 1 - i mod n), which run on n HIP streams (the context's own and n - 1 more), each driven by
 * its own host thread, so that one part's latency-bound phases (QR panels, inner
 * Jacobi solves) overlap the others' work; n = 1 runs a half-sweep as one batch on the
 * context's stream.  HIP runs a process's streams on a few hardware queues (four
 * unless GPU_MAX_HW_QUEUES says otherwise), and busy streams that share one run in
 * turn: more lanes than free queues make the step slower, not faster.  A half-sweep
 * is split only when it has at least 4 bonds per lane and a bond matrix with a side
 * of at least 512 (smaller ones run faster as one batch).  The chains of an ensemble
 * are also distributed over the lanes in tebd_canonical_form_all.  Every bond's
 * result is the same either way (a bond's arithmetic does not depend on the bonds
 * that share its batch). */
int tebd_set_lanes(tebd_ctx* ctx, int n);

/* Autotuner of launch shapes (no counterpart in the reference).  The batched
 * complex GEMMs of the learned-rank path and of CanonicalForm have several tile
 * variants; with autotuning on (the default) the first launch of every shape class
 * (M and N rounded up to 64, batch size to a power of two) times each candidate on
 * the real operands with HIP events and keeps the fastest for that class.  The
 * candidates take the same floating-point operations in the same order per element,
 * so tuning changes the speed, never the result.  on = 0 switches tuning off (new
 * classes take the 64 x 64 MFMA tiles), on = 1 on, on = -1 on and forgets what was
 * tuned.  tebd_autotune_info: out[0] shape classes tuned, out[1] timing launches,
 * out[2 + v] classes that chose variant v (0 VALU 64x64, 1 MFMA 64x64, 2 MFMA
 * 128x64, 3 MFMA 64x96, 4 MFMA 64x64 with k chunks of 32, 5 MFMA 128x96 on eight
 * waves); n <= 8. */
int tebd_set_autotune(tebd_ctx* ctx, int on);
int tebd_autotune_info(tebd_ctx* ctx, double* out, int n);

/* The two-site update of tebd_two_site_op on caller-provided device buffers, with no
 * allocation of any kind.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Inputs, all device pointers, in the packed layout of a context:
 * gamma_l chi_l x d x chi_c and gamma_r chi_c x d x chi_r complex (element a + rows*(i + d*b)),
 * lambda_l/c/r (chi_l, chi_c, chi_r reals), op2 (d*d x d*d complex, column-major, row
 * (a-1)d+b).  Outputs: gamma_l_out (chi_l x d x k) and gamma_r_out (k x d x chi_r, leading
 * dimension k) with k <= chi_cap the kept rank, each with room for chi_cap columns or rows;
 * lambda_c_out (chi_cap reals, zero past k); *chi_c_new = k and *truncerr on the host.  The
 * outputs may alias the inputs.  Truncation is the reference's (at most chi_cap values,
 * normalised |s| > 1e-15).  The workspace is one fixed device array of 6 MiB per GPU, held
 * until the process ends: d*chi_l, d*chi_r, chi_c and chi_cap must each be at most 256
 * (TEBD_ERR_SHAPE otherwise, before any device call); d in {2, 3, 4, 5, 6, 8}.  Every call on
 * a GPU shares that GPU's array and would race with an overlapping one, so the library
 * serialises the calls of a process: a call queues its work on `stream` (a hipStream_t;
 * null = the default stream) and returns only when that work has finished, and concurrent
 * callers wait for each other.  The array keeps the last call's contents; each call clears
 * it first. */
int tebd_two_site_op_noalloc(int d, int chi_l, int chi_c, int chi_r, int chi_cap, const double* gamma_l,
                             const double* gamma_r, const double* lambda_l, const double* lambda_c,
                             const double* lambda_r, const double* op2, double* gamma_l_out, double* gamma_r_out,
                             double* lambda_c_out, int* chi_c_new, double* truncerr, void* stream);

/* Device memory of a context (no counterpart in the reference, whose workspace is the
 * module-global SVD arrays of SVDInit, local_operations_module.f90:84-122): out[0] the
 * per-bond workspace slots, i.e. the bonds one batched launch sequence processes (every
 * bond of a half-sweep, up to a budget of 48 GiB of slots; longer half-sweeps run in
 * batches of this many bonds), out[1] bytes per slot, out[2] bytes of the Vidal state
 * and gates, out[3] out[2] + all slots; n <= 4. */
int tebd_memory_info(tebd_ctx* ctx, double* out, int n);

/* Blocks until the context's stream is idle. */
int tebd_sync(tebd_ctx* ctx);

const char* tebd_last_error(const tebd_ctx* ctx);

/* ---- number conservation (U(1) block-sparse update) ----------------------------
 * Replaces TwoSiteOpNC / TrotterStepNC / SwappingNC (local_operations_module.f90:748-2082,
 * propagation_module.f90:890-1942).  With conservation on, every two-site update of
 * tebd_two_site_op and tebd_trotter_step (orders 2 and 5, open or periodic boundaries; the
 * periodic schedules' swaps as SwappingNC, without normalising theta) keeps the numbers of
 * the Schmidt states: theta is block diagonal in the number on the left,
 * each block is factorised on its own (tasks of the batched SVD), the values of all blocks
 * are ordered together, states at or below 10.0**(-10) (the reference's single precision
 * literal) are dropped, and the labels of the new states are updated as UpdateLabelLeft /
 * UpdateLabelRight do.
 * tebd_set_conserved: conserv[i] the number of local state i (NULL: i, the reference
 * without intDegFree; the model's Conserv with it), tot_num the total (totNum);
 * conserv = NULL and tot_num < 0 switches conservation off.  One chain only (no ensemble).
 * tebd_set_labels / tebd_get_labels: LabelLeft / LabelRight ((L+1) x chi ints, row l the
 * reference's link l+1, 10000 = an unused state), e.g. from ProductStateLabels
 * (Hamiltonian_tools_module.f90:124-178).  tebd_canonical_form_all and tebd_imag_time_prop
 * then run their number-conserving forms; tebd_canonical_form (one bond) is refused.
 * In the batched mode the first number-conserving update also allocates a second SVD
 * workspace of 512 slots for the number blocks of up to min(512, d chi_max) rows (about
 * 8.8 GiB when d chi_max >= 512), held until tebd_destroy. */
int tebd_set_conserved(tebd_ctx* ctx, const int* conserv, int tot_num);
int tebd_set_labels(tebd_ctx* ctx, const int* label_left, const int* label_right, int chi);
int tebd_get_labels(tebd_ctx* ctx, int* label_left, int* label_right, int chi);

/* ---- rank model of the learned-rank path -------------------------------------
 * No counterpart in the reference, which always takes the full SVD of every bond
 * (local_operations_module.f90:322-345).  The learned-rank path takes the SVD of theta
 * from a sketch theta Omega, Omega = [pre-gate right Schmidt vectors | Gaussian] with
 * width l: without a model l = chi_c + max(4, chi_c / 8) + 8 (chi_c: the bond's
 * dimension); with one, a 66 -> 128 -> 128 -> 1 FP32 perceptron predicts the rank k the
 * update needs, as log2(k / chi_c): the rank it keeps, or 2 kmax (kmax = min(chi_max,
 * d chi_l, d chi_r)) when no sketch up to the cap passes (the update discards more
 * than tol_trunc, or its sketch kept kmax values and failed the probe check).
 * Then l = k + max(4, rank_slack k) + 8, rounded up to 32; bonds with k above sqrt(2)
 * kmax or 2 l above the smaller side of theta take the full SVD at once.  Inputs: the
 * last (smallest) 63 Schmidt values of the bond before the update (log10 / 10, -3 in
 * front when fewer), log2(kmax / chi_c) / 4, log2(chi_c) / 12 and log10(||G - 1||_F) / 4
 * of the bond's gate G.  Parameters: W1[128][66], b1[128], W2[128][128], b2[128],
 * W3[128], b3 (row-major floats). */
#define TEBD_MLP_INPUTS 66
#define TEBD_MLP_HIDDEN 128
#define TEBD_MLP_PARAMS (TEBD_MLP_HIDDEN * TEBD_MLP_INPUTS + TEBD_MLP_HIDDEN + TEBD_MLP_HIDDEN * TEBD_MLP_HIDDEN + \
                         TEBD_MLP_HIDDEN + TEBD_MLP_HIDDEN + 1)

/* Learned-rank path settings.  A sketch is kept when its discarded weight (missed
 * weight included) is at most tol_trunc ||theta||^2 (default 1e-10) and it passes the
 * probe check.  With Q the sketch's range basis, E = theta - Q Q^H theta and kept triplets
 * (s_j, u_j, v_j), j < k, two sets of 16 Gaussian probes estimate eta = sum_j ||E v_j||^2
 * (how far the kept vectors are from singular vectors of theta: to first order the
 * state's squared error) and w = ||E||_F^2 (the weight outside the sketch); with 10x
 * each estimate (a tenth of the true value is undershot with probability < 2e-11):
 * eta <= tol_missed ||theta||^2 (default 1e-20), and w <= tol_missed ||theta||^2 -- or,
 * when k = kmax = min(chi_max, d chi_l, d chi_r) (the cap binds, and the full SVD
 * discards the weight beyond it too), the unscaled estimate of w at most s_k^2, so that
 * no missed direction ten times heavier than the smallest kept one goes undetected.
 * Otherwise the bond is redone by the full SVD.  rank_slack (default 1/16) widens the
 * model's predicted rank. */
int tebd_set_learned_params(tebd_ctx* ctx, double tol_missed, double tol_trunc, double rank_slack);
/* Install (n = TEBD_MLP_PARAMS) or remove (params = NULL or n = 0) the context's rank
 * model; without one every sketch has the fixed width above. */
int tebd_rank_model_set(tebd_ctx* ctx, const float* params, int n);
int tebd_rank_model_get(tebd_ctx* ctx, float* params, int n);
/* Record (features, log2(k / chi_c before)) of every bond update of the learned-rank
 * path while on, k the needed rank as above;
 * switching on clears the record.  tebd_rank_log_get copies up to max_n samples (X: max_n x 66,
 * y: max_n) and returns the number recorded (or a negative status). */
int tebd_rank_log(tebd_ctx* ctx, int on);
int tebd_rank_log_get(tebd_ctx* ctx, float* X, float* y, int max_n);
/* The probe checks of the same record: per sketch whose discarded weight passed, 8 values
 * (k, kmax, ||theta||^2, discarded fraction, smallest kept s^2, estimate of
 * sum_j ||E v_j||^2, estimate of ||E||_F^2, accepted); returns the number of records. */
int tebd_sketch_log_get(tebd_ctx* ctx, double* out, int max_n);
/* Stand-alone model kernels on the current device (no context):
 * mean loss sum w (model(X_s) - y_s)^2 / n, w = w_under where the model is below y;
 * its gradient; Adam training for `epochs` full-batch steps (loss_history: epochs
 * values or NULL); inference. */
int tebd_mlp_gradient(const float* params, const float* X, const float* y, int n, float w_under, float* grad,
                      float* loss);
int tebd_mlp_train(float* params, const float* X, const float* y, int n, int epochs, float lr, float w_under,
                   float* loss_history);
int tebd_mlp_predict(const float* params, const float* X, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif /* TEBD_HIP_H */
