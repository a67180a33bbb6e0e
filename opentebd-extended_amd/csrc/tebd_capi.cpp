// Host runtime behind the C ABI (include/tebd_hip.h): device-resident Vidal
// state, the 2nd-order Trotter schedule and the per-bond kernel pipeline
//   theta+gate -> one-sided block-Jacobi sweeps -> truncation -> Gamma rebuild.
#include "../../include/tebd_hip.h"
#include "tebd_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int kMaxLanes = 4;
constexpr int kLaneMinSide = 512;  // lanes split a batch only if a bond matrix has a side >= this
constexpr int kStages = 6;  // task arrays per batch: 0-4 the learned path's stages, 5 the QR subset
constexpr int kGemmRegions = 8;  // GEMM descriptor arrays per batch (learned path: 0-2 sketch, 3-7 probe check)
// Safety factor of the sketch's probe check.  With p = kProbes complex Gaussian probes an
// estimate ||M G||_F^2 / p of ||M||_F^2 is sum_i mu_i^2 Gamma_i(p, 1) / p over M's
// singular values mu_i, least concentrated when one mu_i dominates; then it falls below
// 1/kProbeSafety of the true value with probability
// P(Gamma(16, 1) < 1.6) = P(Poisson(1.6) >= 16) < 2e-11.
constexpr double kProbeSafety = 10.0;
constexpr int kSweepHist = 16;                      // sweeps tallied one by one in the histograms
constexpr int kNumStats = 14 + 2 * kSweepHist;
constexpr int kNumProf = 28;
constexpr double kQrTol = 1e-17;  // QR stops when every remaining column norm is <= kQrTol ||theta||_F
constexpr int kQrPanel = 16;      // panel width of the blocked pivoted QR (tebd_qrp.hip)
constexpr int kQrCheck = 8;       // panels between checks whether every task has stopped
constexpr int kQrMinCols = 64;    // bond matrices with fewer columns go straight to the Jacobi
constexpr int kMaxSweeps = 40;
constexpr double kTolRel = 1e-13;   // relative orthogonality of converged column pairs
constexpr double kTolAbs = 1e-13;   // rounding floor: |A_pq| <= kTolAbs ||theta|| (|w_p| + |w_q|)
// Every outer sweep solves each pair's 32 x 32 problem with at most kInnerSweepCap inner
// sweeps (fewer when one finds nothing to rotate): a pair left coupled is revisited by the
// next outer sweep.  Measured on Fermi-Hubbard bonds at 1024 x 288: one inner sweep per
// visit doubles the outer sweeps; three take 4.9 outer sweeps per solve against 6.0 with
// two, and more than three add inner rounds without saving outer sweeps.
constexpr int kInnerSweepCap = 3;
constexpr double kSlotBudget = 48.0 * (1ull << 30);  // bytes of per-bond workspace slots per context
constexpr size_t kMaxSweepGraphs = 512;  // executable Jacobi sweep graphs kept per context or lane
constexpr int kGraphRowQuantum = 64;     // = the rotation kernel's slab height (tebd_jacobi.hip)
// The number-conserving update's block SVDs in the batched mode: blocks of up to
// kNcSmallSide rows run kNcSmallSlots at a time in a workspace of that many small slots
// (a half-sweep of Fermi-Hubbard bonds at chi = 256 has about a dozen blocks per bond,
// which the bond-sized slots would take ntask_max = L / 2 at a time).
constexpr int kNcSmallSlots = 512;
constexpr int kNcSmallSide = 512;

int round_up(int x, int m) { return ((x + m - 1) / m) * m; }

}  // namespace

struct tebd_ctx {
  int L = 0, d = 0, chi_max = 0;
  // an ensemble holds nchain independent chains of chain_len sites as one chain of
  // L = nchain * chain_len sites whose bonds between chains are never updated
  int chain_len = 0, nchain = 1;
  // Trotter schedule: order 2 or 5 (Forest-Ruth), open ('O') or periodic ('P') boundaries;
  // `table` gates per chain (the reference's propagator table for that schedule)
  int order = 2;
  char bc = 'O';
  int table = 0;
  unsigned mode = 0;
  double cutoff = 1e-30;
  // chi stage of imaginary-time propagation (tebd_imag_time_prop): the truncation keeps at
  // most this many values while it is below chi_max (0 = chi_max)
  int stage_chi = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device state
  zc* gam = nullptr;      // L slabs of chi_max*d*chi_max
  double* lam = nullptr;  // (L+1) x chi_max
  zc* gates = nullptr;    // nchain x table gates of d^4, then tebd_two_site_op's operator, then SWAP
  bool have_gates = false, have_state = false;
  std::vector<int> chi;   // bond dimensions, host mirror
  // per-task workspace
  int ntask_max = 0;
  int Rmax = 0, Cpmax = 0;
  zc* W = nullptr;
  zc* V = nullptr;
  // pivoted-QR preconditioning workspace, per task slot
  zc* X = nullptr;               // Cpmax x Cpmax: R^dagger, the Jacobi input
  zc* Yr = nullptr;              // Rmax x chi_max
  zc* Zc = nullptr;              // Cpmax x chi_max
  double* qnrm = nullptr;        // Cpmax
  int* qpiv = nullptr;           // Cpmax
  zc* tau = nullptr;             // Cpmax
  double* qstate = nullptr;      // 2
  zc* Tq = nullptr;              // per slot: the compact-WY factors of up to Cpmax / 16 + 1 panels
  double* h_qstate = nullptr;    // pinned, 2 per task slot
  double* s = nullptr;
  int* perm = nullptr;
  int* counters = nullptr;       // 4 per task slot (BondTask::counter)
  unsigned* blkver = nullptr;    // nblk_max per task slot
  unsigned long long* pairseen = nullptr;  // nblk_max^2 per task slot
  int nblk_max = 0;
  zc* Qg = nullptr;              // per task slot: (nblk_max / 2) pair rotations of 32 x 32
  int* wl_count = nullptr;       // 2: rotation queue lengths by round parity
  int2* wl_items = nullptr;      // 2 x wl_cap queued (task, pair)
  int wl_cap = 0;
  double* results = nullptr;
  double* results2 = nullptr;    // learned-rank path: per task ||theta||^2 and column norms of theta
  zc* obs_ops = nullptr;         // (L-1) x d^4 operators of tebd_expval_two_site, allocated on first use
  // all-pairs reduced density matrices (tebd_two_point_rdms), allocated on first use:
  // environments Theta (two, chi^2 d^2), T (chi^2 d^3), three chi d chi factors, row scales,
  // the phase, and the device copies of the pair (d^4) and site (d^2) matrices
  zc* cr_theta[2] = {nullptr, nullptr};
  zc *cr_T = nullptr, *cr_K = nullptr, *cr_B = nullptr, *cr_A = nullptr, *cr_rho2 = nullptr, *cr_rho1 = nullptr;
  double *cr_s = nullptr, *cr_phase = nullptr;
  zc* ident = nullptr;           // d^4 identity gate
  // rank model (learned-rank path) and the training record of the full path
  float* mlp = nullptr;          // TEBD_MLP_PARAMS on the device
  bool have_model = false;
  float* d_feat = nullptr;       // ntask_max x TEBD_MLP_INPUTS
  float* d_pred = nullptr;       // ntask_max
  float* h_feat = nullptr;       // pinned copies
  float* h_pred = nullptr;
  bool rank_log = false;
  std::vector<float> log_X, log_y;
  // with the record on, per probe-checked sketch: k, kmax, ||theta||^2, discarded fraction,
  // smallest kept s^2, the two probe estimates, accepted
  std::vector<double> log_check;
  BondTask* d_tasks = nullptr;   // kStages x ntask_max
  BondTask* h_tasks = nullptr;   // pinned, kStages x ntask_max
  // sketched (learned-rank) path workspace
  SplitWs split{};               // per task: slices of the split-operand theta (TEBD_MODE_SPLIT)
  zc* Zb = nullptr;              // per task: range basis Q of the sketch (m x l)
  zc* Bb = nullptr;              // per task: theta Omega and its QR, then B^dagger
  double* lam_scratch = nullptr; // per task chi_max: lambda before it is committed
  zc* Pb = nullptr;              // per task: probe workspace of the sketch check (probe_stride())
  ProbeTask* d_probe = nullptr;  // ntask_max probe descriptors
  ProbeTask* h_probe = nullptr;  // pinned
  GemmTask* d_gemm = nullptr;
  GemmTask* h_gemm = nullptr;    // pinned
  int* d_kprime = nullptr;
  int* h_kprime = nullptr;       // pinned
  unsigned long long sketch_seed = 0x5DEECE66Dull;
  // a sketch is accepted when the probe estimate of sum_j s_j^2 ||E v_j||^2 (E: theta outside
  // the sketch, v_j: kept right vectors), times kProbeSafety, is at most tol_missed ||theta||^4,
  // and the discarded weight at most tol_trunc ||theta||^2
  double tol_missed = 1e-20;
  double tol_trunc = 1e-10;
  int oversample = 8;             // sketch columns beyond the expected rank
  // with a model: width = predicted rank k + max(4, rank_slack k) + oversample
  double rank_slack = 1.0 / 16.0;
  long long step_index = 0;
  // CanonicalForm workspace (allocated on first use): 8 matrices, 3 (s, perm) pairs
  zc* cf = nullptr;
  double* cf_s = nullptr;
  int* cf_perm = nullptr;
  double* cf_lam = nullptr;
  int cf_sets = 0;               // bonds the CanonicalForm workspace holds at once
  // number-conserving update (tebd_set_conserved): the number of every local state, the
  // total, and on the host the labels of every Schmidt state -- number left / right of each
  // link, (L+1) x chi_max, 10000 = unused (the reference's LabelLeft / LabelRight)
  bool nc = false;
  int nc_tot = 0, nc_maxf = 0;
  std::vector<int> nc_conserv, nc_LL, nc_LR;
  // NC workspace (allocated on first use): per bond of a batch theta and the blocks'
  // singular vectors; index lists, block and kept-state descriptors; per slot the
  // truncation's lambda; pinned staging of singular values and lambdas
  zc* nc_theta = nullptr;
  zc* nc_vec = nullptr;
  int nc_bonds = 0;
  // the small-block SVD workspace (batched mode, made on first use): a context of its own
  // of which only the task slots, the Jacobi / QR state, the stream and the pinned
  // staging of singular values are used
  tebd_ctx* nc_small = nullptr;
  int* d_nc_idx = nullptr;
  size_t nc_idx_cap = 0;
  NcBlock* d_nc_blk = nullptr;
  size_t nc_blk_cap = 0;
  NcKept* d_nc_kept = nullptr;
  size_t nc_kept_cap = 0;
  double* d_nc_part = nullptr;
  size_t nc_part_cap = 0;
  double* nc_lscratch = nullptr;
  double* h_nc_s = nullptr;      // pinned, ntask_max x Cpmax
  double* h_nc_lam = nullptr;    // pinned, ntask_max x chi_max
  double* h_nc_part = nullptr;   // pinned, grows with d_nc_part
  std::vector<int> last_rank;    // kept rank of each bond's last update (rank model input)
  std::vector<double> last_trunc;  // discarded weight (fraction) of each bond's last update
  double* h_results = nullptr;   // pinned
  int* h_counters = nullptr;     // pinned, 4 per task slot
  std::vector<zc> staging;
  std::string err;
  // [0] bond updates, [1] Jacobi sweeps, [2] sketch fallbacks, [3] device ms of trotter steps,
  // [4] sketch columns requested, [5] sketches accepted, [6] sketch ranks (sum of r_y), [7] rank-model predictions,
  // [8] pair Gram matrices formed, [9] pairs rotated, [10] pair visits of the sweeps run,
  // [11] bonds the rank model sent to the full SVD at once (predicted rank beyond the cap or too wide
  // to sketch), [12] sketches
  // rejected for discarded weight (the others of [2] by the probe check), [13] half-sweeps split
  // over lanes, [14 + s] pairs rotated in sweep s of a Jacobi call (s < kSweepHist - 1; the
  // last entry sums the later sweeps), [14 + kSweepHist + s] Jacobi calls that ran s + 1
  // sweeps (the last entry: that many or more)
  double stats[kNumStats] = {};
  // optional per-kernel timing with HIP events on the context's stream:
  // prof[k*4 + 0] launches, [1] milliseconds, [2] algorithmic flops, [3] algorithmic HBM bytes,
  // for k = 0 theta+gate, 1 Jacobi round (Gram + 32x32 eigenproblem), 2 truncation + Gamma
  // rebuild, 3 pivoted-QR panel factorisation, 4 sketch construction and GEMMs of the
  // learned-rank path, 5 Jacobi rotation of W and V, 6 pivoted-QR trailing update
  bool profiling = false;
  double prof[kNumProf] = {};
  // autotuner: launch variant of each tunable kernel per shape class (key), chosen by
  // timing every candidate on the first launch of that class
  bool autotune = true;
  std::map<uint64_t, int> tuned;
  hipEvent_t tev0 = nullptr, tev1 = nullptr;
  double tune_trials = 0;
  // lanes split the bonds of a batched half-sweep between up to kMaxLanes streams, each
  // driven by its own host thread (lanes[] are shallow copies whose per-task workspace
  // starts at their first slot); a lane's own lanes[] are null
  // hipGraphs of whole Jacobi sweeps (sweep start, every round and rotation), keyed by
  // the launch arguments; used while per-launch profiling is off
  bool graphs = true;
  std::map<std::array<long long, 7>, hipGraphExec_t> sweep_graphs;
  bool is_lane = false;
  bool owns_gemm = false;  // a lane that allocated its own GEMM descriptors
  bool owns_stream = true;  // false for lane 0, which runs on its context's stream
  int nlanes = 3;
  int lanes_made = 0;
  tebd_ctx* lanes[kMaxLanes] = {};
  struct Pending { hipEvent_t a, b; int kind; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;

  size_t gslab() const { return (size_t)chi_max * d * chi_max; }
  int cap() const { return stage_chi > 0 && stage_chi < chi_max ? stage_chi : chi_max; }
  size_t d4() const { return (size_t)d * d * d * d; }
};

namespace {

int fail(tebd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                               \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail((c), TEBD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

hipEvent_t pool_event(tebd_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with events when profiling is on.
struct ProfScope {
  tebd_ctx* c;
  int kind;
  hipEvent_t a = nullptr;
  ProfScope(tebd_ctx* cc, int k) : c(cc), kind(k) {
    if (c->profiling && (a = pool_event(c))) (void)hipEventRecord(a, c->st);
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = pool_event(c);
    if (!b) { c->event_pool.push_back(a); return; }
    (void)hipEventRecord(b, c->st);
    c->pending.push_back({a, b, kind});
  }
};

// Called after a stream synchronisation: folds finished event pairs into prof[].
void drain_profile(tebd_ctx* c) {
  for (auto& p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->prof[4 * p.kind + 0] += 1;
      c->prof[4 * p.kind + 1] += ms;
    }
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
}

// Bond `link` (0-based, sites link and link + 1) joins two chains of an ensemble.
bool is_cut(const tebd_ctx* c, int link) { return (link + 1) % c->chain_len == 0; }

// theta + gate of a batch: FP32 matrix cores in the mixed mode, int8 matrix cores with exact
// accumulation in the split mode, FP64 matrix cores otherwise.
hipError_t theta_gate(tebd_ctx* c, const BondTask* d_t, int nb, int max_chil, int max_chir) {
  if (c->mode & TEBD_MODE_MIXED) return launch_theta_gate_mixed(d_t, nb, c->d, max_chil, max_chir, c->st);
  if (c->mode & TEBD_MODE_SPLIT) {
    if (nb > c->ntask_max) return hipErrorInvalidValue;
    return launch_theta_gate_split(d_t, nb, c->d, max_chil, max_chir, c->split, c->st);
  }
  return launch_theta_gate_mfma(d_t, nb, c->d, max_chil, max_chir, c->st);
}

// Gives task `slot` of a batch its Jacobi counters and block bookkeeping.
void attach_jacobi(tebd_ctx* c, BondTask& t, int slot) {
  t.counter = c->counters + 4 * slot;
  t.blkver = c->blkver + (size_t)slot * c->nblk_max;
  t.pairseen = c->pairseen + (size_t)slot * c->nblk_max * c->nblk_max;
  t.Qg = c->Qg + (size_t)slot * (c->nblk_max / 2) * JC * JC;
}

void clear_graphs(tebd_ctx* c) {
  for (auto& kv : c->sweep_graphs) (void)hipGraphExecDestroy(kv.second);
  c->sweep_graphs.clear();
}

// One Jacobi sweep (sweep start, then every round's Gram/eigenproblem kernel and its
// rotation kernel) captured from the context's stream as an executable graph; null if
// the capture fails.  The captured launches are the eager ones with the same arguments.
hipGraphExec_t capture_sweep(tebd_ctx* c, BondTask* d_t, int nb, int Pmax, int cls, int inner, int rows_w,
                             int rows_v) {
  if (hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  hipError_t e = launch_jacobi_sweep_begin(d_t, nb, cls, c->wl_count, c->st);
  for (int r = 0; r < 2 * Pmax - 1 && e == hipSuccess; ++r) {
    e = launch_jacobi_round(d_t, nb, Pmax, cls, r, kTolRel, kTolAbs, rows_w, rows_v, c->wl_count, c->wl_items,
                            c->wl_cap, inner, c->st);
    if (e == hipSuccess)
      e = launch_jacobi_rotate(d_t, r, rows_w, rows_v, c->wl_count, c->wl_items, c->wl_cap, c->st);
  }
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(c->st, &g);
  hipGraphExec_t ge = nullptr;
  if (e == hipSuccess && ee == hipSuccess && g && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
    ge = nullptr;
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();  // a failed capture leaves no sticky error behind
  return ge;
}

// Jacobi sweeps over the tasks of one stage (task i uses slot i) until a sweep
// rotates nothing.
int jacobi_sweeps(tebd_ctx* c, BondTask* d_t, const BondTask* h_t, int nb, int Pmax) {
  HIPCHK(c, launch_jacobi_begin(d_t, nb, c->st));
  int rows_w = 1, rows_v = 0;
  for (int i = 0; i < nb; ++i) {
    rows_w = std::max(rows_w, h_t[i].R);
    if (h_t[i].V) rows_v = std::max(rows_v, h_t[i].NV);
  }
  // The rotation kernel reads the rows only as a count of 64-row slabs and every slab
  // checks its task's own row count, so rounding up to whole slabs launches the same
  // work and lets one captured graph serve every row count of a slab class.
  rows_w = round_up(rows_w, kGraphRowQuantum);
  rows_v = round_up(rows_v, kGraphRowQuantum);
  int sweeps = 0;
  bool converged = false;
  while (sweeps < kMaxSweeps && !converged) {
    const int inner = kInnerSweepCap;
    // the kernels read the sweep index only as (sweep > 0) and its parity, so 0, 1 and 2
    // stand for every sweep of their class and one graph serves the whole class
    const int cls = sweeps == 0 ? 0 : ((sweeps & 1) ? 1 : 2);
    bool done = false;
    if (c->graphs && !c->profiling) {
      const std::array<long long, 7> key = {(long long)(uintptr_t)d_t, nb, Pmax, rows_w, rows_v, cls, inner};
      auto it = c->sweep_graphs.find(key);
      // A full cache is not emptied while the context runs: executable graphs are
      // destroyed only when the context (or its lane) is freed, after its stream has
      // drained.  Sweeps of shapes beyond the cap launch eagerly.
      if (it == c->sweep_graphs.end() && c->sweep_graphs.size() < kMaxSweepGraphs) {
        hipGraphExec_t ge = capture_sweep(c, d_t, nb, Pmax, cls, inner, rows_w, rows_v);
        if (ge) it = c->sweep_graphs.emplace(key, ge).first;
        else c->graphs = false;  // capture unavailable: launch eagerly from now on
      }
      if (it != c->sweep_graphs.end()) {
        HIPCHK(c, hipGraphLaunch(it->second, c->st));
        done = true;
      }
    }
    if (!done) {
      HIPCHK(c, launch_jacobi_sweep_begin(d_t, nb, cls, c->wl_count, c->st));
      for (int r = 0; r < 2 * Pmax - 1; ++r) {
        {
          ProfScope ps(c, 1);
          HIPCHK(c, launch_jacobi_round(d_t, nb, Pmax, cls, r, kTolRel, kTolAbs, rows_w, rows_v, c->wl_count,
                                        c->wl_items, c->wl_cap, inner, c->st));
        }
        ProfScope ps(c, 5);
        HIPCHK(c, launch_jacobi_rotate(d_t, r, rows_w, rows_v, c->wl_count, c->wl_items, c->wl_cap, c->st));
      }
    }
    HIPCHK(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(int) * 4 * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    int rot = 0;
    for (int i = 0; i < nb; ++i) rot += c->h_counters[4 * i + (sweeps & 1)];
    converged = rot == 0;
    c->stats[14 + std::min(sweeps, kSweepHist - 1)] += rot;
    ++sweeps;
  }
  c->stats[1] += sweeps;
  c->stats[14 + kSweepHist + std::min(sweeps, kSweepHist) - 1] += 1;
  for (int i = 0; i < nb; ++i) {
    const BondTask& t = h_t[i];
    const double visits = c->h_counters[4 * i + 2], rots = c->h_counters[4 * i + 3];
    const double rows = t.R + (t.V ? t.NV : 0);
    c->stats[8] += visits;
    c->stats[9] += rots;
    c->stats[10] += (double)sweeps * t.P * (2 * t.P - 1);
    if (c->profiling) {
      // k_jacobi_round, per formed Gram: the 16x16 cross block over R rows (8 flops per
      // complex MAC) reading the pair's JC columns once, and per rotated pair its JC x JC
      // rotation written; k_jacobi_rotate, per rotated pair: the JC columns of W and V
      // read and written once, JC complex MACs per element
      c->prof[6] += visits * 8.0 * t.R * JB * JB;
      c->prof[7] += visits * 16.0 * t.R * JC + rots * 16.0 * JC * JC;
      c->prof[22] += rots * 8.0 * rows * JC * JC;
      c->prof[23] += rots * 32.0 * rows * JC;
    }
  }
  if (!converged) return fail(c, TEBD_ERR_NOCONV, "one-sided Jacobi did not converge");
  return TEBD_OK;
}

// Compact-WY factors kept per slot: one kQrPanel^2 block per panel, min(Rb, Cb) <= Cpmax.
size_t tq_stride(const tebd_ctx* c) { return (size_t)(c->Cpmax / kQrPanel + 1) * kQrPanel * kQrPanel; }
// Probe workspace of one slot (sketch check): Om and Z (n, m x 2p), T (r x p), C1, C2 (k x p).
size_t probe_stride(const tebd_ctx* c) { return (size_t)(4 * c->Rmax + 3 * c->Cpmax) * kProbes; }

// Gives task `slot` its pivoted-QR buffers; Wq is the bond matrix slot W.
void attach_qr(tebd_ctx* c, BondTask& t, int slot) {
  t.Wq = t.W;
  t.Rb = t.R;
  t.Cb = t.C;
  t.qnrm = c->qnrm + (size_t)slot * c->Cpmax;
  t.qpiv = c->qpiv + (size_t)slot * c->Cpmax;
  t.tau = c->tau + (size_t)slot * c->Cpmax;
  t.qstate = c->qstate + 2 * slot;
  t.Tq = c->Tq + (size_t)slot * tq_stride(c);
  t.Yr = c->Yr + (size_t)slot * c->Rmax * c->chi_max;
  t.Zc = c->Zc + (size_t)slot * c->Cpmax * c->chi_max;
}

// Workspace slot of a task (its QR state lives at qstate + 2 slot).
int slot_of(const tebd_ctx* c, const BondTask& t) { return (int)((t.qstate - c->qstate) / 2); }

// Column-pivoted Householder QR of every task's bond matrix, stopped at the
// numerical rank.  Returns the ranks in rank[].
int qrp_factor(tebd_ctx* c, BondTask* D, const BondTask* H, int nb, int max_cb, std::vector<int>& rank) {
  int max_rb = 1;
  for (int i = 0; i < nb; ++i) max_rb = std::max(max_rb, H[i].Rb);
  HIPCHK(c, launch_qrp_begin(D, nb, max_cb, kQrTol, c->st));
  // every task's last panel (j0 + width >= min(Rb, Cb)) records its rank, so after the
  // panel with j0 + kQrPanel >= max_cb all tasks have finished
  bool done = false;
  for (int j0 = 0, panel = 0; !done; j0 += kQrPanel, ++panel) {
    {
      ProfScope ps(c, 3);
      HIPCHK(c, launch_qrp_panel(D, nb, max_rb, max_cb, j0, c->st));
    }
    {
      ProfScope ps(c, 6);
      HIPCHK(c, launch_qrp_trailing(D, nb, max_rb, max_cb, j0, c->st));
    }
    const bool last = j0 + kQrPanel >= max_cb;
    if ((panel + 1) % kQrCheck == 0 || last) {
      HIPCHK(c, hipMemcpyAsync(c->h_qstate, c->qstate, sizeof(double) * 2 * c->ntask_max, hipMemcpyDeviceToHost,
                               c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
      done = true;
      for (int i = 0; i < nb; ++i) done = done && c->h_qstate[2 * slot_of(c, H[i])] >= 0.0;
    }
    if (last && !done) return fail(c, TEBD_ERR_STATE, "pivoted QR did not finish");
  }
  rank.resize(nb);
  for (int i = 0; i < nb; ++i) {
    rank[i] = (int)c->h_qstate[2 * slot_of(c, H[i])];
    if (c->profiling) {
      // per panel j0 of width w: the panel kernel reads and writes the (Rb - j0) x w panel
      // and forms w reflectors of it (~8 rows w^2 flops); the trailing update reads and
      // writes the (Rb - j0) x (Cb - j0 - w) block once, 2 w complex MACs per element
      const BondTask& t = H[i];
      for (int j0 = 0; j0 < rank[i]; j0 += kQrPanel) {
        const int w = std::min(kQrPanel, rank[i] - j0);
        const double rows = t.Rb - j0, trail = std::max(0, t.Cb - j0 - w);
        c->prof[14] += 8.0 * rows * w * w;
        c->prof[15] += 32.0 * rows * w;
        c->prof[26] += 2.0 * 8.0 * w * rows * trail;
        c->prof[27] += 32.0 * rows * trail;
      }
    }
  }
  return TEBD_OK;
}

// SVD + truncation of the tasks' bond matrices (in W; tasks D / H, task i uses slot
// i and results[i]): singular values, kept rank, lambda (into lc_out) and truncerr.
// Bond matrices with at least kQrMinCols columns are first reduced by the pivoted
// QR; the one-sided Jacobi then runs on R^dagger and the kept singular vectors are
// formed into Yr / Zc.  On return h_results holds every task's results and *max_k
// the largest kept rank.
int svd_factor(tebd_ctx* c, BondTask* D, BondTask* H, const int* links, int nb, int Cpmax, int Pmax, bool force_qr,
               bool* used_qr, int* max_k) {
  int max_cb = 1, max_rb = 1;
  for (int i = 0; i < nb; ++i) {
    max_cb = std::max(max_cb, H[i].Cb);
    max_rb = std::max(max_rb, H[i].Rb);
  }
  // The QR preconditioning is decided per task (bond matrices with at least
  // kQrMinCols columns, or all with force_qr), so a bond's result does not depend on
  // which other bonds share its batch.
  std::vector<int> qidx;
  for (int i = 0; i < nb; ++i)
    if (force_qr || H[i].Cb >= kQrMinCols) qidx.push_back(i);
  const int nq = (int)qidx.size();
  const bool use_qr = nq > 0;
  int max_rank = 1;
  if (use_qr) {
    BondTask* Hq = H;
    BondTask* Dq = D;
    int max_cb_q = 1;
    if (nq < nb) {  // the QR tasks as a compact batch of their own
      Hq = c->h_tasks + 5 * c->ntask_max;
      Dq = c->d_tasks + 5 * c->ntask_max;
      for (int j = 0; j < nq; ++j) Hq[j] = H[qidx[j]];
      HIPCHK(c, hipMemcpyAsync(Dq, Hq, sizeof(BondTask) * nq, hipMemcpyHostToDevice, c->st));
    }
    for (int j = 0; j < nq; ++j) max_cb_q = std::max(max_cb_q, Hq[j].Cb);
    std::vector<int> rank;
    int e = qrp_factor(c, Dq, Hq, nq, max_cb_q, rank);
    if (e) return e;
    for (int j = 0; j < nq; ++j) {
      const int i = qidx[j];
      if (rank[j] < 1) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "bond %d: theta is zero or not finite", links[i] + 1);
        return fail(c, TEBD_ERR_NAN, buf);
      }
      max_rank = std::max(max_rank, rank[j]);
      BondTask& t = H[i];
      t.qr = 1;
      t.W = c->X + (size_t)slot_of(c, t) * c->Cpmax * c->Cpmax;
      t.R = t.Cb;
      t.C = rank[j];
      t.Cpad = round_up(t.C, JC);
      t.P = t.Cpad / JC;
      t.NV = t.Cpad;
    }
    Cpmax = 1;
    Pmax = 1;
    for (int i = 0; i < nb; ++i) {
      Cpmax = std::max(Cpmax, H[i].Cpad);
      Pmax = std::max(Pmax, H[i].P);
    }
    HIPCHK(c, hipMemcpyAsync(D, H, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_qrp_build_x(D, nb, c->st));
  }
  HIPCHK(c, launch_prepare_svd(D, nb, c->st));
  HIPCHK(c, launch_frobenius(D, nb, Cpmax, c->st, 0));
  {
    int e = jacobi_sweeps(c, D, H, nb, Pmax);
    if (e) return e;
  }
  {
    ProfScope ps(c, 2);
    HIPCHK(c, launch_finalize(D, nb, Cpmax, c->cutoff, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(c->h_results, c->results, sizeof(double) * RESULT_STRIDE * nb, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  *max_k = 1;
  for (int i = 0; i < nb; ++i) {
    const double* r = c->h_results + RESULT_STRIDE * i;
    const int k = (int)r[0];
    if (!(std::isfinite(r[2])) || r[2] <= 0.0 || k < 1) {
      char buf[160];
      std::snprintf(buf, sizeof(buf), "bond %d: theta is zero or not finite (||theta||^2=%g)", links[i] + 1, r[2]);
      return fail(c, TEBD_ERR_NAN, buf);
    }
    *max_k = std::max(*max_k, k);
    if (c->profiling) {
      // column norms read W; Gamma_l, Gamma_r written from k singular vectors
      const BondTask& t = H[i];
      c->prof[11] += 16.0 * ((double)t.R * t.C + 2.0 * k * t.d * ((double)t.chil + t.chir));
    }
  }
  if (use_qr) {
    ProfScope ps(c, 2);
    HIPCHK(c, launch_qrp_vectors(D, nb, *max_k, max_rank, c->st));
  }
  *used_qr = use_qr;
  return TEBD_OK;
}

// svd_factor of the batch (stage-0 tasks) followed by the new Gammas and lambda.
int svd_truncate(tebd_ctx* c, const int* links, int nb, int Cpmax, int Pmax) {
  bool used_qr = false;
  int max_k = 1;
  int e = svd_factor(c, c->d_tasks, c->h_tasks, links, nb, Cpmax, Pmax, false, &used_qr, &max_k);
  if (e) return e;
  {
    ProfScope ps(c, 2);
    HIPCHK(c, launch_write_gammas(c->d_tasks, nb, c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  drain_profile(c);
  return TEBD_OK;
}

// Algorithmic work of one batched complex GEMM launch (k_zgemm): 8 M N K flops, and
// op(A), B read and C written once.
void account_gemm(tebd_ctx* c, const GemmTask* g, int n) {
  if (!c->profiling) return;
  for (int i = 0; i < n; ++i) {
    c->prof[18] += 8.0 * g[i].M * (double)g[i].N * g[i].K;
    c->prof[19] += 16.0 * ((double)g[i].M * g[i].K + (double)g[i].K * g[i].N + (double)g[i].M * g[i].N);
  }
}

// Shape class of a batched GEMM launch: conj flag, M and N rounded up to 64, the batch
// size rounded up to a power of two.
uint64_t gemm_key(int max_m, int max_n, int ntasks, bool conj) {
  int nb = 1;
  while (nb < ntasks) nb <<= 1;
  return (1ull << 62) | ((uint64_t)conj << 61) | ((uint64_t)((max_m + 63) / 64) << 40) |
         ((uint64_t)((max_n + 63) / 64) << 20) | (uint64_t)nb;
}

// Batched complex GEMM with the tile variant the autotuner chose for this shape class.
// The candidates (MFMA 64x64, 128x64 and 64x96 tiles with k chunks of 16, 64x64 tiles
// with k chunks of 32, and 128x96 tiles on eight waves) take the same floating-point operations
// in the same order for every element, so the choice changes the speed, never the
// result; C = op(A) B is overwritten, so the timing runs leave the same C behind.
hipError_t tuned_zgemm(tebd_ctx* c, const GemmTask* d_t, int nt, int max_m, int max_n, bool conj) {
  static const int cand[] = {ZGEMM_MFMA_64x64, ZGEMM_MFMA_128x64, ZGEMM_MFMA_64x96, ZGEMM_MFMA_64x64_K32,
                             ZGEMM_MFMA_128x96_W8};
  const uint64_t key = gemm_key(max_m, max_n, nt, conj);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) return launch_zgemm(d_t, nt, max_m, max_n, conj, c->st, it->second);
  if (!c->autotune) return launch_zgemm(d_t, nt, max_m, max_n, conj, c->st, ZGEMM_MFMA_64x64);
  if (!c->tev0 && (hipEventCreate(&c->tev0) != hipSuccess || hipEventCreate(&c->tev1) != hipSuccess))
    return hipErrorOutOfMemory;
  int best = cand[0];
  float best_ms = 1e30f;
  for (int v : cand) {
    hipError_t e = hipEventRecord(c->tev0, c->st);
    if (e == hipSuccess) e = launch_zgemm(d_t, nt, max_m, max_n, conj, c->st, v);
    if (e == hipSuccess) e = hipEventRecord(c->tev1, c->st);
    if (e == hipSuccess) e = hipEventSynchronize(c->tev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->tev0, c->tev1);
    if (e != hipSuccess) return e;
    c->tune_trials += 1;
    if (ms < best_ms) { best_ms = ms; best = v; }
  }
  c->tuned[key] = best;
  return hipSuccess;
}

// Sketch width for one bond without a rank model: the bond's current dimension with
// 1/8 growth, plus oversampling.
int sketch_width(tebd_ctx* c, int link) {
  const int chic = c->chi[link + 1];
  const int grow = std::max(4, chic / 8);
  return std::min(c->cap(), chic + grow) + c->oversample;
}

// Process `nb` independent bonds (0-based links) as one batch of tasks.
// Task of a full-SVD bond update of `link` in workspace slot `slot`: theta (or
// theta^dagger when it has more columns than rows) into W, the new Gammas and lambda
// written in place.
void make_full_task(tebd_ctx* c, BondTask& t, int link, const zc* gate, int slot) {
  const int d = c->d;
  std::memset(&t, 0, sizeof(t));
  t.chil = c->chi[link];
  t.chic = c->chi[link + 1];
  t.chir = c->chi[link + 2];
  t.d = d;
  t.link = link;
  const int m = d * t.chil, n = d * t.chir;
  t.transposed = n > m ? 1 : 0;
  t.R = std::max(m, n);
  t.C = std::min(m, n);
  t.Cpad = round_up(t.C, JC);
  t.P = t.Cpad / JC;
  t.gl = c->gam + (size_t)link * c->gslab();
  t.gr = c->gam + (size_t)(link + 1) * c->gslab();
  t.ll = c->lam + (size_t)link * c->chi_max;
  t.lc = c->lam + (size_t)(link + 1) * c->chi_max;
  t.lr = c->lam + (size_t)(link + 2) * c->chi_max;
  t.gate = gate;
  t.W = c->W + (size_t)slot * c->Rmax * c->Cpmax;
  t.V = c->V + (size_t)slot * c->Cpmax * c->Cpmax;
  t.s = c->s + (size_t)slot * c->Cpmax;
  t.perm = c->perm + (size_t)slot * c->Cpmax;
  t.NV = t.Cpad;
  attach_jacobi(c, t, slot);
  attach_qr(c, t, slot);
  t.result = c->results + (size_t)RESULT_STRIDE * slot;
  t.gl_out = const_cast<zc*>(t.gl);
  t.gr_out = const_cast<zc*>(t.gr);
  t.lc_out = const_cast<double*>(t.lc);
  t.chi_max = c->cap();
}

// Algorithmic work of the theta + gate kernel on the tasks (profiling only).
void account_theta(tebd_ctx* c, const BondTask* h, int nb) {
  if (!c->profiling) return;
  const double d = c->d;
  for (int i = 0; i < nb; ++i) {
    const BondTask& t = h[i];
    // theta: 8 d^2 chi_l chi_c chi_r, gate: 8 chi_l chi_r d^4 (complex MAC = 8 flops)
    c->prof[2] += 8.0 * d * d * t.chil * (double)t.chic * t.chir + 8.0 * t.chil * (double)t.chir * d * d * d * d;
    // Gamma_l, Gamma_r read, theta written
    c->prof[3] += 16.0 * ((double)t.chil * d * t.chic + (double)t.chic * d * t.chir + (double)t.chil * d * t.chir * d);
  }
}

int run_bonds_full(tebd_ctx* c, const int* links, int nb, const zc* const* gate_ptrs, double* trunc_sum) {
  int max_chil = 1, max_chir = 1, Pmax = 1, Cpmax = 1;
  for (int i = 0; i < nb; ++i) {
    BondTask& t = c->h_tasks[i];
    make_full_task(c, t, links[i], gate_ptrs[i], i);
    max_chil = std::max(max_chil, t.chil);
    max_chir = std::max(max_chir, t.chir);
    Pmax = std::max(Pmax, t.P);
    Cpmax = std::max(Cpmax, t.Cpad);
  }
  HIPCHK(c, hipMemcpyAsync(c->d_tasks, c->h_tasks, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
  {
    ProfScope ps(c, 0);
    HIPCHK(c, theta_gate(c, c->d_tasks, nb, max_chil, max_chir));
  }
  account_theta(c, c->h_tasks, nb);
  int e = svd_truncate(c, links, nb, Cpmax, Pmax);
  if (e) return e;
  double tsum = 0.0;
  for (int i = 0; i < nb; ++i) {
    const double* r = c->h_results + RESULT_STRIDE * i;
    c->chi[links[i] + 1] = (int)r[0];
    c->last_trunc[links[i] + 1] = r[1];
    tsum += r[1];
  }
  c->stats[0] += nb;
  for (int i = 0; i < nb; ++i) c->last_rank[links[i] + 1] = c->chi[links[i] + 1];
  *trunc_sum = tsum;
  return TEBD_OK;
}


int run_bonds_learned(tebd_ctx* c, const int* links, int nb, const zc* const* gate_ptrs, double* trunc_sum) {
  const int d = c->d, nt = c->ntask_max;
  std::vector<int> sk, full;
  std::vector<int> ell;
  std::vector<int> chic_before(nb);
  for (int i = 0; i < nb; ++i) chic_before[i] = c->chi[links[i] + 1];
  std::vector<char> capped_reject(nb, 0);  // sketch at the cap rejected by the probe check
  // rank model: features of every bond of the batch (for the prediction and / or the
  // record), one launch; the predicted rank sizes each bond's sketch
  const bool use_model = c->have_model;
  if (c->have_model || c->rank_log) {
    BondTask* F = c->h_tasks + 4 * nt;
    for (int i = 0; i < nb; ++i) {
      std::memset(&F[i], 0, sizeof(BondTask));
      F[i].chil = c->chi[links[i]];
      F[i].chic = c->chi[links[i] + 1];
      F[i].chir = c->chi[links[i] + 2];
      F[i].d = d;
      F[i].chi_max = c->cap();
      F[i].lc = c->lam + (size_t)(links[i] + 1) * c->chi_max;
      F[i].gate = gate_ptrs[i];
    }
    BondTask* DF = c->d_tasks + 4 * nt;
    HIPCHK(c, hipMemcpyAsync(DF, F, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_mlp_features(DF, nb, c->d_feat, c->st));
    if (use_model) HIPCHK(c, launch_mlp_predict(c->mlp, c->d_feat, nb, c->d_pred, c->st));
    HIPCHK(c, hipMemcpyAsync(c->h_feat, c->d_feat, sizeof(float) * TEBD_MLP_INPUTS * nb, hipMemcpyDeviceToHost, c->st));
    if (use_model) HIPCHK(c, hipMemcpyAsync(c->h_pred, c->d_pred, sizeof(float) * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (use_model) c->stats[7] += nb;
  }
  for (int i = 0; i < nb; ++i) {
    const int link = links[i];
    const int m = d * c->chi[link], n = d * c->chi[link + 2], side = std::min(m, n);
    const int l_formula = round_up(std::min(side, sketch_width(c, link)), JC);
    int l = l_formula;
    bool beyond = false;
    if (use_model) {
      // predicted rank k = chi_c 2^y; beyond sqrt(2) kmax the model expects the update to
      // discard more than tol_trunc even at the cap, which rejects any sketch
      const int kmax = std::min(c->cap(), side);
      const double y = std::min(std::max((double)c->h_pred[i], -8.0), 12.0);
      const double kraw = c->chi[link + 1] * std::exp2(y);
      beyond = kraw > std::sqrt(2.0) * kmax;
      const double khat = std::min((double)kmax, kraw);
      const int want = std::min(kmax, (int)std::ceil(khat + std::max(4.0, c->rank_slack * khat))) + c->oversample;
      l = round_up(std::min(side, want), JC);
    }
    if (beyond || 2 * l > side) {
      full.push_back(i);
      if (use_model && 2 * l_formula <= side) c->stats[11] += 1;
    } else {
      sk.push_back(i);
      ell.push_back(l);
    }
  }
  double tsum = 0.0;
  const int ns = (int)sk.size();
  // Bonds that need the full SVD and are large enough for its QR preconditioning join
  // the sketched bonds' final SVD stage (same QR and Jacobi launches); the others, and
  // sketches that are rejected, go through run_bonds_full afterwards.
  std::vector<int> merged;
  std::vector<int> fb_links;
  std::vector<const zc*> fb_gates;
  for (int i : full) {
    const int link = links[i];
    const int C = d * std::min(c->chi[link], c->chi[link + 2]);
    if (ns > 0 && C >= kQrMinCols) {
      merged.push_back(i);
    } else {
      fb_links.push_back(link);
      fb_gates.push_back(gate_ptrs[i]);
    }
  }
  const int nf = (int)merged.size();
  if (ns > 0) {
    BondTask *S0 = c->h_tasks, *S1 = S0 + nt, *S2 = S0 + 2 * nt, *S3 = S0 + 3 * nt, *S4 = S0 + 4 * nt;
    BondTask *D0 = c->d_tasks, *D1 = D0 + nt, *D2 = D0 + 2 * nt, *D3 = D0 + 3 * nt, *D4 = D0 + 4 * nt;
    GemmTask* HG = c->h_gemm;
    GemmTask* DG = c->d_gemm;
    int chil_max = 1, chir_max = 1, max_m = 1, max_n = 1, max_l = 1;
    std::vector<int> sk_links(ns);
    for (int q = 0; q < ns; ++q) {
      const int i = sk[q], link = links[i];
      sk_links[q] = link;
      const int chil = c->chi[link], chic = c->chi[link + 1], chir = c->chi[link + 2];
      const int m = d * chil, n = d * chir, l = ell[q];
      BondTask base;
      std::memset(&base, 0, sizeof(base));
      base.gl = c->gam + (size_t)link * c->gslab();
      base.gr = c->gam + (size_t)(link + 1) * c->gslab();
      base.ll = c->lam + (size_t)link * c->chi_max;
      base.lc = c->lam + (size_t)(link + 1) * c->chi_max;
      base.lr = c->lam + (size_t)(link + 2) * c->chi_max;
      base.gate = gate_ptrs[i];
      base.chil = chil; base.chic = chic; base.chir = chir; base.d = d;
      base.link = link;
      base.rng_key = link % c->chain_len;
      base.chi_max = c->cap();
      base.s = c->s + (size_t)q * c->Cpmax;
      base.perm = c->perm + (size_t)q * c->Cpmax;
      zc* Wt = c->W + (size_t)q * c->Rmax * c->Cpmax;   // theta, later the left vectors
      zc* Vt = c->V + (size_t)q * c->Cpmax * c->Cpmax;  // Omega, later the Jacobi V
      zc* Bt = c->Bb + (size_t)q * c->Rmax * c->Cpmax;  // Y and its QR, later B^dagger
      zc* Qt = c->Zb + (size_t)q * c->Rmax * c->Cpmax;  // Q
      // stage 0: theta (m x n) and ||theta||^2 into results2[q][3]
      S0[q] = base;
      S0[q].W = Wt; S0[q].R = m; S0[q].C = n; S0[q].Cpad = round_up(n, JC);
      S0[q].result = c->results2 + (size_t)RESULT_STRIDE * q;
      // stage 1: Omega (n x l)
      S1[q] = base;
      S1[q].V = Vt; S1[q].Cpad = l;
      c->h_kprime[q] = l;
      // stage 2: pivoted QR of Y (m x l), explicit Q
      S2[q] = base;
      S2[q].W = Bt; S2[q].R = m; S2[q].C = l;
      attach_qr(c, S2[q], q);
      S2[q].Qs = Qt;
      S2[q].result = c->results + (size_t)RESULT_STRIDE * q;
      // GEMM 0: Y = theta Omega
      HG[0 * nt + q] = {Wt, Vt, Bt, m, l, n, m, n, m, 1.0, 0};
      max_m = std::max(max_m, m); max_n = std::max(max_n, n); max_l = std::max(max_l, l);
      chil_max = std::max(chil_max, chil); chir_max = std::max(chir_max, chir);
    }
    HIPCHK(c, hipMemcpyAsync(c->d_tasks, c->h_tasks, sizeof(BondTask) * nt * kStages, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_kprime, c->h_kprime, sizeof(int) * ns, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(DG, HG, sizeof(GemmTask) * nt, hipMemcpyHostToDevice, c->st));
    {
      ProfScope ps(c, 0);
      HIPCHK(c, theta_gate(c, D0, ns, chil_max, chir_max));
    }
    account_theta(c, S0, ns);
    HIPCHK(c, launch_frobenius(D0, ns, round_up(max_n, JC), c->st, 1));
    {
      ProfScope ps(c, 2);
      HIPCHK(c, launch_build_sketch(D1, c->d_kprime, ns, c->sketch_seed ^ (unsigned long long)c->step_index, c->st));
    }
    {
      ProfScope ps(c, 4);
      HIPCHK(c, tuned_zgemm(c, DG, ns, max_m, max_l, false));
    }
    account_gemm(c, HG, ns);
    std::vector<int> ry;
    int e = qrp_factor(c, D2, S2, ns, max_l, ry);
    if (e) return e;
    int max_ry = 1;
    for (int q = 0; q < ns; ++q) {
      S2[q].C = std::max(ry[q], 1);
      max_ry = std::max(max_ry, S2[q].C);
    }
    HIPCHK(c, hipMemcpyAsync(D2, S2, sizeof(BondTask) * ns, hipMemcpyHostToDevice, c->st));
    for (int q = 0; q < ns; ++q) {
      const BondTask& b0 = S0[q];
      const int m = b0.R, n = b0.C, r = S2[q].C;
      zc* Bt = c->Bb + (size_t)q * c->Rmax * c->Cpmax;
      // GEMM 1: B^dagger (n x r) = theta^dagger Q
      HG[1 * nt + q] = {b0.W, S2[q].Qs, Bt, n, r, m, m, m, n, 1.0, 0};
      // stage 3: SVD of B^dagger (rows n, columns r)
      BondTask& t = S3[q];
      t = S1[q];
      t.V = c->V + (size_t)q * c->Cpmax * c->Cpmax;
      t.W = Bt; t.R = n; t.C = r; t.Cpad = round_up(r, JC); t.P = t.Cpad / JC; t.NV = t.Cpad;
      attach_jacobi(c, t, q);
      attach_qr(c, t, q);
      t.result = c->results + (size_t)RESULT_STRIDE * q;
      t.lc_out = c->lam_scratch + (size_t)q * c->chi_max;
      t.wtot = c->results2 + (size_t)RESULT_STRIDE * q + 3;
    }
    HIPCHK(c, hipMemcpyAsync(DG + nt, HG + nt, sizeof(GemmTask) * ns, hipMemcpyHostToDevice, c->st));
    {
      ProfScope ps(c, 2);
      HIPCHK(c, launch_qrp_form_q(D2, ns, max_ry, c->st));
    }
    {
      ProfScope ps(c, 4);
      HIPCHK(c, tuned_zgemm(c, DG + nt, ns, max_n, max_ry, true));
    }
    account_gemm(c, HG + nt, ns);
    std::vector<int> all_links = sk_links;
    int fl_max = 1, fr_max = 1;
    for (int j = 0; j < nf; ++j) {
      const int i = merged[j];
      make_full_task(c, S3[ns + j], links[i], gate_ptrs[i], ns + j);
      all_links.push_back(links[i]);
      fl_max = std::max(fl_max, S3[ns + j].chil);
      fr_max = std::max(fr_max, S3[ns + j].chir);
    }
    HIPCHK(c, hipMemcpyAsync(D3, S3, sizeof(BondTask) * (ns + nf), hipMemcpyHostToDevice, c->st));
    if (nf > 0) {
      ProfScope ps(c, 0);
      HIPCHK(c, theta_gate(c, D3 + ns, nf, fl_max, fr_max));
    }
    account_theta(c, S3 + ns, nf);
    int max_k = 1;
    bool used_qr = false;
    e = svd_factor(c, D3, S3, all_links.data(), ns + nf, round_up(max_ry, JC), round_up(max_ry, JC) / JC, true,
                   &used_qr, &max_k);
    if (e) return e;
    for (int j = 0; j < nf; ++j) {
      const double* r = c->h_results + RESULT_STRIDE * (ns + j);
      const int link = links[merged[j]];
      c->chi[link + 1] = (int)r[0];
      c->last_rank[link + 1] = (int)r[0];
      c->last_trunc[link + 1] = r[1];
      tsum += r[1];
    }
    c->stats[0] += nf;
    // Every sketch whose discarded weight passes gets its left vectors Q U_B (U_B: the
    // factor's Cb-side output, r x k) and the probe check of ProbeTask: with E = theta -
    // Q Q^dagger theta, set 1 gives Z1 = theta V G1 - (Q U_B) S G1 = E V G1 (theta^dagger Q U_B
    // = V S holds exactly, so theta v_j - s_j u_j = E v_j), set 2 gives Z2 = E G2.  theta is
    // read by the probe product before Q U_B overwrites it.
    std::vector<int> cand;
    int max_kc = 1, max_kn = 1, max_r = 1;
    for (int q = 0; q < ns; ++q) {
      const double* r = c->h_results + RESULT_STRIDE * q;
      c->stats[4] += ell[q];
      c->stats[6] += S3[q].C;
      if (used_qr && r[1] <= c->tol_trunc) {
        cand.push_back(q);
        max_kc = std::max(max_kc, (int)r[0]);
        max_kn = std::max(max_kn, (int)r[0] + S0[q].C);
        max_r = std::max(max_r, S2[q].C);
      } else {
        c->stats[12] += 1;
      }
    }
    const int nc = (int)cand.size();
    GemmTask* G[kGemmRegions];
    for (int g = 0; g < kGemmRegions; ++g) G[g] = HG + g * nt;
    for (int j = 0; j < nc; ++j) {
      const int q = cand[j];
      const int k = (int)c->h_results[RESULT_STRIDE * q];
      const BondTask& b0 = S0[q];
      const int m = b0.R, n = b0.C, r = S2[q].C;
      zc* Om = c->Pb + (size_t)q * probe_stride(c);  // n x 2p
      zc* Z = Om + (size_t)c->Rmax * 2 * kProbes;      // m x 2p
      zc* T = Z + (size_t)c->Rmax * 2 * kProbes;       // r x p
      zc* C1 = T + (size_t)c->Cpmax * kProbes;         // k x p
      zc* C2 = C1 + (size_t)c->Cpmax * kProbes;        // k x p
      zc* Z2 = Z + (size_t)m * kProbes;
      c->h_probe[j] = {S3[q].s, k, m, n, sk_links[q] % c->chain_len, C1, C2, Om, Z,
                       c->results + (size_t)RESULT_STRIDE * q + 6};
      G[3][j] = {S3[q].Yr, C1, Om, n, kProbes, k, n, k, n, 1.0, 0};                 // Om(:, :p) = V G1
      G[4][j] = {b0.W, Om, Z, m, 2 * kProbes, n, m, n, m, 1.0, 0};                  // Z = theta Om
      G[5][j] = {S2[q].Qs, Z2, T, r, kProbes, m, m, m, r, 1.0, 0};                  // T = Q^dagger Z2
      G[6][j] = {S2[q].Qs, T, Z2, m, kProbes, r, m, r, m, -1.0, 1};                 // Z2 -= Q T
      G[2][j] = {S2[q].Qs, S3[q].Zc, b0.W, m, k, S3[q].Cb, m, S3[q].Cb, m, 1.0, 0};  // theta <- Q U_B
      G[7][j] = {b0.W, C2, Z, m, kProbes, k, m, k, m, -1.0, 1};                     // Z1 -= Q U_B S G1
    }
    if (nc > 0) {
      HIPCHK(c, hipMemcpyAsync(c->d_probe, c->h_probe, sizeof(ProbeTask) * nc, hipMemcpyHostToDevice, c->st));
      for (int g = 2; g < kGemmRegions; ++g)
        HIPCHK(c, hipMemcpyAsync(DG + g * nt, G[g], sizeof(GemmTask) * nc, hipMemcpyHostToDevice, c->st));
      {
        ProfScope ps(c, 2);
        HIPCHK(c, launch_probe_coeffs(c->d_probe, nc, max_kn, c->sketch_seed ^ (unsigned long long)c->step_index,
                                      c->st));
      }
      struct Launch { int region, max_m, max_n; bool conj; };
      const Launch order[6] = {{3, max_n, kProbes, false}, {4, max_m, 2 * kProbes, false}, {5, max_r, kProbes, true},
                               {6, max_m, kProbes, false}, {2, max_m, max_kc, false},     {7, max_m, kProbes, false}};
      for (const Launch& o : order) {
        ProfScope ps(c, 4);
        HIPCHK(c, tuned_zgemm(c, DG + o.region * nt, nc, o.max_m, o.max_n, o.conj));
      }
      for (int g = 2; g < kGemmRegions; ++g) account_gemm(c, G[g], nc);
      {
        ProfScope ps(c, 2);
        HIPCHK(c, launch_probe_norm(c->d_probe, nc, c->st));
      }
      HIPCHK(c, hipMemcpyAsync(c->h_results, c->results, sizeof(double) * RESULT_STRIDE * ns, hipMemcpyDeviceToHost,
                               c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
    }
    int na = 0;
    for (int j = 0, q = 0; q < ns; ++q) {
      const double* r = c->h_results + RESULT_STRIDE * q;
      const bool is_cand = j < nc && cand[j] == q;
      if (is_cand) ++j;
      bool ok = false;
      if (is_cand) {
        // r[2] = ||theta||^2, r[5] = the smallest kept s^2, r[6] / r[7] the two probe sums
        const double total = r[2], eta = kProbeSafety * r[6] / kProbes, missed = kProbeSafety * r[7] / kProbes;
        const int k = (int)r[0];
        const int kmax = std::min(c->cap(), std::min(S0[q].R, S0[q].C));
        // The kept vectors must be singular vectors of theta to first order (eta), and the
        // weight outside the sketch negligible -- or, when the cap binds (k = kmax), where
        // the full SVD discards the weight beyond the cap too, its unscaled estimate no
        // larger than the smallest kept value: a missed direction ten times heavier than
        // that would pass with probability < 2e-11, and one within that factor changes the
        // state by no more than the truncation at the cap already does.
        const double floor_w = c->tol_missed * total;
        ok = eta <= floor_w && (missed <= floor_w || (k == kmax && r[7] / kProbes <= r[5]));
        if (!ok && k == kmax) capped_reject[sk[q]] = 1;
        if (c->rank_log) {
          const double rec[8] = {(double)k, (double)kmax, total, r[1], r[5], r[6] / kProbes, r[7] / kProbes, ok ? 1.0 : 0.0};
          c->log_check.insert(c->log_check.end(), rec, rec + 8);
        }
      }
      if (ok) {
        const int k = (int)r[0];
        const BondTask& b0 = S0[q];
        BondTask& w = S4[na];
        w = S1[q];
        w.qr = 1;
        w.transposed = 0;
        w.Rb = b0.R; w.Cb = b0.C;
        w.Yr = b0.W;            // left vectors of theta (m x k)
        w.Zc = S3[q].Yr;        // right vectors of theta (n x k)
        w.perm = S3[q].perm;
        w.s = S3[q].s;
        w.result = S3[q].result;
        w.gl_out = const_cast<zc*>(w.gl);
        w.gr_out = const_cast<zc*>(w.gr);
        w.lc_out = S3[q].lc_out;
        w.lc_commit = const_cast<double*>(w.lc);
        c->chi[sk_links[q] + 1] = k;
        c->last_rank[sk_links[q] + 1] = k;
        c->last_trunc[sk_links[q] + 1] = r[1];
        tsum += r[1];
        ++na;
      } else {
        fb_links.push_back(sk_links[q]);
        fb_gates.push_back(gate_ptrs[sk[q]]);
        c->stats[2] += 1;
      }
    }
    c->stats[5] += na;
    if (na > 0) {
      HIPCHK(c, hipMemcpyAsync(D4, S4, sizeof(BondTask) * na, hipMemcpyHostToDevice, c->st));
      {
        ProfScope ps(c, 2);
        HIPCHK(c, launch_write_gammas(D4, na, c->st));
      }
      c->stats[0] += na;
    }
    if (nf > 0) {
      ProfScope ps(c, 2);
      HIPCHK(c, launch_write_gammas(D3 + ns, nf, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    drain_profile(c);
  }
  if (!fb_links.empty()) {
    double t = 0.0;
    int e = run_bonds_full(c, fb_links.data(), (int)fb_links.size(), fb_gates.data(), &t);
    if (e) return e;
    tsum += t;
  }
  // the record: (features, log2(k / chi_c before)) of every bond of the batch, with k the
  // kept rank, or 2 kmax when no sketch the model can ask for would pass: the update
  // discarded more than tol_trunc, or its sketch kept the cap and failed the probe check
  if (c->rank_log)
    for (int i = 0; i < nb; ++i) {
      const int link = links[i];
      const float* f = c->h_feat + (size_t)TEBD_MLP_INPUTS * i;
      c->log_X.insert(c->log_X.end(), f, f + TEBD_MLP_INPUTS);
      const int kmax = std::min(c->cap(), d * std::min(c->chi[link], c->chi[link + 2]));
      const bool hard = c->last_trunc[link + 1] > c->tol_trunc || capped_reject[i];
      const double k = hard ? 2.0 * kmax : (double)c->chi[link + 1];
      c->log_y.push_back((float)std::log2(k / (double)chic_before[i]));
    }
  *trunc_sum = tsum;
  return TEBD_OK;
}

// SVD of W (R x C, ld R, R >= C) by one-sided Jacobi.  On return W holds W J
// (columns U s, unsorted), V (Cpad x Cpad) holds J, s[0..C) the singular values
// in descending order and perm[b] the column of W / V that carries s[b].
// SVDs of nb matrices at once by one-sided Jacobi (task i in slot i): W_i (R_i x C_i,
// ld R_i, R_i >= C_i).  On return W_i holds W_i J (columns U s, unsorted), V_i
// (Cpad x Cpad) holds J, s_i[0..C_i) the singular values in descending order and
// perm_i[b] the column of W_i / V_i that carries s_i[b].
struct SvdJob {
  zc* W;
  int R, C;
  zc* V;
  double* s;
  int* perm;
  double* lam_scratch;
};

int device_svd_batch(tebd_ctx* c, const SvdJob* jobs, int nb) {
  int Pmax = 1, Cpmax = 1;
  for (int i = 0; i < nb; ++i) {
    BondTask& t = c->h_tasks[i];
    std::memset(&t, 0, sizeof(t));
    t.W = jobs[i].W; t.R = jobs[i].R; t.C = jobs[i].C; t.Cpad = round_up(t.C, JC); t.P = t.Cpad / JC;
    t.V = jobs[i].V; t.NV = t.Cpad; t.s = jobs[i].s; t.perm = jobs[i].perm;
    attach_jacobi(c, t, i);
    t.result = c->results + (size_t)RESULT_STRIDE * i;
    t.chi_max = t.C;
    t.lc_out = jobs[i].lam_scratch;
    Pmax = std::max(Pmax, t.P);
    Cpmax = std::max(Cpmax, t.Cpad);
  }
  HIPCHK(c, hipMemcpyAsync(c->d_tasks, c->h_tasks, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, launch_prepare_svd(c->d_tasks, nb, c->st));
  HIPCHK(c, launch_frobenius(c->d_tasks, nb, Cpmax, c->st, 0));
  int e = jacobi_sweeps(c, c->d_tasks, c->h_tasks, nb, Pmax);
  if (e) return e;
  HIPCHK(c, launch_col_norms(c->d_tasks, nb, Cpmax, c->st));
  HIPCHK(c, launch_truncate(c->d_tasks, nb, 0.0, c->st));
  return TEBD_OK;
}

// nb GEMMs C_i = op(A_i) B_i as one batched launch; descriptor region `region` (0..4) of
// the pinned descriptors, so the GEMMs of one CanonicalForm batch need no host wait.
int gemm_batch(tebd_ctx* c, const GemmTask* g, int nb, int region, bool conj) {
  GemmTask* h = c->h_gemm + (size_t)region * c->ntask_max;
  GemmTask* d = c->d_gemm + (size_t)region * c->ntask_max;
  int max_m = 1, max_n = 1;
  for (int i = 0; i < nb; ++i) {
    h[i] = g[i];
    max_m = std::max(max_m, g[i].M);
    max_n = std::max(max_n, g[i].N);
  }
  HIPCHK(c, hipMemcpyAsync(d, h, sizeof(GemmTask) * nb, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, tuned_zgemm(c, d, nb, max_m, max_n, conj));
  return TEBD_OK;
}

// CanonicalForm (local_operations_module.f90:496-620) on nb bonds at once, bonds[i]
// 0-based (sites l, l+1); the bonds must not share a site (e.g. the same bond of
// different chains of an ensemble).  Bond i works in CanonicalForm workspace set i.
// The SVDs of the weighted Gram matrices M1 = B1^H B1 and M2 = C2 C2^H are taken as
// one-sided Jacobi SVDs of B1 and C2^H: S = s^2 and the same U, without squaring the
// condition number.  Bond dimensions are kept, as in the reference.  Each bond's
// arithmetic is the same whatever bonds share its batch.
int canonical_form_batch(tebd_ctx* c, const int* bonds, int nb) {
  const int d = c->d;
  const size_t slab = (size_t)std::max(c->Rmax, c->Cpmax) * c->Cpmax;
  const double thr = 1e-8;  // S > 1e-8 (local_operations_module.f90:536,571)
  struct Bond {
    int chi0, chi1, chi2, m1, n2, cp1;
    zc *G1, *G2, *P[8];
    const double *lam0, *lam2;
    double *lam1, *s1, *s2, *s3, *lsc;
    int *p1, *p2, *p3;
  };
  std::vector<Bond> B(nb);
  for (int i = 0; i < nb; ++i) {
    Bond& b = B[i];
    const int l = bonds[i];
    b.chi0 = c->chi[l]; b.chi1 = c->chi[l + 1]; b.chi2 = c->chi[l + 2];
    b.m1 = d * b.chi0; b.n2 = d * b.chi2;
    if (b.m1 < b.chi1 || b.n2 < b.chi1)
      return fail(c, TEBD_ERR_SHAPE, "canonical_form: bond dimension exceeds d x neighbour");
    b.cp1 = round_up(b.chi1, JC);
    for (int j = 0; j < 8; ++j) b.P[j] = c->cf + slab * (8 * (size_t)i + j);
    b.s1 = c->cf_s + (3 * (size_t)i + 0) * c->Cpmax;
    b.s2 = c->cf_s + (3 * (size_t)i + 1) * c->Cpmax;
    b.s3 = c->cf_s + (3 * (size_t)i + 2) * c->Cpmax;
    b.p1 = c->cf_perm + (3 * (size_t)i + 0) * c->Cpmax;
    b.p2 = c->cf_perm + (3 * (size_t)i + 1) * c->Cpmax;
    b.p3 = c->cf_perm + (3 * (size_t)i + 2) * c->Cpmax;
    b.lsc = c->cf_lam + (size_t)i * c->Cpmax;
    b.G1 = c->gam + (size_t)l * c->gslab();
    b.G2 = c->gam + (size_t)(l + 1) * c->gslab();
    b.lam0 = c->lam + (size_t)l * c->chi_max;
    b.lam1 = c->lam + (size_t)(l + 1) * c->chi_max;
    b.lam2 = c->lam + (size_t)(l + 2) * c->chi_max;
  }
  std::vector<SvdJob> J(nb);
  std::vector<GemmTask> G(nb);
  int e;
  // M1 = Gamma1^H lambda0^2 Gamma1  ->  U1, S1
  for (const Bond& b : B)
    HIPCHK(c, launch_scaled_copy(b.G1, b.m1, b.P[0], b.m1, b.m1, b.chi1, b.lam0, b.chi0, nullptr, 1, 0, c->st));
  for (int i = 0; i < nb; ++i) J[i] = {B[i].P[0], B[i].m1, B[i].chi1, B[i].P[1], B[i].s1, B[i].p1, B[i].lsc};
  if ((e = device_svd_batch(c, J.data(), nb))) return e;
  for (const Bond& b : B) {
    HIPCHK(c, launch_perm_cols(b.P[1], b.cp1, b.P[2], b.chi1, b.chi1, b.chi1, b.p1, b.s1, 0, thr, c->st));  // U1
    HIPCHK(c, launch_perm_cols(b.P[1], b.cp1, b.P[3], b.chi1, b.chi1, b.chi1, b.p1, b.s1, 2, thr, c->st));  // U1 S1^-1/2
  }
  for (int i = 0; i < nb; ++i)  // Gamma1Temp = Gamma1 U1 S1^-1/2
    G[i] = {B[i].G1, B[i].P[3], B[i].P[4], B[i].m1, B[i].chi1, B[i].chi1, B[i].m1, B[i].chi1, B[i].m1, 1.0, 0};
  if ((e = gemm_batch(c, G.data(), nb, 0, false))) return e;
  // M2 = Gamma2 lambda2^2 Gamma2^H  ->  U2, S2
  for (const Bond& b : B)
    HIPCHK(c, launch_scaled_copy(b.G2, b.chi1, b.P[0], b.n2, b.chi1, b.n2, nullptr, 1, b.lam2, d, 1, c->st));
  for (int i = 0; i < nb; ++i) J[i] = {B[i].P[0], B[i].n2, B[i].chi1, B[i].P[1], B[i].s2, B[i].p2, B[i].lsc};
  if ((e = device_svd_batch(c, J.data(), nb))) return e;
  for (const Bond& b : B) {
    HIPCHK(c, launch_perm_cols(b.P[1], b.cp1, b.P[5], b.chi1, b.chi1, b.chi1, b.p2, b.s2, 0, thr, c->st));  // U2
    HIPCHK(c, launch_perm_cols(b.P[1], b.cp1, b.P[6], b.chi1, b.chi1, b.chi1, b.p2, b.s2, 2, thr, c->st));  // U2 S2^-1/2
  }
  for (int i = 0; i < nb; ++i)  // Gamma2Temp = (U2 S2^-1/2)^H Gamma2
    G[i] = {B[i].P[6], B[i].G2, B[i].P[7], B[i].chi1, B[i].n2, B[i].chi1, B[i].chi1, B[i].chi1, B[i].chi1, 1.0, 0};
  if ((e = gemm_batch(c, G.data(), nb, 1, true))) return e;
  // X = S1^1/2 U1^H, Y = lambda1 U2 S2^1/2, M = X Y -> U, S, V^T
  for (const Bond& b : B) HIPCHK(c, launch_cf_xy(b.P[2], b.s1, b.P[5], b.s2, b.lam1, b.P[3], b.P[6], b.chi1, c->st));
  for (int i = 0; i < nb; ++i)
    G[i] = {B[i].P[3], B[i].P[6], B[i].P[0], B[i].chi1, B[i].chi1, B[i].chi1, B[i].chi1, B[i].chi1, B[i].chi1, 1.0, 0};
  if ((e = gemm_batch(c, G.data(), nb, 2, false))) return e;
  for (int i = 0; i < nb; ++i) J[i] = {B[i].P[0], B[i].chi1, B[i].chi1, B[i].P[1], B[i].s3, B[i].p3, B[i].lsc};
  if ((e = device_svd_batch(c, J.data(), nb))) return e;
  for (const Bond& b : B) {
    HIPCHK(c, launch_perm_cols(b.P[0], b.chi1, b.P[2], b.chi1, b.chi1, b.chi1, b.p3, b.s3, 1, 0.0, c->st));  // U
    HIPCHK(c, launch_perm_cols(b.P[1], b.cp1, b.P[5], b.chi1, b.chi1, b.chi1, b.p3, b.s3, 0, 0.0, c->st));   // V
    HIPCHK(c, launch_cf_lambda(b.s3, b.chi1, b.lam1, c->chi_max, c->st));
  }
  for (int i = 0; i < nb; ++i)  // Gamma1 = Gamma1Temp U
    G[i] = {B[i].P[4], B[i].P[2], B[i].G1, B[i].m1, B[i].chi1, B[i].chi1, B[i].m1, B[i].chi1, B[i].m1, 1.0, 0};
  if ((e = gemm_batch(c, G.data(), nb, 3, false))) return e;
  for (int i = 0; i < nb; ++i)  // Gamma2 = V^T Gamma2Temp
    G[i] = {B[i].P[5], B[i].P[7], B[i].G2, B[i].chi1, B[i].n2, B[i].chi1, B[i].chi1, B[i].chi1, B[i].chi1, 1.0, 0};
  if ((e = gemm_batch(c, G.data(), nb, 4, true))) return e;
  // the pinned GEMM descriptors are rewritten by the next batch
  HIPCHK(c, hipStreamSynchronize(c->st));
  return TEBD_OK;
}

// CanonicalForm workspace for `sets` bonds at once (8 matrices and 3 (s, perm) pairs
// per bond), kept and grown as needed.
int ensure_cf_workspace(tebd_ctx* c, int sets) {
  if (c->cf && c->cf_sets >= sets) return TEBD_OK;
  if (sets > c->ntask_max) return fail(c, TEBD_ERR_SHAPE, "canonical_form: more bonds than workspace slots");
  (void)hipFree(c->cf);
  (void)hipFree(c->cf_s);
  (void)hipFree(c->cf_perm);
  (void)hipFree(c->cf_lam);
  c->cf = nullptr; c->cf_s = nullptr; c->cf_perm = nullptr; c->cf_lam = nullptr;
  c->cf_sets = 0;
  const size_t slab = (size_t)std::max(c->Rmax, c->Cpmax) * c->Cpmax;
  HIPCHK(c, hipMalloc(&c->cf, sizeof(zc) * slab * 8 * sets));
  HIPCHK(c, hipMalloc(&c->cf_s, sizeof(double) * 3 * c->Cpmax * sets));
  HIPCHK(c, hipMalloc(&c->cf_perm, sizeof(int) * 3 * c->Cpmax * sets));
  HIPCHK(c, hipMalloc(&c->cf_lam, sizeof(double) * c->Cpmax * sets));
  c->cf_sets = sets;
  if (!c->h_gemm) {
    HIPCHK(c, hipMalloc(&c->d_gemm, sizeof(GemmTask) * c->ntask_max * kGemmRegions));
    HIPCHK(c, hipHostMalloc(&c->h_gemm, sizeof(GemmTask) * c->ntask_max * kGemmRegions));
    c->owns_gemm = c->is_lane;
  }
  return TEBD_OK;
}

// ---- number-conserving update (TwoSiteOpNC_c, local_operations_module.f90:1845-1964) -----
// The reference's 10.0**(-10): a default-kind (single precision) literal, promoted when
// compared with double precision values (UpdateLabelLeft, FormLambdaNC, FormGamma1/2NC).
const double kNcCut = (double)std::pow(10.0f, -10.0f);

// Grows a device buffer to hold `need` elements (the stream is drained first, so nothing
// in flight reads the old one).
template <class T>
hipError_t nc_grow(tebd_ctx* c, T** p, size_t* cap, size_t need) {
  if (need <= *cap) return hipSuccess;
  hipError_t e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return e;
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  const size_t n = std::max(need, *cap * 2);
  e = hipMalloc((void**)p, sizeof(T) * n);
  if (e == hipSuccess) *cap = n;
  return e;
}

// Per-bond theta and singular-vector stores for batches of up to ntask_max bonds, the
// per-slot truncation lambdas and the pinned staging.
int ensure_nc_workspace(tebd_ctx* c) {
  if (c->nc_theta) return TEBD_OK;
  const size_t nt = c->ntask_max, sq = (size_t)c->Rmax * c->Rmax;
  HIPCHK(c, hipMalloc(&c->nc_theta, sizeof(zc) * sq * nt));
  HIPCHK(c, hipMalloc(&c->nc_vec, sizeof(zc) * 2 * sq * nt));
  HIPCHK(c, hipMalloc(&c->nc_lscratch, sizeof(double) * c->Cpmax * nt));
  HIPCHK(c, hipHostMalloc(&c->h_nc_s, sizeof(double) * c->Cpmax * nt));
  HIPCHK(c, hipHostMalloc(&c->h_nc_lam, sizeof(double) * c->chi_max * nt));
  c->nc_bonds = (int)nt;
  return TEBD_OK;
}

void free_all(tebd_ctx* c);
extern "C" int create_impl(tebd_ctx** out, int nchain, int Lc, int d, int chi_max, int order, char bc,
                           unsigned mode);

// The small-block SVD workspace: a batched context with d = 2 and chi = side / 2, whose
// kNcSmallSlots slots hold bond matrices of side = min(kNcSmallSide, Rmax) rows and
// columns with side / 2 singular vectors each.  Its state arrays are never used (its
// Gammas are freed at once).
int ensure_nc_small(tebd_ctx* c) {
  if (c->nc_small) return TEBD_OK;
  const int side = std::min(kNcSmallSide, round_up(c->Rmax, 2));
  tebd_ctx* sm = nullptr;
  if (create_impl(&sm, 1, 2 * kNcSmallSlots, 2, side / 2, 2, 'O', TEBD_MODE_BATCHED) != TEBD_OK || !sm)
    return fail(c, TEBD_ERR_HIP, "number-conserving update: cannot allocate the small-block workspace");
  (void)hipFree(sm->gam);
  sm->gam = nullptr;
  const size_t nt = sm->ntask_max;
  if (hipMalloc(&sm->nc_lscratch, sizeof(double) * sm->Cpmax * nt) != hipSuccess ||
      hipHostMalloc(&sm->h_nc_s, sizeof(double) * sm->Cpmax * nt) != hipSuccess) {
    free_all(sm);
    delete sm;
    return fail(c, TEBD_ERR_HIP, "number-conserving update: cannot allocate the small-block workspace");
  }
  c->nc_small = sm;
  return TEBD_OK;
}

// One bond's number blocks (minmaxNLR, SizeOfBlocks, IndexLeft, IndexRight,
// local_operations_module.f90:1201-1443), rows as theta rows a + chi_l i and columns as
// theta columns g + chi_r j, in the reference's traversal order.
struct NcBondBlocks {
  int minNL = 0;
  std::vector<std::vector<int>> rows, cols;
};

int nc_blocks(tebd_ctx* c, int link, NcBondBlocks& B) {
  const int chi = c->chi_max, d = c->d, tot = c->nc_tot, maxf = c->nc_maxf;
  const int chil = c->chi[link], chir = c->chi[link + 2];
  const int* LLl = c->nc_LL.data() + (size_t)link * chi;
  const int* LRr = c->nc_LR.data() + (size_t)(link + 2) * chi;
  int minNL = 1 << 30, maxNL = 0, minNR = 1 << 30, maxNR = 0;
  for (int a = 0; a < chi; ++a) {
    minNL = std::min(minNL, LLl[a]);
    maxNL = std::max(maxNL, LLl[a] == 10000 ? 0 : LLl[a]);
    minNR = std::min(minNR, LRr[a]);
    maxNR = std::max(maxNR, LRr[a] == 10000 ? 0 : LRr[a]);
  }
  maxNL += maxf;
  maxNR += maxf;
  if (maxNL > tot - minNR) maxNL = tot - minNR; else minNR = tot - maxNL;
  if (maxNR > tot - minNL) maxNR = tot - minNL; else minNL = tot - maxNR;
  const int nblk = maxNL - minNL + 1;
  if (nblk < 1) return fail(c, TEBD_ERR_STATE, "number-conserving update: no number block on bond " + std::to_string(link + 1));
  B.minNL = minNL;
  B.rows.assign(nblk, {});
  B.cols.assign(nblk, {});
  for (int a = 0; a < chi; ++a)
    for (int i = 0; i < d; ++i) {
      const int lab = LLl[a], nl = lab + c->nc_conserv[i];
      if (lab < 1000 && minNL <= nl && nl <= maxNL) {
        if (a >= chil) return fail(c, TEBD_ERR_STATE, "labels: a label past the bond dimension on link " + std::to_string(link + 1));
        B.rows[nl - minNL].push_back(a + chil * i);
      }
    }
  for (int g = 0; g < chi; ++g)
    for (int j = 0; j < d; ++j) {
      const int lab = LRr[g], nr = lab + c->nc_conserv[j];
      if (lab < 1000 && minNR <= nr && nr <= maxNR) {
        if (g >= chir) return fail(c, TEBD_ERR_STATE, "labels: a label past the bond dimension on link " + std::to_string(link + 3));
        B.cols[tot - nr - minNL].push_back(g + chir * j);
      }
    }
  return TEBD_OK;
}

// The number-conserving update of nb independent bonds (0-based links): theta of each bond
// by the ordinary theta kernel, its number blocks gathered and factorised as tasks of the
// batched SVD (in chunks of ntask_max work slots), the singular values of every block of a
// bond ordered together (Ordering, JudgePosition), the labels, lambda and truncation error
// as UpdateLabelLeft / UpdateLabelRight / FormLambdaNC set them, and the Gammas written from
// the kept states' block vectors.  renorm_theta = false skips the normalisation of theta
// before the cut, as SwappingNC does (local_operations_module.f90:1966-2082: no
// RenormThetaNC); the swaps of the periodic schedules run that way.
int run_bonds_nc(tebd_ctx* c, const int* links, int nb, const zc* const* gate_ptrs, double* trunc_sum,
                 bool renorm_theta) {
  const int d = c->d, chi = c->chi_max, nt = c->ntask_max;
  if (nb > nt) return fail(c, TEBD_ERR_ARG, "number-conserving update: batch larger than the workspace");
  int e = ensure_nc_workspace(c);
  if (e) return e;
  const size_t sq = (size_t)c->Rmax * c->Rmax;
  // 1. theta (m x n, ld m) of every bond into the NC theta store
  BondTask* TH = c->h_tasks + 4 * nt;
  BondTask* DTH = c->d_tasks + 4 * nt;
  int max_chil = 1, max_chir = 1;
  for (int i = 0; i < nb; ++i) {
    BondTask& t = TH[i];
    make_full_task(c, t, links[i], gate_ptrs[i], 0);
    t.transposed = 0;
    t.R = d * t.chil;
    t.C = d * t.chir;
    t.W = c->nc_theta + (size_t)i * sq;
    max_chil = std::max(max_chil, t.chil);
    max_chir = std::max(max_chir, t.chir);
  }
  HIPCHK(c, hipMemcpyAsync(DTH, TH, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
  {
    ProfScope ps(c, 0);
    HIPCHK(c, theta_gate(c, DTH, nb, max_chil, max_chir));
  }
  account_theta(c, TH, nb);
  // 2. the blocks of every bond and their index lists
  std::vector<NcBondBlocks> bb(nb);
  for (int i = 0; i < nb; ++i)
    if ((e = nc_blocks(c, links[i], bb[i]))) return e;
  std::vector<int> idx;
  std::vector<NcBlock> blocks;       // blocks with rows and columns
  std::vector<int> blk_bond, blk_k;  // their bond and block number
  std::vector<std::vector<int>> blk_of(nb);  // per bond and block number: index in `blocks` or -1
  std::vector<size_t> row_off, col_off;
  for (int i = 0; i < nb; ++i) {
    const int m = d * c->chi[links[i]];
    size_t voff = (size_t)i * 2 * sq;
    blk_of[i].assign(bb[i].rows.size(), -1);
    for (size_t k = 0; k < bb[i].rows.size(); ++k) {
      const int lgt = (int)bb[i].rows[k].size(), trv = (int)bb[i].cols[k].size();
      if (lgt == 0 || trv == 0) continue;
      NcBlock b{};
      b.theta = c->nc_theta + (size_t)i * sq;
      b.m = m;
      row_off.push_back(idx.size());
      idx.insert(idx.end(), bb[i].rows[k].begin(), bb[i].rows[k].end());
      col_off.push_back(idx.size());
      idx.insert(idx.end(), bb[i].cols[k].begin(), bb[i].cols[k].end());
      b.lgt = lgt;
      b.trv = trv;
      b.transposed = lgt < trv ? 1 : 0;
      b.R = std::max(lgt, trv);
      const int mins = std::min(lgt, trv);
      b.U = c->nc_vec + voff;
      voff += (size_t)lgt * mins;
      b.V = c->nc_vec + voff;
      voff += (size_t)trv * mins;
      blk_of[i][k] = (int)blocks.size();
      blocks.push_back(b);
      blk_bond.push_back(i);
      blk_k.push_back((int)k);
    }
  }
  const int nbk = (int)blocks.size();
  HIPCHK(c, nc_grow(c, &c->d_nc_idx, &c->nc_idx_cap, std::max<size_t>(idx.size(), 1)));
  HIPCHK(c, nc_grow(c, &c->d_nc_blk, &c->nc_blk_cap, std::max(nbk, 1)));
  const int split = nc_gather_split();
  if ((size_t)nbk * split > c->nc_part_cap) {
    HIPCHK(c, nc_grow(c, &c->d_nc_part, &c->nc_part_cap, (size_t)nbk * split));
    (void)hipHostFree(c->h_nc_part);
    c->h_nc_part = nullptr;
    HIPCHK(c, hipHostMalloc(&c->h_nc_part, sizeof(double) * c->nc_part_cap));
  }
  for (int q = 0; q < nbk; ++q) {
    blocks[q].rows = c->d_nc_idx + row_off[q];
    blocks[q].cols = c->d_nc_idx + col_off[q];
  }
  HIPCHK(c, hipMemcpyAsync(c->d_nc_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c->st));
  // 3. the weight of every block (blocks without weight skip the SVD; a non-finite one is
  // reported)
  std::vector<double> wgt(nbk, 0.0);
  if (nbk > 0) {
    HIPCHK(c, hipMemcpyAsync(c->d_nc_blk, blocks.data(), sizeof(NcBlock) * nbk, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_nc_gather(c->d_nc_blk, nbk, c->d_nc_part, c->st));
    HIPCHK(c, hipMemcpyAsync(c->h_nc_part, c->d_nc_part, sizeof(double) * nbk * split, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int q = 0; q < nbk; ++q) {
      for (int x = 0; x < split; ++x) wgt[q] += c->h_nc_part[(size_t)q * split + x];
      if (!std::isfinite(wgt[q]))
        return fail(c, TEBD_ERR_NAN, "bond " + std::to_string(links[blk_bond[q]] + 1) + ": theta is not finite");
    }
  }
  // 4. SVDs of the blocks with weight; singular values to the host.  In the batched mode the
  // blocks that fit the small-block workspace run there, kNcSmallSlots at a time, the
  // others ntask_max at a time in the context's slots.  A block's result does not depend on
  // the slots or the batch it runs in.
  std::vector<std::vector<double>> sv(nbk);  // per block: min(lgt, trv) values, descending
  for (int q = 0; q < nbk; ++q) sv[q].assign(std::min(blocks[q].lgt, blocks[q].trv), 0.0);
  tebd_ctx* sm = nullptr;
  if (c->mode & TEBD_MODE_BATCHED) {
    if ((e = ensure_nc_small(c))) return e;
    sm = c->nc_small;
  }
  std::vector<int> live_ctx, live_small;
  for (int q = 0; q < nbk; ++q) {
    if (!(wgt[q] > 0.0)) continue;
    const NcBlock& b = blocks[q];
    const int mins = std::min(b.lgt, b.trv);
    const bool fits = sm && b.R <= sm->Rmax && round_up(mins, JC) <= sm->Cpmax &&
                      std::min(mins, c->chi_max) <= sm->chi_max;
    (fits ? live_small : live_ctx).push_back(q);
  }
  // largest first, so that each batch's sweeps run about as many rounds as its blocks need
  std::stable_sort(live_small.begin(), live_small.end(), [&](int x, int y) {
    const int cx = std::min(blocks[x].lgt, blocks[x].trv), cy = std::min(blocks[y].lgt, blocks[y].trv);
    return cx != cy ? cx > cy : blocks[x].R > blocks[y].R;
  });
  std::vector<NcBlock> chunk_blk;
  std::vector<int> err_links;
  // the blocks `live` in the slots of w (c or its small-block workspace), on w's stream
  auto factor_blocks = [&](tebd_ctx* w, const std::vector<int>& live) -> int {
    const int wn = w->ntask_max;
    BondTask* S = w->h_tasks;
    BondTask* DS = w->d_tasks;
    if (w != c) {
      w->cutoff = c->cutoff;
      w->graphs = c->graphs;
      w->profiling = c->profiling;
    }
    for (size_t q0 = 0; q0 < live.size(); q0 += wn) {
      const int n = (int)std::min(live.size() - q0, (size_t)wn);
      chunk_blk.assign(n, NcBlock{});
      err_links.assign(n, 0);
      int Cpm = 1, Pm = 1;
      for (int q = 0; q < n; ++q) {
        const NcBlock& b = blocks[live[q0 + q]];
        BondTask& t = S[q];
        std::memset(&t, 0, sizeof(t));
        t.d = d;
        t.link = links[blk_bond[live[q0 + q]]];
        t.R = b.R;
        t.C = std::min(b.lgt, b.trv);
        t.Cpad = round_up(t.C, JC);
        t.P = t.Cpad / JC;
        t.W = w->W + (size_t)q * w->Rmax * w->Cpmax;
        t.V = w->V + (size_t)q * w->Cpmax * w->Cpmax;
        t.NV = t.Cpad;
        t.s = w->s + (size_t)q * w->Cpmax;
        t.perm = w->perm + (size_t)q * w->Cpmax;
        attach_jacobi(w, t, q);
        attach_qr(w, t, q);
        t.result = w->results + (size_t)RESULT_STRIDE * q;
        t.lc_out = w->nc_lscratch + (size_t)q * w->Cpmax;
        // at most chi_max values of a block can be among the chi_max kept ones (and the
        // vector panels hold chi_max columns)
        t.chi_max = std::min(t.C, c->chi_max);
        chunk_blk[q] = b;
        chunk_blk[q].W = t.W;
        err_links[q] = t.link;
        Cpm = std::max(Cpm, t.Cpad);
        Pm = std::max(Pm, t.P);
      }
      HIPCHK(c, hipMemcpyAsync(c->d_nc_blk, chunk_blk.data(), sizeof(NcBlock) * n, hipMemcpyHostToDevice, w->st));
      HIPCHK(c, hipMemcpyAsync(DS, S, sizeof(BondTask) * n, hipMemcpyHostToDevice, w->st));
      HIPCHK(c, launch_nc_gather(c->d_nc_blk, n, c->d_nc_part, w->st));
      bool used_qr = false;
      int max_k = 1;
      int ef = svd_factor(w, DS, S, err_links.data(), n, Cpm, Pm, false, &used_qr, &max_k);
      if (!ef) {
        HIPCHK(c, launch_nc_vectors(c->d_nc_blk, DS, n, max_k, w->st));
        HIPCHK(c, hipMemcpyAsync(w->h_nc_s, w->s, sizeof(double) * w->Cpmax * n, hipMemcpyDeviceToHost, w->st));
      }
      HIPCHK(c, hipStreamSynchronize(w->st));
      if (w != c) {  // the workspace's counters and timings belong to the context
        drain_profile(w);
        for (int k = 0; k < kNumStats; ++k) c->stats[k] += w->stats[k], w->stats[k] = 0.0;
        for (int k = 0; k < kNumProf; ++k) c->prof[k] += w->prof[k], w->prof[k] = 0.0;
        if (ef) return fail(c, ef, w->err);
      }
      if (ef) return ef;
      for (int q = 0; q < n; ++q) {
        const int k = (int)w->h_results[RESULT_STRIDE * q];
        std::vector<double>& v = sv[live[q0 + q]];
        for (int a = 0; a < k && a < (int)v.size(); ++a) v[a] = w->h_nc_s[(size_t)q * w->Cpmax + a];
      }
    }
    return TEBD_OK;
  };
  if ((e = factor_blocks(c, live_ctx))) return e;
  if (sm && (e = factor_blocks(sm, live_small))) return e;
  // 5. per bond: all blocks' values in block order (FlattenSS), normalised as RenormThetaNC
  // normalises theta, ordered (Ordering, JudgePosition), kept above the cut (UpdateLabelLeft,
  // FormLambdaNC), and the kept states' Gammas
  std::vector<NcKept> kept;
  double tsum = 0.0;
  for (int i = 0; i < nb; ++i) {
    const int link = links[i];
    std::vector<double> ssfl;
    std::vector<int> fl_blk, fl_pos;  // block number and position of each flat value
    double wsum = 0.0;
    for (size_t k = 0; k < bb[i].rows.size(); ++k) {
      const int q = blk_of[i][k];
      if (q < 0) continue;  // min(lgt, trv) = 0: no values
      for (size_t a = 0; a < sv[q].size(); ++a) {
        ssfl.push_back(sv[q][a]);
        fl_blk.push_back((int)k);
        fl_pos.push_back((int)a);
        wsum += sv[q][a] * sv[q][a];
      }
    }
    if (!(wsum > 0.0)) return fail(c, TEBD_ERR_NAN, "bond " + std::to_string(link + 1) + ": theta is zero in every number block");
    const double inv = renorm_theta ? 1.0 / std::sqrt(wsum) : 1.0;
    for (double& x : ssfl) x *= inv;
    // Ordering: backwards, the first occurrence of the smallest remaining value each time
    const int sz = (int)ssfl.size();
    std::vector<int> order(sz);
    {
      std::vector<int> ix(sz);
      for (int a = 0; a < sz; ++a) ix[a] = a;
      // descending by value, equal values with the higher flat index first
      std::stable_sort(ix.begin(), ix.end(), [&](int x, int y) {
        if (ssfl[x] != ssfl[y]) return ssfl[x] > ssfl[y];
        return x > y;
      });
      order = ix;
    }
    int knew = 0;
    while (knew < std::min(c->cap(), sz) && ssfl[order[knew]] > kNcCut) ++knew;
    if (knew < 1) return fail(c, TEBD_ERR_NAN, "bond " + std::to_string(link + 1) + ": no singular value above the cut");
    double renorm = 0.0;
    for (int b = 0; b < knew; ++b) renorm += ssfl[order[b]] * ssfl[order[b]];
    const double trunc = std::fabs(1.0 - renorm);
    tsum += trunc;
    int* LLc = c->nc_LL.data() + (size_t)(link + 1) * chi;
    int* LRc = c->nc_LR.data() + (size_t)(link + 1) * chi;
    double* lh = c->h_nc_lam + (size_t)i * chi;
    for (int b = 0; b < chi; ++b) {
      if (b < knew) {
        const int lab = bb[i].minNL + fl_blk[order[b]];
        LLc[b] = lab;
        LRc[b] = c->nc_tot - lab;
        lh[b] = ssfl[order[b]] / std::sqrt(renorm);
      } else {
        LLc[b] = 10000;
        LRc[b] = 10000;
        lh[b] = 0.0;
      }
    }
    const int chil = c->chi[link], chir = c->chi[link + 2];
    for (int b = 0; b < knew; ++b) {
      const int q = blk_of[i][fl_blk[order[b]]];
      NcKept o{};
      o.rows = blocks[q].rows;
      o.cols = blocks[q].cols;
      o.lgt = blocks[q].lgt;
      o.trv = blocks[q].trv;
      o.U = blocks[q].U;
      o.V = blocks[q].V;
      o.pos = fl_pos[order[b]];
      o.gl = c->gam + (size_t)link * c->gslab();
      o.gr = c->gam + (size_t)(link + 1) * c->gslab();
      o.ll = c->lam + (size_t)link * c->chi_max;
      o.lr = c->lam + (size_t)(link + 2) * c->chi_max;
      o.chil = chil;
      o.chir = chir;
      o.d = d;
      o.knew = knew;
      o.b = b;
      kept.push_back(o);
    }
    c->chi[link + 1] = knew;
    c->last_rank[link + 1] = knew;
    c->last_trunc[link + 1] = trunc;
  }
  HIPCHK(c, nc_grow(c, &c->d_nc_kept, &c->nc_kept_cap, kept.size()));
  HIPCHK(c, hipMemcpyAsync(c->d_nc_kept, kept.data(), sizeof(NcKept) * kept.size(), hipMemcpyHostToDevice, c->st));
  {
    ProfScope ps(c, 2);
    HIPCHK(c, launch_nc_gammas(c->d_nc_kept, (int)kept.size(), kNcCut, c->st));
  }
  for (int i = 0; i < nb; ++i)
    HIPCHK(c, hipMemcpyAsync(c->lam + (size_t)(links[i] + 1) * c->chi_max, c->h_nc_lam + (size_t)i * chi,
                             sizeof(double) * chi, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  drain_profile(c);
  c->stats[0] += nb;
  *trunc_sum = tsum;
  return TEBD_OK;
}

int run_bonds(tebd_ctx* c, const int* links, int nb, const zc* const* gate_ptrs, double* trunc_sum) {
  if (c->nc) return run_bonds_nc(c, links, nb, gate_ptrs, trunc_sum, true);
  if (c->mode & TEBD_MODE_LEARNED_RANK) return run_bonds_learned(c, links, nb, gate_ptrs, trunc_sum);
  return run_bonds_full(c, links, nb, gate_ptrs, trunc_sum);
}

void free_lane(tebd_ctx* l);

// Lane `lane` of context c: workspace slots [slot0, slot0 + nslots), its own Jacobi
// queue counters and events, and a stream: lane 0 uses the context's own stream, which
// is idle while the lanes run, the others one each.  HIP maps streams onto a few
// hardware queues per process (four by default), and two busy streams that share a
// queue run one after the other, so the lanes take as few streams as they can.
// Everything else is the context's.
tebd_ctx* make_lane(tebd_ctx* c, int lane, int slot0, int nslots) {
  tebd_ctx* l = new tebd_ctx(*c);
  l->is_lane = true;
  for (tebd_ctx*& x : l->lanes) x = nullptr;
  l->lanes_made = 0;
  l->event_pool.clear();
  l->pending.clear();
  l->staging.clear();
  l->staging.shrink_to_fit();
  l->tev0 = l->tev1 = l->ev0 = l->ev1 = nullptr;
  l->sweep_graphs.clear();  // graphs are captured per stream
  l->cf = nullptr;           // CanonicalForm workspace: a lane allocates its own
  l->cf_s = nullptr;
  l->cf_lam = nullptr;
  l->cf_perm = nullptr;
  l->cf_sets = 0;
  l->cr_theta[0] = l->cr_theta[1] = l->cr_T = l->cr_K = l->cr_B = l->cr_A = l->cr_rho2 = l->cr_rho1 = nullptr;
  l->cr_s = l->cr_phase = nullptr;
  l->nc_theta = l->nc_vec = nullptr;   // the number-conserving update runs on the context only
  l->nc_small = nullptr;
  l->d_nc_idx = nullptr;
  l->d_nc_blk = nullptr;
  l->d_nc_kept = nullptr;
  l->d_nc_part = l->nc_lscratch = l->h_nc_s = l->h_nc_lam = l->h_nc_part = nullptr;
  l->owns_gemm = false;
  l->owns_stream = lane != 0;
  l->st = lane == 0 ? c->st : nullptr;
  l->wl_count = nullptr;
  if ((lane != 0 && hipStreamCreateWithFlags(&l->st, hipStreamNonBlocking) != hipSuccess) ||
      hipMalloc(&l->wl_count, sizeof(int) * 2) != hipSuccess) {
    free_lane(l);
    return nullptr;
  }
  const size_t s0 = (size_t)slot0;
  l->ntask_max = nslots;
  l->W = c->W + s0 * c->Rmax * c->Cpmax;
  l->V = c->V + s0 * c->Cpmax * c->Cpmax;
  l->s = c->s + s0 * c->Cpmax;
  l->perm = c->perm + s0 * c->Cpmax;
  l->X = c->X + s0 * c->Cpmax * c->Cpmax;
  l->Yr = c->Yr + s0 * c->Rmax * c->chi_max;
  l->Zc = c->Zc + s0 * c->Cpmax * c->chi_max;
  l->qnrm = c->qnrm + s0 * c->Cpmax;
  l->qpiv = c->qpiv + s0 * c->Cpmax;
  l->tau = c->tau + s0 * c->Cpmax;
  l->qstate = c->qstate + 2 * s0;
  l->h_qstate = c->h_qstate + 2 * s0;
  l->Tq = c->Tq + s0 * tq_stride(c);
  l->counters = c->counters + 4 * s0;
  l->h_counters = c->h_counters + 4 * s0;
  l->blkver = c->blkver + s0 * c->nblk_max;
  l->pairseen = c->pairseen + s0 * c->nblk_max * c->nblk_max;
  l->Qg = c->Qg + s0 * (c->nblk_max / 2) * JC * JC;
  l->wl_cap = nslots * (c->nblk_max / 2);
  l->wl_items = c->wl_items + 2 * s0 * (c->nblk_max / 2);
  l->results = c->results + RESULT_STRIDE * s0;
  l->results2 = c->results2 + RESULT_STRIDE * s0;
  l->h_results = c->h_results + RESULT_STRIDE * s0;
  l->d_feat = c->d_feat + TEBD_MLP_INPUTS * s0;
  l->h_feat = c->h_feat + TEBD_MLP_INPUTS * s0;
  l->d_pred = c->d_pred + s0;
  l->h_pred = c->h_pred + s0;
  l->d_tasks = c->d_tasks + kStages * s0;
  l->h_tasks = c->h_tasks + kStages * s0;
  if (c->split.a) {
    l->split.a = c->split.a + s0 * c->split.stride;
    l->split.b = c->split.b + s0 * c->split.stride;
    l->split.ea = c->split.ea + s0 * c->split.rows;
    l->split.eb = c->split.eb + s0 * c->split.rows;
    l->split.eg = c->split.eg + s0 * c->split.pitch;
  }
  if (c->Zb) {
    l->Zb = c->Zb + s0 * c->Rmax * c->Cpmax;
    l->Bb = c->Bb + s0 * c->Rmax * c->Cpmax;
    l->lam_scratch = c->lam_scratch + s0 * c->chi_max;
    l->Pb = c->Pb + s0 * probe_stride(c);
    l->d_probe = c->d_probe + s0;
    l->h_probe = c->h_probe + s0;
    l->d_kprime = c->d_kprime + s0;
    l->h_kprime = c->h_kprime + s0;
  }
  // GEMM descriptors: the lane's slice of the context's, or (none yet) its own later
  l->d_gemm = c->d_gemm ? c->d_gemm + kGemmRegions * s0 : nullptr;
  l->h_gemm = c->h_gemm ? c->h_gemm + kGemmRegions * s0 : nullptr;
  return l;
}

// The settings and host mirrors a lane reads, from its context, before a run.
void lane_pull(tebd_ctx* l, const tebd_ctx* c) {
  l->chi = c->chi;
  l->last_rank = c->last_rank;
  l->last_trunc = c->last_trunc;
  l->tuned = c->tuned;
  l->mode = c->mode;
  l->cutoff = c->cutoff;
  l->stage_chi = c->stage_chi;
  l->have_model = c->have_model;
  l->rank_log = c->rank_log;
  l->tol_missed = c->tol_missed;
  l->tol_trunc = c->tol_trunc;
  l->rank_slack = c->rank_slack;
  l->oversample = c->oversample;
  l->step_index = c->step_index;
  l->profiling = c->profiling;
  l->autotune = c->autotune;
  l->graphs = c->graphs;
  l->have_gates = c->have_gates;
  l->have_state = c->have_state;
  for (double& v : l->stats) v = 0.0;
  for (double& v : l->prof) v = 0.0;
  l->tune_trials = 0.0;
  l->log_X.clear();
  l->log_y.clear();
  l->log_check.clear();
  l->err.clear();
}

// What a lane's run changed, back into its context: the bonds it updated, counters,
// timings, autotuner choices and the rank-model record.
void lane_push(tebd_ctx* c, const tebd_ctx* l, const int* links, int nb) {
  for (int i = 0; i < nb; ++i) {
    c->chi[links[i] + 1] = l->chi[links[i] + 1];
    c->last_rank[links[i] + 1] = l->last_rank[links[i] + 1];
    c->last_trunc[links[i] + 1] = l->last_trunc[links[i] + 1];
  }
  for (int k = 0; k < kNumStats; ++k) c->stats[k] += l->stats[k];
  for (int k = 0; k < kNumProf; ++k) c->prof[k] += l->prof[k];
  for (const auto& kv : l->tuned) c->tuned.insert(kv);
  c->tune_trials += l->tune_trials;
  c->log_X.insert(c->log_X.end(), l->log_X.begin(), l->log_X.end());
  c->log_y.insert(c->log_y.end(), l->log_y.begin(), l->log_y.end());
  c->log_check.insert(c->log_check.end(), l->log_check.begin(), l->log_check.end());
}

// n lanes of ntask_max / n workspace slots each (made again when n changes); false if
// they cannot be made.
bool ensure_lanes(tebd_ctx* c, int n) {
  if (c->lanes_made == n) return true;
  for (tebd_ctx*& l : c->lanes) {
    free_lane(l);
    l = nullptr;
  }
  c->lanes_made = 0;
  const int per = c->ntask_max / n;
  for (int k = 0; k < n; ++k) {
    c->lanes[k] = make_lane(c, k, k * per, k + 1 < n ? per : c->ntask_max - k * per);
    if (!c->lanes[k]) return false;
  }
  c->lanes_made = n;
  return true;
}

// The bonds of one half-sweep: with n lanes, interleaved parts of the batch run on n
// streams, part 0 from this thread and the others from their own threads, so that one
// part's latency-bound phases (QR panels, inner Jacobi solves) overlap the others'
// work.  Every bond's result is independent of which bonds share its batch, so the
// split changes the timing, never the numbers.
int run_bonds_split(tebd_ctx* c, const int* links, int nb, const zc* const* gate_ptrs, double* trunc_sum) {
  // small bond matrices leave nothing to overlap: the extra launches and threads only
  // cost (measured: 64-site chains at chi = 64, d = 2 run slower with two lanes)
  int side = 0;
  for (int i = 0; i < nb; ++i) side = std::max(side, c->d * std::max(c->chi[links[i]], c->chi[links[i] + 2]));
  const int n = std::min(c->nlanes, nb / 4);
  if (n < 2 || c->ntask_max < 2 * kMaxLanes || side < kLaneMinSide) return run_bonds(c, links, nb, gate_ptrs, trunc_sum);
  if (!ensure_lanes(c, n)) return run_bonds(c, links, nb, gate_ptrs, trunc_sum);
  // bond i of the batch goes to lane n - 1 - (i mod n): every lane gets every n-th bond,
  // so the chain's cheap edge bonds are spread over the lanes instead of making the lanes
  // that hold them finish early; the last lane, which has the spare workspace slots, takes
  // the largest share
  std::vector<int> lk[kMaxLanes];
  std::vector<const zc*> gk[kMaxLanes];
  for (int i = 0; i < nb; ++i) {
    const int k = n - 1 - i % n;
    lk[k].push_back(links[i]);
    gk[k].push_back(gate_ptrs[i]);
  }
  for (int k = 0; k < n; ++k)
    if (lk[k].empty() || (int)lk[k].size() > c->lanes[k]->ntask_max) return run_bonds(c, links, nb, gate_ptrs, trunc_sum);
  // the lanes start after everything already queued on the context's stream
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int k = 0; k < n; ++k) lane_pull(c->lanes[k], c);
  double tr[kMaxLanes] = {};
  int err[kMaxLanes] = {};
  std::vector<std::thread> workers;
  for (int k = 1; k < n; ++k)
    workers.emplace_back([&, k] {
      err[k] = run_bonds(c->lanes[k], lk[k].data(), (int)lk[k].size(), gk[k].data(), &tr[k]);
    });
  err[0] = run_bonds(c->lanes[0], lk[0].data(), (int)lk[0].size(), gk[0].data(), &tr[0]);
  for (auto& w : workers) w.join();
  for (int k = 0; k < n; ++k)
    if (err[k]) return fail(c, err[k], c->lanes[k]->err);
  double tsum = 0.0;
  for (int k = 0; k < n; ++k) {
    lane_push(c, c->lanes[k], lk[k].data(), (int)lk[k].size());
    tsum += tr[k];
  }
  c->stats[13] += 1;
  *trunc_sum = tsum;
  return TEBD_OK;
}

int ensure_state(tebd_ctx* c, bool need_gates) {
  if (!c) return TEBD_ERR_ARG;
  if (!c->have_state) return fail(c, TEBD_ERR_STATE, "no state uploaded");
  if (need_gates && !c->have_gates) return fail(c, TEBD_ERR_STATE, "no gates set");
  return TEBD_OK;
}

int upload_impl(tebd_ctx* c, const double* const* gamma, const double* const* lambda, int chi_in) {
  if (!c || !gamma || !lambda || chi_in < 1) return fail(c, TEBD_ERR_ARG, "upload: bad arguments");
  const int L = c->L, d = c->d;
  std::vector<int> chi(L + 1);
  for (int l = 0; l <= L; ++l) {
    int k = 0;
    for (int a = 0; a < chi_in; ++a)
      if (lambda[l][a] != 0.0) k = a + 1;
    if (k < 1) k = 1;
    if (k > c->chi_max) return fail(c, TEBD_ERR_SHAPE, "upload: a bond dimension exceeds chi_max");
    chi[l] = k;
  }
  for (int r = 1; r < c->nchain; ++r)
    if (chi[r * c->chain_len] != 1)
      return fail(c, TEBD_ERR_SHAPE, "upload: the bond between two chains of an ensemble must have dimension 1");
  std::vector<double> lam((size_t)(L + 1) * c->chi_max, 0.0);
  for (int l = 0; l <= L; ++l)
    for (int a = 0; a < chi[l]; ++a) lam[(size_t)l * c->chi_max + a] = lambda[l][a];
  HIPCHK(c, hipMemcpy(c->lam, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice));
  c->staging.resize(c->gslab());
  for (int s = 0; s < L; ++s) {
    const zc* src = reinterpret_cast<const zc*>(gamma[s]);
    const int cl = chi[s], cr = chi[s + 1];
    for (int b = 0; b < cr; ++b)
      for (int i = 0; i < d; ++i)
        for (int a = 0; a < cl; ++a)
          c->staging[a + (size_t)cl * (i + (size_t)d * b)] = src[a + (size_t)chi_in * (i + (size_t)d * b)];
    HIPCHK(c, hipMemcpy(c->gam + (size_t)s * c->gslab(), c->staging.data(), sizeof(zc) * (size_t)cl * d * cr,
                        hipMemcpyHostToDevice));
  }
  c->chi = chi;
  c->have_state = true;
  return TEBD_OK;
}

int download_impl(tebd_ctx* c, double* const* gamma, double* const* lambda, int chi_out) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (!gamma || !lambda || chi_out < 1) return fail(c, TEBD_ERR_ARG, "download: bad arguments");
  const int L = c->L, d = c->d;
  for (int l = 0; l <= L; ++l)
    if (c->chi[l] > chi_out) return fail(c, TEBD_ERR_SHAPE, "download: chi_out smaller than a bond dimension");
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::vector<double> lam((size_t)(L + 1) * c->chi_max);
  HIPCHK(c, hipMemcpy(lam.data(), c->lam, lam.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int l = 0; l <= L; ++l)
    for (int a = 0; a < chi_out; ++a) lambda[l][a] = a < c->chi[l] ? lam[(size_t)l * c->chi_max + a] : 0.0;
  c->staging.resize(c->gslab());
  for (int s = 0; s < L; ++s) {
    const int cl = c->chi[s], cr = c->chi[s + 1];
    HIPCHK(c, hipMemcpy(c->staging.data(), c->gam + (size_t)s * c->gslab(), sizeof(zc) * (size_t)cl * d * cr,
                        hipMemcpyDeviceToHost));
    zc* dst = reinterpret_cast<zc*>(gamma[s]);
    std::memset(dst, 0, sizeof(zc) * (size_t)chi_out * d * chi_out);
    for (int b = 0; b < cr; ++b)
      for (int i = 0; i < d; ++i)
        for (int a = 0; a < cl; ++a)
          dst[a + (size_t)chi_out * (i + (size_t)d * b)] = c->staging[a + (size_t)cl * (i + (size_t)d * b)];
  }
  return TEBD_OK;
}

void free_lane(tebd_ctx* l) {
  if (!l) return;
  if (l->st) (void)hipStreamSynchronize(l->st);
  for (auto& p : l->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : l->event_pool) (void)hipEventDestroy(e);
  if (l->tev0) (void)hipEventDestroy(l->tev0);
  if (l->tev1) (void)hipEventDestroy(l->tev1);
  clear_graphs(l);
  (void)hipFree(l->cf);
  (void)hipFree(l->cf_s);
  (void)hipFree(l->cf_perm);
  (void)hipFree(l->cf_lam);
  if (l->owns_gemm) {
    (void)hipFree(l->d_gemm);
    (void)hipHostFree(l->h_gemm);
  }
  (void)hipFree(l->wl_count);
  if (l->st && l->owns_stream) (void)hipStreamDestroy(l->st);
  delete l;
}

void free_all(tebd_ctx* c) {
  for (tebd_ctx*& l : c->lanes) {
    free_lane(l);
    l = nullptr;
  }
  if (tebd_ctx* sm = c->nc_small) {
    c->nc_small = nullptr;
    if (sm->st) (void)hipStreamSynchronize(sm->st);
    free_all(sm);
    delete sm;
  }
  c->lanes_made = 0;
  clear_graphs(c);
  if (c->tev0) (void)hipEventDestroy(c->tev0);
  if (c->tev1) (void)hipEventDestroy(c->tev1);
  (void)hipFree(c->gam);
  (void)hipFree(c->lam);
  (void)hipFree(c->gates);
  (void)hipFree(c->W);
  (void)hipFree(c->V);
  (void)hipFree(c->s);
  (void)hipFree(c->perm);
  (void)hipFree(c->counters);
  (void)hipFree(c->blkver);
  (void)hipFree(c->X);
  (void)hipFree(c->Yr);
  (void)hipFree(c->Zc);
  (void)hipFree(c->qnrm);
  (void)hipFree(c->qpiv);
  (void)hipFree(c->tau);
  (void)hipFree(c->qstate);
  (void)hipFree(c->Tq);
  (void)hipHostFree(c->h_qstate);
  (void)hipFree(c->pairseen);
  (void)hipFree(c->Qg);
  (void)hipFree(c->wl_count);
  (void)hipFree(c->wl_items);
  (void)hipFree(c->results);
  (void)hipFree(c->results2);
  (void)hipFree(c->obs_ops);
  for (zc* p : {c->cr_theta[0], c->cr_theta[1], c->cr_T, c->cr_K, c->cr_B, c->cr_A, c->cr_rho2, c->cr_rho1})
    (void)hipFree(p);
  (void)hipFree(c->cr_s);
  (void)hipFree(c->cr_phase);
  (void)hipFree(c->mlp);
  (void)hipFree(c->d_feat);
  (void)hipFree(c->d_pred);
  (void)hipHostFree(c->h_feat);
  (void)hipHostFree(c->h_pred);
  (void)hipFree(c->ident);
  (void)hipFree(c->d_tasks);
  (void)hipFree(c->cf);
  (void)hipFree(c->cf_s);
  (void)hipFree(c->cf_perm);
  (void)hipFree(c->cf_lam);
  (void)hipFree(c->split.a);
  (void)hipFree(c->split.b);
  (void)hipFree(c->split.ea);
  (void)hipFree(c->split.eb);
  (void)hipFree(c->split.eg);
  (void)hipFree(c->Zb);
  (void)hipFree(c->Bb);
  (void)hipFree(c->lam_scratch);
  (void)hipFree(c->Pb);
  (void)hipFree(c->d_probe);
  (void)hipHostFree(c->h_probe);
  (void)hipFree(c->nc_theta);
  (void)hipFree(c->nc_vec);
  (void)hipFree(c->d_nc_idx);
  (void)hipFree(c->d_nc_blk);
  (void)hipFree(c->d_nc_kept);
  (void)hipFree(c->d_nc_part);
  (void)hipFree(c->nc_lscratch);
  (void)hipHostFree(c->h_nc_s);
  (void)hipHostFree(c->h_nc_lam);
  (void)hipHostFree(c->h_nc_part);
  (void)hipFree(c->d_gemm);
  (void)hipHostFree(c->h_gemm);
  (void)hipFree(c->d_kprime);
  (void)hipHostFree(c->h_kprime);
  (void)hipHostFree(c->h_tasks);
  (void)hipHostFree(c->h_results);
  (void)hipHostFree(c->h_counters);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->st) (void)hipStreamDestroy(c->st);
}

// ---- gates on the host: ConstructPropagators -----------------------------------------
using cplx = std::complex<double>;

// A = V diag(w) V^H for a Hermitian n x n matrix (column-major) by cyclic complex Jacobi
// rotations; each rotation takes the phase off A_pq and then zeroes it with a real plane
// rotation, so V stays unitary and w real.
void herm_eig(int n, std::vector<cplx> A, std::vector<double>& w, std::vector<cplx>& V) {
  V.assign((size_t)n * n, cplx(0.0));
  for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int q = 0; q < n; ++q)
      for (int p = 0; p < n; ++p) {
        const double a2 = std::norm(A[p + (size_t)n * q]);
        tot += a2;
        if (p != q) off += a2;
      }
    if (off <= 1e-34 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const cplx apq = A[p + (size_t)n * q];
        const double g = std::abs(apq);
        if (g == 0.0) continue;
        const cplx ph = std::conj(apq) / g;  // D = diag(1, ph) makes the (p, q) entry real
        const double tau = (A[q + (size_t)n * q].real() - A[p + (size_t)n * p].real()) / (2.0 * g);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        // G = D J, J = [[cs, sn], [-sn, cs]] in the (p, q) plane
        const cplx gpp = cs, gpq = sn, gqp = -sn * ph, gqq = cs * ph;
        for (int r = 0; r < n; ++r) {  // A <- A G, V <- V G
          const cplx ap = A[r + (size_t)n * p], aq = A[r + (size_t)n * q];
          A[r + (size_t)n * p] = ap * gpp + aq * gqp;
          A[r + (size_t)n * q] = ap * gpq + aq * gqq;
          const cplx vp = V[r + (size_t)n * p], vq = V[r + (size_t)n * q];
          V[r + (size_t)n * p] = vp * gpp + vq * gqp;
          V[r + (size_t)n * q] = vp * gpq + vq * gqq;
        }
        for (int col = 0; col < n; ++col) {  // A <- G^H A
          const cplx ap = A[p + (size_t)n * col], aq = A[q + (size_t)n * col];
          A[p + (size_t)n * col] = std::conj(gpp) * ap + std::conj(gqp) * aq;
          A[q + (size_t)n * col] = std::conj(gpq) * ap + std::conj(gqq) * aq;
        }
        A[p + (size_t)n * q] = A[q + (size_t)n * p] = 0.0;
        A[p + (size_t)n * p] = A[p + (size_t)n * p].real();
        A[q + (size_t)n * q] = A[q + (size_t)n * q].real();
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A[i + (size_t)n * i].real();
}

// ---- Trotter schedules ------------------------------------------------------------------
// Size of the reference's propagator table (AllocateProp, TEBDtools_module.f90:937-998).
int gate_table_size(int L, int order, char bc) {
  const bool pbc = bc == 'P';
  if (order == 2) return pbc ? L : L - 1;
  return !pbc ? 2 * (L - 1) : (L % 2 == 0 ? 2 * L : 3 * L);
}

// 0-based table entry of bond i (1-based; i = L is the periodic edge, sites L and 1) in
// slot s (1-based), as NEWConstructPropagators fills the table
// (TEBDtools_module.f90:1079-1193).
int gate_index(int L, int order, char bc, int i, int s) {
  if (order == 2) return i - 1;
  if (bc != 'P' || L % 2 == 0) return 2 * (i - 1) + s - 1;
  return (i % 2 == 1) ? 3 * (i - 1) + s - 1 : 2 + 3 * (i - 2) + s - 1;  // odd L: i = L is odd
}

// Time factor of table entry (i, s) (dt times this is the gate's exponent), with the
// Forest-Ruth theta as the reference evaluates it, 1 / (2 - 2**(1.0/3.0)) with a
// single-precision exponent 1.0/3.0 (TEBDtools_module.f90:1135).
double gate_factor(int L, int order, char bc, int i, int s) {
  const bool edge = i == L, odd = i % 2 == 1;
  if (order == 2) {
    if (edge) return 1.0;
    if (odd) return 0.5;
    return (bc == 'P' && L % 2 == 1) ? 0.5 : 1.0;
  }
  const double th = 1.0 / (2.0 - std::pow(2.0, (double)(1.0f / 3.0f)));
  if (bc != 'P' || L % 2 == 0) {
    if (edge || !odd) return s == 1 ? th : 1.0 - 2.0 * th;
    return s == 1 ? 0.5 * th : 0.5 * (1.0 - th);
  }
  if (edge) return s == 1 ? th * th : (s == 2 ? (1.0 - 2.0 * th) * th : (1.0 - 2.0 * th) * (1.0 - 2.0 * th));
  if (odd) return s == 1 ? 0.5 * th : 0.5 * (1.0 - th);
  const double f[4] = {0.5 * th * th, 0.5 * (1.0 - th) * th, 0.5 * (1.0 - 2.0 * th) * th,
                       0.5 * (1.0 - th) * (1.0 - 2.0 * th)};
  return f[s - 1];
}

// One stage of a TrotterStep: 'G' every odd (parity 1) or even (parity 2) bond with its
// slot-s gate; 'E' the edge gate of slot s, applied on bond 1 after swapping site L next
// to site 1 (swaps on bonds L-1 .. 2) and followed by the swaps back (bonds 2 .. L-1).
struct Stage {
  char kind;
  int parity, slot;
};

// The stage sequence of TrotterStep for each order and boundary
// (propagation_module.f90:42-614).
std::vector<Stage> trotter_schedule(int L, int order, char bc) {
  auto G = [](int p, int s) { return Stage{'G', p, s}; };
  auto E = [](int s) { return Stage{'E', 0, s}; };
  if (order == 2 && bc != 'P') return {G(1, 1), G(2, 1), G(1, 1)};          // :82-114
  if (order == 2) {                                                         // :116-198
    if (L % 2 == 1) return {G(1, 1), G(2, 1), E(1), G(2, 1), G(1, 1)};
    return {G(1, 1), E(1), G(2, 1), G(1, 1)};
  }
  if (bc != 'P')                                                            // :200-245
    return {G(1, 1), G(2, 1), G(1, 2), G(2, 2), G(1, 2), G(2, 1), G(1, 1)};
  if (L % 2 == 0)                                                           // :541-612
    return {G(1, 1), E(1), G(2, 1), G(1, 2), E(2), G(2, 2), G(1, 2), E(1), G(2, 1), G(1, 1)};
  return {G(1, 1), G(2, 1), E(1), G(2, 2), E(2), G(2, 2), E(1), G(2, 1), G(1, 2),  // :260-538
          G(2, 3), E(2), G(2, 4), E(3), G(2, 4), E(2), G(2, 3), G(1, 2), G(2, 1), E(1),
          G(2, 2), E(2), G(2, 2), E(1), G(2, 1), G(1, 1)};
}

// The propagator table of one chain of L sites (NEWConstructPropagators,
// TEBDtools_module.f90:1079-1193, with matrix_exponential_c :176-222): entry (i, s) is
// exp(-i f dt H(i)) with the factor f of gate_factor; dt complex (imaginary time passes
// dt = -i dtau).  H: L-1 (open) or L (periodic, the last one the edge) matrices of n x n
// complex, column-major; U: gate_table_size entries.
void construct_propagators_host(const double* H, int L, int n, int order, char bc, cplx dt, double* U) {
  const int nb = bc == 'P' ? L : L - 1;
  const int slots_odd = order == 2 ? 1 : 2;
  const int slots_even = order == 2 ? 1 : ((bc == 'P' && L % 2 == 1) ? 4 : 2);
  const int slots_edge = order == 2 ? 1 : ((L % 2 == 1) ? 3 : 2);
  std::vector<cplx> A((size_t)n * n), V;
  std::vector<double> w;
  std::vector<cplx> f(n);
  for (int i = 1; i <= nb; ++i) {
    const cplx* h = reinterpret_cast<const cplx*>(H) + (size_t)(i - 1) * n * n;
    for (size_t e = 0; e < (size_t)n * n; ++e) A[e] = h[e];
    herm_eig(n, A, w, V);
    const int ns = i == L ? slots_edge : (i % 2 == 1 ? slots_odd : slots_even);
    for (int sl = 1; sl <= ns; ++sl) {
      const cplx t = gate_factor(L, order, bc, i, sl) * dt;
      for (int k = 0; k < n; ++k) f[k] = std::exp(cplx(0.0, -1.0) * t * w[k]);
      cplx* u = reinterpret_cast<cplx*>(U) + (size_t)gate_index(L, order, bc, i, sl) * n * n;
      for (int col = 0; col < n; ++col)
        for (int r = 0; r < n; ++r) {
          cplx acc = 0.0;
          for (int k = 0; k < n; ++k) acc += V[r + (size_t)n * k] * f[k] * std::conj(V[col + (size_t)n * k]);
          u[r + (size_t)n * col] = acc;
        }
    }
  }
}

// ---- imaginary-time chi stages ----------------------------------------------------------
// Cuts every bond to at most chi values, as the reference's copy of the state into chi-sized
// arrays does at the start of a stage (CopyGamLam, TEBDtools_module.f90:523-553: entries past
// chi dropped, lambda not renormalised), and caps later truncations at chi.
int set_stage(tebd_ctx* c, int chi) {
  const int L = c->L, d = c->d;
  std::vector<int> nd(c->chi);
  bool cut = false;
  for (int l = 0; l <= L; ++l) {
    nd[l] = std::min(c->chi[l], chi);
    cut = cut || nd[l] != c->chi[l];
  }
  if (cut) {
    zc* tmp = c->W;  // slot 0 holds Rmax x Cpmax >= chi_max d chi_max entries
    for (int s = 0; s < L; ++s) {
      if (nd[s] == c->chi[s] && nd[s + 1] == c->chi[s + 1]) continue;
      zc* slab = c->gam + (size_t)s * c->gslab();
      HIPCHK(c, hipMemcpyAsync(tmp, slab, sizeof(zc) * (size_t)c->chi[s] * d * c->chi[s + 1], hipMemcpyDeviceToDevice,
                               c->st));
      HIPCHK(c, launch_shrink_slab(tmp, c->chi[s], d, slab, nd[s], nd[s + 1], c->st));
    }
    HIPCHK(c, launch_zero_lambda_tail(c->lam, L + 1, chi, c->chi_max, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->chi = nd;
    for (int l = 0; l <= L; ++l) c->last_rank[l] = std::min(c->last_rank[l], chi);
  }
  if (c->nc)  // labels past chi reset, as CopyLabel does (TEBDtools_module.f90:623-642)
    for (int l = 0; l <= L; ++l)
      for (int a = chi; a < c->chi_max; ++a) {
        c->nc_LL[(size_t)l * c->chi_max + a] = 10000;
        c->nc_LR[(size_t)l * c->chi_max + a] = 10000;
      }
  c->stage_chi = chi;
  return TEBD_OK;
}

// lambda_l(1) of every bond l = 0..L (the quantity ImagTimeProp's convergence test follows).
int largest_lambdas(tebd_ctx* c, std::vector<double>& out) {
  out.resize(c->L + 1);
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy2D(out.data(), sizeof(double), c->lam, sizeof(double) * c->chi_max, sizeof(double), c->L + 1,
                        hipMemcpyDeviceToHost));
  return TEBD_OK;
}

// ---- allocation-free two-site update (tebd_two_site_op_noalloc) -------------------
// One bond's workspace lives in a file-scope device array instead of hipMalloc'd slots:
// bond matrices of up to kNoAllocSide x kNoAllocSide (d chi_l, d chi_r <= 256) and at
// most kNoAllocSide kept values.  Every device holds its own copy; its address is
// looked up once per device with hipGetSymbolAddress, and only that address is passed
// to kernels and copies.
constexpr int kNoAllocSide = 256;
constexpr size_t kNoAllocAlign = 256;
constexpr size_t na_up(size_t b) { return (b + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; }
constexpr size_t kNoAllocNblk = kNoAllocSide / JB;
constexpr size_t kNoAllocTq = (kNoAllocSide / kQrPanel + 1) * kQrPanel * kQrPanel;  // = tq_stride at Cpmax 256
constexpr size_t kNoAllocBytes =
    5 * na_up(sizeof(zc) * kNoAllocSide * kNoAllocSide) +                      // W, V, X, Yr, Zc
    2 * na_up(sizeof(double) * kNoAllocSide) + 2 * na_up(sizeof(int) * kNoAllocSide) +  // s, qnrm, perm, qpiv
    na_up(sizeof(zc) * kNoAllocSide) + na_up(sizeof(double) * 2) +              // tau, qstate
    na_up(sizeof(zc) * kNoAllocTq) + na_up(sizeof(int) * 4) +                   // Tq, counters
    na_up(sizeof(unsigned) * kNoAllocNblk) + na_up(sizeof(unsigned long long) * kNoAllocNblk * kNoAllocNblk) +
    na_up(sizeof(zc) * (kNoAllocNblk / 2) * JC * JC) + na_up(sizeof(int) * 2) +  // Qg, wl_count
    na_up(sizeof(int2) * 2 * (kNoAllocNblk / 2)) + 2 * na_up(sizeof(double) * RESULT_STRIDE) +
    na_up(sizeof(BondTask) * kStages);
constexpr size_t kNoAllocScratch = 6u << 20;
static_assert(kNoAllocBytes <= kNoAllocScratch, "no-alloc workspace exceeds its array");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;                        // one call at a time per process (shared host staging)
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};   // device address of `scratch` on each device

__global__ void k_zero_words(unsigned long long* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

// Points the workspace members of `c` (one slot) into the array at `base`.
void noalloc_layout(tebd_ctx* c, unsigned char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* q = base + off;
    off += na_up(bytes);
    return q;
  };
  const size_t n = kNoAllocSide, zsq = sizeof(zc) * n * n;
  c->W = (zc*)take(zsq);
  c->V = (zc*)take(zsq);
  c->X = (zc*)take(zsq);
  c->Yr = (zc*)take(zsq);
  c->Zc = (zc*)take(zsq);
  c->s = (double*)take(sizeof(double) * n);
  c->qnrm = (double*)take(sizeof(double) * n);
  c->perm = (int*)take(sizeof(int) * n);
  c->qpiv = (int*)take(sizeof(int) * n);
  c->tau = (zc*)take(sizeof(zc) * n);
  c->qstate = (double*)take(sizeof(double) * 2);
  c->Tq = (zc*)take(sizeof(zc) * kNoAllocTq);
  c->counters = (int*)take(sizeof(int) * 4);
  c->blkver = (unsigned*)take(sizeof(unsigned) * kNoAllocNblk);
  c->pairseen = (unsigned long long*)take(sizeof(unsigned long long) * kNoAllocNblk * kNoAllocNblk);
  c->Qg = (zc*)take(sizeof(zc) * (kNoAllocNblk / 2) * JC * JC);
  c->wl_count = (int*)take(sizeof(int) * 2);
  c->wl_items = (int2*)take(sizeof(int2) * 2 * (kNoAllocNblk / 2));
  c->results = (double*)take(sizeof(double) * RESULT_STRIDE);
  c->results2 = (double*)take(sizeof(double) * RESULT_STRIDE);
  c->d_tasks = (BondTask*)take(sizeof(BondTask) * kStages);
}

// Device bytes of one per-bond workspace slot (every per-slot array tebd_create
// allocates), for bond matrices of up to Rmax x Cpmax.
// Split-operand theta workspace per task: row capacity and bytes of one operand's slices.
size_t split_rows(const tebd_ctx* c) { return (size_t)c->d * round_up(c->chi_max, 16); }
int split_pitch(const tebd_ctx* c) { return round_up(c->chi_max, 64); }
size_t split_stride(const tebd_ctx* c) { return (size_t)2 * kSplitSlices * split_rows(c) * split_pitch(c); }

double slot_bytes(const tebd_ctx* c) {
  const double R = c->Rmax, Cp = c->Cpmax, chi = c->chi_max, nblk = c->nblk_max, z = sizeof(zc);
  double b = z * (R * Cp + 2.0 * Cp * Cp + R * chi + Cp * chi + Cp) + 8.0 * Cp * 2 + 4.0 * Cp * 2 + 16.0 +
             z * (double)tq_stride(c) + 16.0 + 4.0 * nblk + 8.0 * nblk * nblk + z * (nblk / 2) * JC * JC +
             16.0 * (nblk / 2) + 2.0 * 8 * RESULT_STRIDE + 4.0 * TEBD_MLP_INPUTS + 4.0 +
             2.0 * kStages * sizeof(BondTask);
  if (c->mode & TEBD_MODE_LEARNED_RANK)
    b += 2.0 * z * R * Cp + 8.0 * chi + 2.0 * kGemmRegions * sizeof(GemmTask) + 8.0 + z * (double)probe_stride(c) +
         2.0 * sizeof(ProbeTask);
  if (c->mode & TEBD_MODE_SPLIT) b += 2.0 * (double)split_stride(c) + 8.0 * split_rows(c) + 4.0 * split_pitch(c);
  return b;
}

}  // namespace

extern "C" {

// nchain chains of Lc sites (L = nchain Lc) with the schedule (order, bc).
int create_impl(tebd_ctx** out, int nchain, int Lc, int d, int chi_max, int order, char bc, unsigned mode) {
  if (!out) return TEBD_ERR_ARG;
  *out = nullptr;
  if (bc == 'o') bc = 'O';
  if (bc == 'p') bc = 'P';
  if (nchain < 1 || Lc < 2 || d < 2 || chi_max < 1 || (long long)nchain * Lc > (1 << 24)) return TEBD_ERR_ARG;
  if ((order != 2 && order != 5) || (bc != 'O' && bc != 'P')) return TEBD_ERR_ARG;
  if (bc == 'P' && (Lc < 3 || nchain > 1)) return TEBD_ERR_ARG;  // periodic: one chain of at least 3 sites
  if (!(d == 2 || d == 3 || d == 4 || d == 5 || d == 6 || d == 8)) return TEBD_ERR_ARG;
  if (d * chi_max > 4096) return TEBD_ERR_ARG;
  if ((mode & TEBD_MODE_MIXED) && (mode & TEBD_MODE_SPLIT)) return TEBD_ERR_ARG;  // one theta precision
  const int L = nchain * Lc;
  tebd_ctx* c = new tebd_ctx();
  c->L = L;
  c->d = d;
  c->chi_max = chi_max;
  c->chain_len = Lc;
  c->nchain = nchain;
  c->order = order;
  c->bc = bc;
  c->table = gate_table_size(Lc, order, bc);
  c->mode = mode;
  c->chi.assign(L + 1, 1);
  c->Rmax = d * chi_max;
  c->Cpmax = round_up(d * chi_max, JC);
  c->nblk_max = c->Cpmax / JB;
  // Batched mode gives every bond of a half-sweep a workspace slot, up to kSlotBudget
  // bytes of slots; longer half-sweeps (ensembles, very large bond matrices) run in
  // batches of that many bonds.
  const int want = (mode & TEBD_MODE_BATCHED) ? std::max(1, L / 2) : 1;
  c->ntask_max = (int)std::max(1.0, std::min((double)want, std::floor(kSlotBudget / slot_bytes(c))));
  auto bad = [&](hipError_t e) {
    if (e != hipSuccess) {
      free_all(c);
      delete c;
      return true;
    }
    return false;
  };
  const int nt = c->ntask_max;
  if (bad(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking))) return TEBD_ERR_HIP;
  if (bad(hipEventCreate(&c->ev0)) || bad(hipEventCreate(&c->ev1))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->gam, sizeof(zc) * c->gslab() * L))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->lam, sizeof(double) * (size_t)(L + 1) * chi_max))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->gates, sizeof(zc) * c->d4() * ((size_t)nchain * c->table + 2)))) return TEBD_ERR_HIP;
  {  // the SWAP gate of the periodic schedules' swaps: U(i d + j, k d + l) = delta(i,l) delta(j,k)
    const int d2 = d * d;
    std::vector<zc> sw((size_t)d2 * d2, zc{0.0, 0.0});
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) sw[(i * d + j) + (size_t)d2 * (j * d + i)] = zc{1.0, 0.0};
    if (bad(hipMemcpy(c->gates + c->d4() * ((size_t)nchain * c->table + 1), sw.data(), sizeof(zc) * sw.size(),
                      hipMemcpyHostToDevice)))
      return TEBD_ERR_HIP;
  }
  if (bad(hipMalloc(&c->W, sizeof(zc) * (size_t)c->Rmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->V, sizeof(zc) * (size_t)c->Cpmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->s, sizeof(double) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->perm, sizeof(int) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->X, sizeof(zc) * (size_t)c->Cpmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->Yr, sizeof(zc) * (size_t)c->Rmax * chi_max * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->Zc, sizeof(zc) * (size_t)c->Cpmax * chi_max * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->qnrm, sizeof(double) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->qpiv, sizeof(int) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->tau, sizeof(zc) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->qstate, sizeof(double) * 2 * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->Tq, sizeof(zc) * tq_stride(c) * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_qstate, sizeof(double) * 2 * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->counters, sizeof(int) * 4 * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->blkver, sizeof(unsigned) * c->nblk_max * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->pairseen, sizeof(unsigned long long) * c->nblk_max * c->nblk_max * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->Qg, sizeof(zc) * (size_t)(c->nblk_max / 2) * JC * JC * nt))) return TEBD_ERR_HIP;
  c->wl_cap = nt * (c->nblk_max / 2);
  if (bad(hipMalloc(&c->wl_count, sizeof(int) * 2))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->wl_items, sizeof(int2) * 2 * c->wl_cap))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->results, sizeof(double) * RESULT_STRIDE * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->results2, sizeof(double) * RESULT_STRIDE * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->mlp, sizeof(float) * TEBD_MLP_PARAMS))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->d_feat, sizeof(float) * TEBD_MLP_INPUTS * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->d_pred, sizeof(float) * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_feat, sizeof(float) * TEBD_MLP_INPUTS * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_pred, sizeof(float) * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->d_tasks, sizeof(BondTask) * nt * kStages))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_tasks, sizeof(BondTask) * nt * kStages))) return TEBD_ERR_HIP;
  if (mode & TEBD_MODE_LEARNED_RANK) {
    if (bad(hipMalloc(&c->Zb, sizeof(zc) * (size_t)c->Rmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->Bb, sizeof(zc) * (size_t)c->Rmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->lam_scratch, sizeof(double) * (size_t)chi_max * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->Pb, sizeof(zc) * probe_stride(c) * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->d_probe, sizeof(ProbeTask) * nt))) return TEBD_ERR_HIP;
    if (bad(hipHostMalloc(&c->h_probe, sizeof(ProbeTask) * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->d_gemm, sizeof(GemmTask) * nt * kGemmRegions))) return TEBD_ERR_HIP;
    if (bad(hipHostMalloc(&c->h_gemm, sizeof(GemmTask) * nt * kGemmRegions))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->d_kprime, sizeof(int) * nt))) return TEBD_ERR_HIP;
    if (bad(hipHostMalloc(&c->h_kprime, sizeof(int) * nt))) return TEBD_ERR_HIP;
  }
  if (mode & TEBD_MODE_SPLIT) {
    c->split.rows = split_rows(c);
    c->split.pitch = split_pitch(c);
    c->split.stride = split_stride(c);
    if (bad(hipMalloc(&c->split.a, c->split.stride * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->split.b, c->split.stride * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->split.ea, sizeof(int) * c->split.rows * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->split.eb, sizeof(int) * c->split.rows * nt))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&c->split.eg, sizeof(int) * c->split.pitch * nt))) return TEBD_ERR_HIP;
  }
  c->last_rank.assign(L + 1, 0);
  c->last_trunc.assign(L + 1, 0.0);
  if (bad(hipHostMalloc(&c->h_results, sizeof(double) * RESULT_STRIDE * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_counters, sizeof(int) * 4 * nt))) return TEBD_ERR_HIP;
  *out = c;
  return TEBD_OK;
}

int tebd_create(tebd_ctx** out, int L, int d, int chi_max, int order, char bc, unsigned mode) {
  return create_impl(out, 1, L, d, chi_max, order, bc, mode);
}

int tebd_create_ensemble(tebd_ctx** out, int nchain, int L, int d, int chi_max, int order, char bc, unsigned mode) {
  return create_impl(out, nchain, L, d, chi_max, order, bc, mode);
}

void tebd_destroy(tebd_ctx* c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  free_all(c);
  delete c;
}

int tebd_set_gates(tebd_ctx* c, const double* U, int nb) {
  if (!c) return TEBD_ERR_ARG;
  if (!U || nb != c->nchain * c->table)
    return fail(c, TEBD_ERR_ARG, "set_gates: need the schedule's propagator table (L-1 gates per chain for "
                                 "2nd order with open boundaries)");
  // chain after chain, each its whole table
  HIPCHK(c, hipMemcpy(c->gates, U, sizeof(zc) * c->d4() * nb, hipMemcpyHostToDevice));
  c->have_gates = true;
  return TEBD_OK;
}

int tebd_upload(tebd_ctx* c, const double* const* gamma, const double* const* lambda, int chi_in) {
  return upload_impl(c, gamma, lambda, chi_in);
}

int tebd_upload_packed(tebd_ctx* c, const double* gamma, const double* lambda, int chi_in) {
  if (!c || !gamma || !lambda || chi_in < 1) return fail(c, TEBD_ERR_ARG, "upload: bad arguments");
  std::vector<const double*> g(c->L), l(c->L + 1);
  for (int s = 0; s < c->L; ++s) g[s] = gamma + (size_t)2 * chi_in * c->d * chi_in * s;
  for (int b = 0; b <= c->L; ++b) l[b] = lambda + (size_t)chi_in * b;
  return upload_impl(c, g.data(), l.data(), chi_in);
}

int tebd_download(tebd_ctx* c, double* const* gamma, double* const* lambda, int chi_out) {
  return download_impl(c, gamma, lambda, chi_out);
}

int tebd_download_packed(tebd_ctx* c, double* gamma, double* lambda, int chi_out) {
  if (!c || !gamma || !lambda || chi_out < 1) return fail(c, TEBD_ERR_ARG, "download: bad arguments");
  std::vector<double*> g(c->L), l(c->L + 1);
  for (int s = 0; s < c->L; ++s) g[s] = gamma + (size_t)2 * chi_out * c->d * chi_out * s;
  for (int b = 0; b <= c->L; ++b) l[b] = lambda + (size_t)chi_out * b;
  return download_impl(c, g.data(), l.data(), chi_out);
}

int tebd_download_lambdas(tebd_ctx* c, double* lambda, int chi_out) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (!lambda || chi_out < 1) return fail(c, TEBD_ERR_ARG, "download_lambdas: bad arguments");
  for (int l = 0; l <= c->L; ++l)
    if (c->chi[l] > chi_out) return fail(c, TEBD_ERR_SHAPE, "download_lambdas: chi_out smaller than a bond dimension");
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::vector<double> lam((size_t)(c->L + 1) * c->chi_max);
  HIPCHK(c, hipMemcpy(lam.data(), c->lam, lam.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int l = 0; l <= c->L; ++l)
    for (int a = 0; a < chi_out; ++a)
      lambda[(size_t)l * chi_out + a] = a < c->chi[l] ? lam[(size_t)l * c->chi_max + a] : 0.0;
  return TEBD_OK;
}

int tebd_trotter_step(tebd_ctx* c, double* truncerr) {
  int e = ensure_state(c, true);
  if (e) return e;
  const int L = c->L, Lc = c->chain_len;
  double total = 0.0;
  HIPCHK(c, hipEventRecord(c->ev0, c->st));
  std::vector<int> links;
  std::vector<const zc*> gp;
  c->step_index++;
  const zc* swap = c->gates + c->d4() * ((size_t)c->nchain * c->table + 1);
  // bonds one after another (the swaps and edge gates of the periodic schedules, and
  // every bond in the per-bond FP64 mode)
  auto one = [&](int link, const zc* g) {
    double t = 0.0;
    int r = run_bonds(c, &link, 1, &g, &t);
    total += t;
    return r;
  };
  // a swap: with number conservation SwappingNC, which keeps theta unnormalised
  auto swp = [&](int link) {
    double t = 0.0;
    int r = c->nc ? run_bonds_nc(c, &link, 1, &swap, &t, false) : run_bonds(c, &link, 1, &swap, &t);
    total += t;
    return r;
  };
  for (const Stage& stg : trotter_schedule(Lc, c->order, c->bc)) {
    if (stg.kind == 'E') {  // site L next to site 1, the edge gate on bond 1, and back
      for (int i = Lc - 1; i >= 2 && !e; --i) e = swp(i - 1);
      if (!e) e = one(0, c->gates + c->d4() * (size_t)gate_index(Lc, c->order, c->bc, Lc, stg.slot));
      for (int i = 2; i <= Lc - 1 && !e; ++i) e = swp(i - 1);
      if (e) return e;
      continue;
    }
    links.clear();
    gp.clear();
    for (int b = 0; b < L - 1; ++b) {
      // the reference's bond numbering restarts in every chain of an ensemble
      const int i = b % Lc + 1, r = b / Lc;
      if (is_cut(c, b) || (i % 2 == 1 ? 1 : 2) != stg.parity) continue;
      links.push_back(b);
      gp.push_back(c->gates + c->d4() * ((size_t)r * c->table + gate_index(Lc, c->order, c->bc, i, stg.slot)));
    }
    if (c->mode & TEBD_MODE_BATCHED) {
      // batches of at most ntask_max bonds (one workspace slot each)
      for (size_t b0 = 0; b0 < links.size(); b0 += c->ntask_max) {
        const int nb = (int)std::min(links.size() - b0, (size_t)c->ntask_max);
        double t = 0.0;
        e = c->nc ? run_bonds_nc(c, links.data() + b0, nb, gp.data() + b0, &t, true)
                  : run_bonds_split(c, links.data() + b0, nb, gp.data() + b0, &t);
        if (e) return e;
        total += t;
      }
    } else {
      for (size_t k = 0; k < links.size(); ++k)
        if ((e = one(links[k], gp[k]))) return e;
    }
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->st));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats[3] += ms;
  if (truncerr) *truncerr = total;
  return TEBD_OK;
}

int tebd_two_site_op(tebd_ctx* c, int link, const double* op2, double* truncerr) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (link < 1 || link > c->L - 1 || !op2) return fail(c, TEBD_ERR_ARG, "two_site_op: bad link or operator");
  if (is_cut(c, link - 1)) return fail(c, TEBD_ERR_ARG, "two_site_op: bond joins two chains of the ensemble");
  zc* slot = c->gates + c->d4() * (size_t)c->nchain * c->table;
  HIPCHK(c, hipMemcpyAsync(slot, op2, sizeof(zc) * c->d4(), hipMemcpyHostToDevice, c->st));
  const int l0 = link - 1;
  const zc* gp = slot;
  double t = 0.0;
  e = run_bonds(c, &l0, 1, &gp, &t);
  if (e) return e;
  if (truncerr) *truncerr = t;
  return TEBD_OK;
}

// CanonicalFormAll of chains r0, r0 + dr, ... of the context together: per batch the
// same bond of every one of them, left to right, then right to left, as
// propagation_module.f90:617-634 does for one chain.
int canonical_form_chains(tebd_ctx* c, int r0, int dr) {
  std::vector<int> all;
  for (int r = r0; r < c->nchain; r += dr) all.push_back(r);
  if (all.empty()) return TEBD_OK;
  // chains in groups of at most ntask_max (CanonicalForm workspace sets); each group runs
  // its chains' full left-to-right and right-to-left sweeps
  int e = TEBD_OK;
  const int Lc = c->chain_len;
  for (size_t g0 = 0; g0 < all.size(); g0 += c->ntask_max) {
    const std::vector<int> rs(all.begin() + g0, all.begin() + std::min(all.size(), g0 + c->ntask_max));
    const int n = (int)rs.size();
    if ((e = ensure_cf_workspace(c, n))) return e;
    std::vector<int> bonds(n);
    for (int k = 0; k < Lc - 1; ++k) {
      for (int j = 0; j < n; ++j) bonds[j] = rs[j] * Lc + k;
      if ((e = canonical_form_batch(c, bonds.data(), n))) return e;
    }
    for (int k = Lc - 2; k >= 0; --k) {
      for (int j = 0; j < n; ++j) bonds[j] = rs[j] * Lc + k;
      if ((e = canonical_form_batch(c, bonds.data(), n))) return e;
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  drain_profile(c);
  return TEBD_OK;
}

// CanonicalFormAllNC (propagation_module.f90:1943-1983): the number-conserving update with
// the identity on bonds 1 .. L-1, then L-1 .. 1, one bond at a time (each bond's labels and
// Gammas come from the update before it).
int canonical_form_all_nc(tebd_ctx* c) {
  const int n = c->d * c->d;
  std::vector<double> eye((size_t)2 * n * n, 0.0);
  for (int k = 0; k < n; ++k) eye[(size_t)2 * (k + (size_t)n * k)] = 1.0;
  zc* slot = c->gates + c->d4() * (size_t)c->nchain * c->table;
  HIPCHK(c, hipMemcpyAsync(slot, eye.data(), sizeof(zc) * c->d4(), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const zc* gp = slot;
  double t = 0.0;
  int e = TEBD_OK;
  for (int l = 0; l < c->L - 1 && !e; ++l) e = run_bonds_nc(c, &l, 1, &gp, &t, true);
  for (int l = c->L - 2; l >= 0 && !e; --l) e = run_bonds_nc(c, &l, 1, &gp, &t, true);
  return e;
}

int tebd_canonical_form_all(tebd_ctx* c) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (c->nc) return canonical_form_all_nc(c);
  // the chains of an ensemble are independent: with lanes, chain r goes to lane r mod n,
  // each lane on its own stream and host thread; each chain sees the same sequence of
  // bond operations either way
  // (a lane needs one workspace slot here, so lanes already made for the Trotter step
  // are used as they are)
  int n = std::min(c->nlanes, c->nchain);
  if (c->lanes_made >= 2 && n >= 2) n = std::min(c->lanes_made, c->nchain);
  if (n >= 2 && c->ntask_max >= n && (c->lanes_made >= n || ensure_lanes(c, n))) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    int err[kMaxLanes] = {};
    auto work = [&](int k) {
      tebd_ctx* l = c->lanes[k];
      lane_pull(l, c);
      err[k] = canonical_form_chains(l, k, n);
    };
    std::vector<std::thread> workers;
    for (int k = 1; k < n; ++k) workers.emplace_back(work, k);
    work(0);
    for (auto& w : workers) w.join();
    for (int k = 0; k < n; ++k) {
      if (err[k]) return fail(c, err[k], c->lanes[k]->err);
      lane_push(c, c->lanes[k], nullptr, 0);
    }
    return TEBD_OK;
  }
  return canonical_form_chains(c, 0, 1);
}

int tebd_canonical_form(tebd_ctx* c, int link) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (c->nc) return fail(c, TEBD_ERR_UNSUPPORTED, "canonical_form: not with number conservation");
  if (link < 1 || link > c->L - 1) return fail(c, TEBD_ERR_ARG, "canonical_form: bad link");
  if (is_cut(c, link - 1)) return fail(c, TEBD_ERR_ARG, "canonical_form: bond joins two chains of the ensemble");
  if ((e = ensure_cf_workspace(c, 1))) return e;
  const int l0 = link - 1;
  if ((e = canonical_form_batch(c, &l0, 1))) return e;
  HIPCHK(c, hipStreamSynchronize(c->st));
  drain_profile(c);
  return TEBD_OK;
}

int tebd_expval_two_site(tebd_ctx* c, const double* ops, double* out) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (!ops || !out) return fail(c, TEBD_ERR_ARG, "expval_two_site: null operator or output");
  const int L = c->L, d = c->d, nt = c->ntask_max, d2 = d * d;
  if (!c->obs_ops) {
    HIPCHK(c, hipMalloc(&c->obs_ops, sizeof(zc) * c->d4() * (L - 1)));
    HIPCHK(c, hipMalloc(&c->ident, sizeof(zc) * c->d4()));
    std::vector<zc> id(c->d4());
    for (int r = 0; r < d2; ++r)
      for (int q = 0; q < d2; ++q) id[r + (size_t)d2 * q] = zc{r == q ? 1.0 : 0.0, 0.0};
    HIPCHK(c, hipMemcpy(c->ident, id.data(), sizeof(zc) * c->d4(), hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpyAsync(c->obs_ops, ops, sizeof(zc) * c->d4() * (L - 1), hipMemcpyHostToDevice, c->st));
  BondTask *S0 = c->h_tasks, *S1 = S0 + nt;
  BondTask *D0 = c->d_tasks, *D1 = D0 + nt;
  for (int b0 = 0; b0 < L - 1; b0 += nt) {
    const int nb = std::min(nt, L - 1 - b0);
    int max_chil = 1, max_chir = 1;
    for (int i = 0; i < nb; ++i) {
      const int link = b0 + i;
      BondTask& t = S0[i];
      std::memset(&t, 0, sizeof(t));
      t.chil = c->chi[link]; t.chic = c->chi[link + 1]; t.chir = c->chi[link + 2]; t.d = d;
      t.gl = c->gam + (size_t)link * c->gslab();
      t.gr = c->gam + (size_t)(link + 1) * c->gslab();
      t.ll = c->lam + (size_t)link * c->chi_max;
      t.lc = c->lam + (size_t)(link + 1) * c->chi_max;
      t.lr = c->lam + (size_t)(link + 2) * c->chi_max;
      t.R = d * t.chil;
      t.C = d * t.chir;
      t.gate = c->ident;
      t.W = c->W + (size_t)i * c->Rmax * c->Cpmax;
      S1[i] = t;
      S1[i].gate = c->obs_ops + (size_t)link * c->d4();
      S1[i].W = c->V + (size_t)i * c->Cpmax * c->Cpmax;
      max_chil = std::max(max_chil, t.chil);
      max_chir = std::max(max_chir, t.chir);
    }
    HIPCHK(c, hipMemcpyAsync(D0, S0, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(D1, S1, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_theta_gate(D0, nb, d, max_chil, max_chir, c->st));
    HIPCHK(c, launch_theta_gate(D1, nb, d, max_chil, max_chir, c->st));
    HIPCHK(c, launch_zdotc(D0, D1, nb, c->results, c->st));
    HIPCHK(c, hipMemcpyAsync(out + 2 * b0, c->results, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  return TEBD_OK;
}

int tebd_set_learned_params(tebd_ctx* c, double tol_missed, double tol_trunc, double rank_slack) {
  if (!c) return TEBD_ERR_ARG;
  if (!(tol_missed > 0.0 && tol_missed < 1.0) || !(tol_trunc > 0.0 && tol_trunc < 1.0) ||
      !(rank_slack >= 0.0 && rank_slack <= 4.0))
    return fail(c, TEBD_ERR_ARG, "learned params: tolerances in (0, 1), rank slack in [0, 4]");
  c->tol_missed = tol_missed;
  c->tol_trunc = tol_trunc;
  c->rank_slack = rank_slack;
  return TEBD_OK;
}

int tebd_rank_model_set(tebd_ctx* c, const float* params, int n) {
  if (!c) return TEBD_ERR_ARG;
  if (!params || n == 0) {
    c->have_model = false;
    return TEBD_OK;
  }
  if (n != TEBD_MLP_PARAMS) return fail(c, TEBD_ERR_ARG, "rank model: wrong parameter count");
  HIPCHK(c, hipMemcpy(c->mlp, params, sizeof(float) * n, hipMemcpyHostToDevice));
  c->have_model = true;
  return TEBD_OK;
}

int tebd_rank_model_get(tebd_ctx* c, float* params, int n) {
  if (!c || !params || n != TEBD_MLP_PARAMS) return TEBD_ERR_ARG;
  if (!c->have_model) return fail(c, TEBD_ERR_STATE, "no rank model installed");
  HIPCHK(c, hipMemcpy(params, c->mlp, sizeof(float) * n, hipMemcpyDeviceToHost));
  return TEBD_OK;
}

int tebd_set_conserved(tebd_ctx* c, const int* conserv, int tot_num) {
  if (!c) return TEBD_ERR_ARG;
  if (!conserv && tot_num < 0) {  // off
    c->nc = false;
    return TEBD_OK;
  }
  if (tot_num < 0) return fail(c, TEBD_ERR_ARG, "set_conserved: negative total");
  if (c->nchain > 1) return fail(c, TEBD_ERR_ARG, "set_conserved: number conservation needs a single chain");
  std::vector<int> cv(c->d);
  int maxf = 0;
  for (int i = 0; i < c->d; ++i) {
    cv[i] = conserv ? conserv[i] : i;
    if (cv[i] < 0 || cv[i] >= 1000) return fail(c, TEBD_ERR_ARG, "set_conserved: local numbers in [0, 1000)");
    maxf = std::max(maxf, cv[i]);
  }
  c->nc_conserv = cv;
  c->nc_tot = tot_num;
  c->nc_maxf = maxf;
  c->nc = true;
  if (c->nc_LL.size() != (size_t)(c->L + 1) * c->chi_max) {
    c->nc_LL.assign((size_t)(c->L + 1) * c->chi_max, 10000);
    c->nc_LR.assign((size_t)(c->L + 1) * c->chi_max, 10000);
  }
  return TEBD_OK;
}

int tebd_set_labels(tebd_ctx* c, const int* label_left, const int* label_right, int chi_in) {
  if (!c || !label_left || !label_right || chi_in < 1) return fail(c, TEBD_ERR_ARG, "set_labels: bad arguments");
  if (!c->nc) return fail(c, TEBD_ERR_STATE, "set_labels: call tebd_set_conserved first");
  const int chi = c->chi_max, L = c->L;
  std::vector<int> LL((size_t)(L + 1) * chi, 10000), LR((size_t)(L + 1) * chi, 10000);
  for (int l = 0; l <= L; ++l)
    for (int a = 0; a < chi_in; ++a) {
      const int x = label_left[(size_t)l * chi_in + a], y = label_right[(size_t)l * chi_in + a];
      if ((x < 1000) != (y < 1000)) return fail(c, TEBD_ERR_ARG, "set_labels: left and right labels disagree on used states");
      if (x >= 1000) continue;
      if (a >= chi) return fail(c, TEBD_ERR_SHAPE, "set_labels: a used state past chi_max");
      LL[(size_t)l * chi + a] = x;
      LR[(size_t)l * chi + a] = y;
    }
  c->nc_LL.swap(LL);
  c->nc_LR.swap(LR);
  return TEBD_OK;
}

int tebd_get_labels(tebd_ctx* c, int* label_left, int* label_right, int chi_out) {
  if (!c || !label_left || !label_right || chi_out < 1) return fail(c, TEBD_ERR_ARG, "get_labels: bad arguments");
  if (!c->nc) return fail(c, TEBD_ERR_STATE, "get_labels: no number conservation set");
  const int chi = c->chi_max;
  for (int l = 0; l <= c->L; ++l)
    for (int a = 0; a < chi_out; ++a) {
      label_left[(size_t)l * chi_out + a] = a < chi ? c->nc_LL[(size_t)l * chi + a] : 10000;
      label_right[(size_t)l * chi_out + a] = a < chi ? c->nc_LR[(size_t)l * chi + a] : 10000;
    }
  return TEBD_OK;
}

int tebd_rank_log(tebd_ctx* c, int on) {
  if (!c) return TEBD_ERR_ARG;
  c->rank_log = on != 0;
  if (c->rank_log) {
    c->log_X.clear();
    c->log_y.clear();
    c->log_check.clear();
  }
  return TEBD_OK;
}

int tebd_sketch_log_get(tebd_ctx* c, double* out, int max_n) {
  if (!c || max_n < 0) return TEBD_ERR_ARG;
  const int n = (int)(c->log_check.size() / 8);
  const int m = std::min(n, max_n);
  if (m > 0 && !out) return fail(c, TEBD_ERR_ARG, "sketch_log_get: null output");
  if (m > 0) std::memcpy(out, c->log_check.data(), sizeof(double) * 8 * m);
  return n;
}

int tebd_rank_log_get(tebd_ctx* c, float* X, float* y, int max_n) {
  if (!c || max_n < 0) return TEBD_ERR_ARG;
  const int n = (int)c->log_y.size();
  const int m = std::min(n, max_n);
  if (m > 0 && (!X || !y)) return fail(c, TEBD_ERR_ARG, "rank_log_get: null output");
  if (m > 0) {
    std::memcpy(X, c->log_X.data(), sizeof(float) * TEBD_MLP_INPUTS * m);
    std::memcpy(y, c->log_y.data(), sizeof(float) * m);
  }
  return n;
}

int tebd_set_cutoff(tebd_ctx* c, double cutoff) {
  if (!c) return TEBD_ERR_ARG;
  if (!(cutoff >= 0.0) || cutoff >= 1.0) return fail(c, TEBD_ERR_ARG, "cutoff must be in [0, 1)");
  c->cutoff = cutoff;
  return TEBD_OK;
}

int tebd_get_bond_dims(tebd_ctx* c, int* chi) {
  if (!c || !chi) return TEBD_ERR_ARG;
  for (int l = 0; l <= c->L; ++l) chi[l] = c->chi[l];
  return TEBD_OK;
}

int tebd_get_stats(tebd_ctx* c, double* stats, int n) {
  if (!c || !stats || n < 0) return TEBD_ERR_ARG;
  for (int i = 0; i < n && i < kNumStats; ++i) stats[i] = c->stats[i];
  return TEBD_OK;
}

int tebd_set_profiling(tebd_ctx* c, int on) {
  if (!c) return TEBD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  drain_profile(c);
  c->profiling = on != 0;
  for (double& v : c->prof) v = 0.0;
  return TEBD_OK;
}

int tebd_get_profile(tebd_ctx* c, double* out, int n) {
  if (!c || !out || n < 0) return TEBD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  drain_profile(c);
  for (int i = 0; i < n && i < kNumProf; ++i) out[i] = c->prof[i];
  return TEBD_OK;
}

int tebd_set_graphs(tebd_ctx* c, int on) {
  if (!c) return TEBD_ERR_ARG;
  c->graphs = on != 0;
  return TEBD_OK;
}

int tebd_set_lanes(tebd_ctx* c, int n) {
  if (!c || n < 1 || n > kMaxLanes) return fail(c, TEBD_ERR_ARG, "set_lanes: 1 to 4 lanes");
  c->nlanes = n;
  return TEBD_OK;
}

int tebd_set_autotune(tebd_ctx* c, int on) {
  if (!c) return TEBD_ERR_ARG;
  c->autotune = on != 0;
  if (on < 0) c->tuned.clear();
  return TEBD_OK;
}

int tebd_autotune_info(tebd_ctx* c, double* out, int n) {
  if (!c || !out || n < 0) return TEBD_ERR_ARG;
  std::vector<double> v(2 + ZGEMM_VARIANTS, 0.0);
  v[0] = (double)c->tuned.size();
  v[1] = c->tune_trials;
  for (const auto& kv : c->tuned) v[2 + kv.second] += 1;
  for (int i = 0; i < n && i < (int)v.size(); ++i) out[i] = v[i];
  return TEBD_OK;
}

int tebd_construct_propagators(const double* H, int L, int d, int order, char bc, double dt_re, double dt_im,
                               double* U) {
  if (bc == 'o') bc = 'O';
  if (bc == 'p') bc = 'P';
  if (!H || !U || L < 2 || d < 2 || d > 8 || (order != 2 && order != 5) || (bc != 'O' && bc != 'P') ||
      (bc == 'P' && L < 3))
    return TEBD_ERR_ARG;
  construct_propagators_host(H, L, d * d, order, bc, cplx(dt_re, dt_im), U);
  return gate_table_size(L, order, bc);
}

int tebd_imag_time_prop(tebd_ctx* c, const double* H, int nh, double dtau, int chi_in, int max_steps,
                        int steps_for_judge, double conv1, double conv2, int* steps) {
  int e = ensure_state(c, false);
  if (e) return e;
  const int Lc = c->chain_len, n = c->d * c->d;
  const int per = c->bc == 'P' ? Lc : Lc - 1;  // bond Hamiltonians per chain
  if (!H || nh != c->nchain * per || !(dtau > 0.0) || chi_in < 1 || chi_in > c->chi_max || max_steps < 0 ||
      steps_for_judge < 1 || !(conv1 > 0.0) || !(conv2 > 0.0) || !steps)
    return fail(c, TEBD_ERR_ARG, "imag_time_prop: bad arguments");
  // Uitp = ConstructPropagators(H, -i dtau), chain by chain (propagation_module.f90:656-662)
  std::vector<double> U((size_t)2 * n * n * c->table * c->nchain);
  for (int r = 0; r < c->nchain; ++r)
    construct_propagators_host(H + (size_t)2 * n * n * per * r, Lc, n, c->order, c->bc, cplx(0.0, -dtau),
                               U.data() + (size_t)2 * n * n * c->table * r);
  if ((e = tebd_set_gates(c, U.data(), c->nchain * c->table))) return e;
  // stages chi_in and chi_max (one when they are equal), the tighter criterion on the last
  // (propagation_module.f90:663-690)
  std::vector<int> stages = {c->chi_max};
  if (chi_in != c->chi_max) stages = {chi_in, c->chi_max};
  for (int k = 0; k < 4; ++k) steps[k] = 0;
  std::vector<double> before, after;
  for (size_t k = 0; k < stages.size(); ++k) {
    const int chi = stages[k];
    const double conv = (chi_in == c->chi_max || chi == c->chi_max) ? conv2 : conv1;
    if ((e = set_stage(c, chi))) break;
    if ((e = largest_lambdas(c, before))) break;
    int j = 1;
    bool converged = false;
    for (; j <= max_steps; ++j) {
      if (j % steps_for_judge == 0) {  // the largest relative change of any bond's lambda_1 (:704-738)
        if ((e = largest_lambdas(c, after))) break;
        double gap = 0.0;
        for (size_t l = 0; l < after.size(); ++l) gap = std::max(gap, std::fabs((before[l] - after[l]) / before[l]));
        if (gap < conv) {
          converged = true;
          break;
        }
        before = after;
      }
      double tr = 0.0;  // TrotterStep, then CanonicalFormAll (:741-743)
      if ((e = tebd_trotter_step(c, &tr)) || (e = tebd_canonical_form_all(c))) break;
    }
    if (e) break;
    steps[2 * k] = converged ? j : max_steps;
    steps[2 * k + 1] = converged ? 1 : 0;
  }
  c->stage_chi = 0;
  return e;
}

int tebd_two_site_op_noalloc(int d, int chi_l, int chi_c, int chi_r, int chi_cap, const double* gamma_l,
                             const double* gamma_r, const double* lambda_l, const double* lambda_c,
                             const double* lambda_r, const double* op2, double* gamma_l_out, double* gamma_r_out,
                             double* lambda_c_out, int* chi_c_new, double* truncerr, void* stream) {
  if (!(d == 2 || d == 3 || d == 4 || d == 5 || d == 6 || d == 8)) return TEBD_ERR_ARG;
  if (chi_l < 1 || chi_c < 1 || chi_r < 1 || chi_cap < 1) return TEBD_ERR_ARG;
  if (d * chi_l > kNoAllocSide || d * chi_r > kNoAllocSide || chi_cap > kNoAllocSide || chi_c > kNoAllocSide)
    return TEBD_ERR_SHAPE;  // beyond the fixed workspace: use a context (tebd_two_site_op)
  if (!gamma_l || !gamma_r || !lambda_l || !lambda_c || !lambda_r || !op2 || !gamma_l_out || !gamma_r_out ||
      !lambda_c_out || !chi_c_new || !truncerr)
    return TEBD_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) return TEBD_ERR_HIP;
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TEBD_ERR_HIP;
    g_noalloc_ptr[dev] = p;
  }
  // host staging of the one task (pageable: no pinned allocation either)
  static BondTask h_tasks[kStages];
  static double h_results[RESULT_STRIDE], h_qstate[2];
  static int h_counters[4];
  tebd_ctx c;
  c.d = d;
  c.chi_max = chi_cap;
  c.Rmax = kNoAllocSide;
  c.Cpmax = kNoAllocSide;
  c.nblk_max = (int)kNoAllocNblk;
  c.ntask_max = 1;
  c.wl_cap = (int)(kNoAllocNblk / 2);
  c.graphs = false;
  c.st = (hipStream_t)stream;
  c.h_tasks = h_tasks;
  c.h_results = h_results;
  c.h_qstate = h_qstate;
  c.h_counters = h_counters;
  noalloc_layout(&c, (unsigned char*)g_noalloc_ptr[dev]);
  // the array keeps the previous call's contents: start from zero, as a fresh workspace
  hipLaunchKernelGGL(k_zero_words, dim3(256), dim3(256), 0, c.st, (unsigned long long*)g_noalloc_ptr[dev],
                     kNoAllocBytes / sizeof(unsigned long long));
  HIPCHK(&c, hipGetLastError());
  BondTask& t = h_tasks[0];
  std::memset(&t, 0, sizeof(t));
  t.chil = chi_l;
  t.chic = chi_c;
  t.chir = chi_r;
  t.d = d;
  const int m = d * chi_l, n = d * chi_r;
  t.transposed = n > m ? 1 : 0;
  t.R = std::max(m, n);
  t.C = std::min(m, n);
  t.Cpad = round_up(t.C, JC);
  t.P = t.Cpad / JC;
  t.gl = (const zc*)gamma_l;
  t.gr = (const zc*)gamma_r;
  t.ll = lambda_l;
  t.lc = lambda_c;
  t.lr = lambda_r;
  t.gate = (const zc*)op2;
  t.W = c.W;
  t.V = c.V;
  t.s = c.s;
  t.perm = c.perm;
  t.NV = t.Cpad;
  attach_jacobi(&c, t, 0);
  attach_qr(&c, t, 0);
  t.result = c.results;
  t.gl_out = (zc*)gamma_l_out;
  t.gr_out = (zc*)gamma_r_out;
  t.lc_out = lambda_c_out;
  t.chi_max = chi_cap;
  if (hipMemcpyAsync(c.d_tasks, h_tasks, sizeof(BondTask), hipMemcpyHostToDevice, c.st) != hipSuccess ||
      theta_gate(&c, c.d_tasks, 1, chi_l, chi_r) != hipSuccess) {
    (void)hipStreamSynchronize(c.st);
    return TEBD_ERR_HIP;
  }
  const int link = 0;
  int e = svd_truncate(&c, &link, 1, t.Cpad, t.P);
  if (e) {
    (void)hipStreamSynchronize(c.st);  // nothing of this call may still run when the next one starts
    return e;
  }
  *chi_c_new = (int)h_results[0];
  *truncerr = h_results[1];
  return TEBD_OK;
}

int tebd_checkpoint(tebd_ctx* c, const char* lambda_path, const char* gamma_path, char kind) {
  int e = ensure_state(c, false);
  if (e) return e;
  const int chi = c->chi_max;  // the reference records its arrays at their allocated size
  std::vector<double> g((size_t)2 * chi * c->d * chi * c->L), l((size_t)(c->L + 1) * chi);
  std::vector<double*> gp(c->L), lp(c->L + 1);
  for (int s = 0; s < c->L; ++s) gp[s] = g.data() + (size_t)2 * chi * c->d * chi * s;
  for (int b = 0; b <= c->L; ++b) lp[b] = l.data() + (size_t)chi * b;
  if ((e = download_impl(c, gp.data(), lp.data(), chi))) return e;
  e = tebd_write_state_files(c->L, c->d, chi, g.data(), l.data(), lambda_path, gamma_path, kind);
  return e ? fail(c, e, "checkpoint: cannot write the state files") : TEBD_OK;
}

int tebd_restart(tebd_ctx* c, const char* lambda_path, const char* gamma_path, char kind, int chi) {
  if (!c) return TEBD_ERR_ARG;
  if (chi < 1 || chi > c->chi_max) return fail(c, TEBD_ERR_SHAPE, "restart: file chi exceeds the context's chi_max");
  std::vector<double> g((size_t)2 * chi * c->d * chi * c->L), l((size_t)(c->L + 1) * chi);
  int e = tebd_read_state_files(c->L, c->d, chi, g.data(), l.data(), lambda_path, gamma_path, kind);
  if (e == TEBD_ERR_ARG) return fail(c, e, "restart: cannot open the state files (or bad kind)");
  if (e) return fail(c, e, "restart: the state files are shorter than L, d and chi require, or unreadable");
  std::vector<const double*> gp(c->L), lp(c->L + 1);
  for (int s = 0; s < c->L; ++s) gp[s] = g.data() + (size_t)2 * chi * c->d * chi * s;
  for (int b = 0; b <= c->L; ++b) lp[b] = l.data() + (size_t)chi * b;
  return upload_impl(c, gp.data(), lp.data(), chi);
}

int tebd_two_point_rdms(tebd_ctx* c, const double* phase, double* rho2, double* rho1) {
  int e = ensure_state(c, false);
  if (e) return e;
  if (!rho2) return fail(c, TEBD_ERR_ARG, "two_point_rdms: null output");
  const int d = c->d, Lc = c->chain_len, chi = c->chi_max, d2 = d * d;
  const size_t np = (size_t)Lc * (Lc - 1) / 2, d4 = (size_t)d2 * d2;
  if (!c->cr_T) {
    const size_t th = (size_t)chi * chi * d2, fac = (size_t)chi * d * chi;
    HIPCHK(c, hipMalloc(&c->cr_theta[0], sizeof(zc) * th));
    HIPCHK(c, hipMalloc(&c->cr_theta[1], sizeof(zc) * th));
    HIPCHK(c, hipMalloc(&c->cr_T, sizeof(zc) * th * d));
    HIPCHK(c, hipMalloc(&c->cr_K, sizeof(zc) * fac));
    HIPCHK(c, hipMalloc(&c->cr_B, sizeof(zc) * fac));
    HIPCHK(c, hipMalloc(&c->cr_A, sizeof(zc) * fac));
    HIPCHK(c, hipMalloc(&c->cr_s, sizeof(double) * chi * d));
    HIPCHK(c, hipMalloc(&c->cr_phase, sizeof(double) * d));
    HIPCHK(c, hipMalloc(&c->cr_rho2, sizeof(zc) * d4 * np * c->nchain));
    HIPCHK(c, hipMalloc(&c->cr_rho1, sizeof(zc) * d2 * c->L));
  }
  if (!c->h_gemm) {
    HIPCHK(c, hipMalloc(&c->d_gemm, sizeof(GemmTask) * c->ntask_max * kGemmRegions));
    HIPCHK(c, hipHostMalloc(&c->h_gemm, sizeof(GemmTask) * c->ntask_max * kGemmRegions));
  }
  const double* ph = nullptr;
  if (phase) {
    HIPCHK(c, hipMemcpy(c->cr_phase, phase, sizeof(double) * d, hipMemcpyHostToDevice));
    ph = c->cr_phase;
  }
  // one GEMM at a time; the pinned descriptors of a region are rewritten only after the
  // stream has consumed them
  int region = 0;
  auto gemm = [&](const zc* A, const zc* B, zc* C, int M, int N, int K, int lda, int ldb, int ldc, int acc) {
    if (region == 5) {
      HIPCHK(c, hipStreamSynchronize(c->st));
      region = 0;
    }
    GemmTask t = {A, B, C, M, N, K, lda, ldb, ldc, 1.0, acc};
    return gemm_batch(c, &t, 1, region++, false);
  };
  auto slab = [&](int s) { return c->gam + (size_t)s * c->gslab(); };
  auto lam = [&](int b) { return c->lam + (size_t)b * chi; };
  for (int r = 0; r < c->nchain; ++r) {
    const int base = r * Lc;
    zc* rho2_chain = c->cr_rho2 + d4 * np * r;
    for (int l2 = Lc - 1; l2 >= 0; --l2) {
      const int s2 = base + l2, chl = c->chi[s2], chr = c->chi[s2 + 1];
      // ThetaKernal of site l2 (observables_module.f90:704-730): ket factor (lambda Gamma
      // lambda as a (chi_l d) x chi_r matrix) times its bra counterpart
      HIPCHK(c, launch_scaled_copy(slab(s2), chl * d, c->cr_K, chl * d, chl * d, chr, lam(s2), chl, lam(s2 + 1), 1, 0,
                                   c->st));
      HIPCHK(c, launch_bra_kernal(slab(s2), lam(s2), lam(s2 + 1), chl, d, chr, c->cr_B, c->st));
      int cur = 0;
      if ((e = gemm(c->cr_K, c->cr_B, c->cr_theta[cur], chl * d, d * chl, chr, chl * d, chr, chl * d, 0))) return e;
      HIPCHK(c, launch_rho1_trace(c->cr_theta[cur], chl, d, c->cr_rho1 + (size_t)d2 * s2, c->st));
      for (int l1 = l2 - 1; l1 >= 0; --l1) {
        const int s1 = base + l1, ca = c->chi[s1], cb = c->chi[s1 + 1];
        // rows of the left site's ket Gamma scaled by lambda(a) phase(i)
        HIPCHK(c, launch_scale_vec(lam(s1), ph, ca, d, c->cr_s, c->st));
        // bra Gamma, conjugate-transposed with lambda(a): Gbar(eta, a + chi_a l)
        HIPCHK(c, launch_scaled_copy(slab(s1), ca * d, c->cr_B, cb, ca * d, cb, lam(s1), ca, nullptr, 1, 1, c->st));
        // T = Theta Gbar (ThetaNext's / TwoSiteRho's first contraction)
        if ((e = gemm(c->cr_theta[cur], c->cr_B, c->cr_T, cb * d2, ca * d, cb, cb * d2, cb, cb * d2, 0))) return e;
        const size_t pair = (size_t)l1 * Lc - (size_t)l1 * (l1 + 1) / 2 + (l2 - l1 - 1);
        HIPCHK(c, launch_rho_reduce(slab(s1), c->cr_s, c->cr_T, ca, cb, d, rho2_chain + d4 * pair, c->st));
        if (l1 == 0) break;  // no site further left: the environment is not needed
        // ket Gamma scaled by lambda(a) phase(k), then Theta' = sum_k A_k T_k
        HIPCHK(c, launch_scaled_copy(slab(s1), ca * d, c->cr_A, ca * d, ca * d, cb, c->cr_s, ca * d, nullptr, 1, 0,
                                     c->st));
        for (int k = 0; k < d; ++k)
          if ((e = gemm(c->cr_A + (size_t)ca * k, c->cr_T + (size_t)cb * d2 * ca * k, c->cr_theta[1 - cur], ca,
                        d2 * ca, cb, ca * d, cb, ca, k > 0)))
            return e;
        cur = 1 - cur;
      }
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(rho2, c->cr_rho2, sizeof(zc) * d4 * np * c->nchain, hipMemcpyDeviceToHost));
  if (rho1) HIPCHK(c, hipMemcpy(rho1, c->cr_rho1, sizeof(zc) * d2 * c->L, hipMemcpyDeviceToHost));
  return TEBD_OK;
}

int tebd_memory_info(tebd_ctx* c, double* out, int n) {
  if (!c || !out || n < 0) return TEBD_ERR_ARG;
  const double state = sizeof(zc) * (double)c->gslab() * c->L + 8.0 * (c->L + 1) * c->chi_max +
                       sizeof(zc) * (double)c->d4() * c->L;
  const double v[4] = {(double)c->ntask_max, slot_bytes(c), state, state + slot_bytes(c) * c->ntask_max};
  for (int i = 0; i < n && i < 4; ++i) out[i] = v[i];
  return TEBD_OK;
}

int tebd_sync(tebd_ctx* c) {
  if (!c) return TEBD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return TEBD_OK;
}

const char* tebd_last_error(const tebd_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"
