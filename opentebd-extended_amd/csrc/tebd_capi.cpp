// Host runtime behind the C ABI (include/tebd_hip.h): device-resident Vidal
// state, the 2nd-order Trotter schedule and the per-bond kernel pipeline
//   theta+gate -> one-sided block-Jacobi sweeps -> truncation -> Gamma rebuild.
#include "../../include/tebd_hip.h"
#include "tebd_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kMaxSweeps = 40;
constexpr double kTolRel = 1e-13;   // relative orthogonality of converged column pairs
constexpr double kTolAbs = 1e-14;   // rounding floor: |A_pq| <= kTolAbs ||theta|| (|w_p| + |w_q|)

int round_up(int x, int m) { return ((x + m - 1) / m) * m; }

}  // namespace

struct tebd_ctx {
  int L = 0, d = 0, chi_max = 0;
  unsigned mode = 0;
  double cutoff = 1e-30;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // device state
  zc* gam = nullptr;      // L slabs of chi_max*d*chi_max
  double* lam = nullptr;  // (L+1) x chi_max
  zc* gates = nullptr;    // (L-1) x d^4, plus one slot for tebd_two_site_op
  bool have_gates = false, have_state = false;
  std::vector<int> chi;   // bond dimensions, host mirror
  // per-task workspace
  int ntask_max = 0;
  int Rmax = 0, Cpmax = 0;
  zc* W = nullptr;
  zc* V = nullptr;
  double* s = nullptr;
  int* perm = nullptr;
  int* counters = nullptr;
  double* results = nullptr;
  BondTask* d_tasks = nullptr;
  BondTask* h_tasks = nullptr;   // pinned
  double* h_results = nullptr;   // pinned
  int* h_counters = nullptr;     // pinned
  std::vector<zc> staging;
  std::string err;
  double stats[4] = {0, 0, 0, 0};

  size_t gslab() const { return (size_t)chi_max * d * chi_max; }
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

// Process `nb` independent bonds (0-based links) as one batch of tasks.
int run_bonds(tebd_ctx* c, const int* links, int nb, const zc* const* gate_ptrs, double* trunc_sum) {
  const int d = c->d;
  int max_chil = 1, max_chir = 1, Pmax = 1, Cpmax = 1;
  for (int i = 0; i < nb; ++i) {
    const int link = links[i];
    BondTask& t = c->h_tasks[i];
    std::memset(&t, 0, sizeof(t));
    t.chil = c->chi[link];
    t.chic = c->chi[link + 1];
    t.chir = c->chi[link + 2];
    t.d = d;
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
    t.gate = gate_ptrs[i];
    t.W = c->W + (size_t)i * c->Rmax * c->Cpmax;
    t.V = c->V + (size_t)i * c->Cpmax * c->Cpmax;
    t.s = c->s + (size_t)i * c->Cpmax;
    t.perm = c->perm + (size_t)i * c->Cpmax;
    t.counter = c->counters + i;
    t.result = c->results + (size_t)4 * i;
    t.gl_out = const_cast<zc*>(t.gl);
    t.gr_out = const_cast<zc*>(t.gr);
    t.lc_out = const_cast<double*>(t.lc);
    t.chi_max = c->chi_max;
    max_chil = std::max(max_chil, t.chil);
    max_chir = std::max(max_chir, t.chir);
    Pmax = std::max(Pmax, t.P);
    Cpmax = std::max(Cpmax, t.Cpad);
  }
  HIPCHK(c, hipMemcpyAsync(c->d_tasks, c->h_tasks, sizeof(BondTask) * nb, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, launch_theta_gate(c->d_tasks, nb, d, max_chil, max_chir, c->st));
  HIPCHK(c, launch_prepare_svd(c->d_tasks, nb, c->st));
  HIPCHK(c, launch_frobenius(c->d_tasks, nb, Cpmax, c->st));
  int sweeps = 0;
  bool converged = false;
  for (; sweeps < kMaxSweeps; ++sweeps) {
    if (sweeps > 0) HIPCHK(c, launch_reset_counters(c->d_tasks, nb, c->st));
    for (int r = 0; r < 2 * Pmax - 1; ++r)
      HIPCHK(c, launch_jacobi_round(c->d_tasks, nb, Pmax, r, kTolRel, kTolAbs, c->st));
    HIPCHK(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(int) * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    int rot = 0;
    for (int i = 0; i < nb; ++i) rot += c->h_counters[i];
    if (rot == 0) { converged = true; ++sweeps; break; }
  }
  c->stats[1] += sweeps;
  if (!converged) return fail(c, TEBD_ERR_NOCONV, "one-sided Jacobi did not converge");
  HIPCHK(c, launch_finalize(c->d_tasks, nb, Cpmax, c->cutoff, c->st));
  HIPCHK(c, launch_write_gammas(c->d_tasks, nb, c->st));
  HIPCHK(c, hipMemcpyAsync(c->h_results, c->results, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  double tsum = 0.0;
  for (int i = 0; i < nb; ++i) {
    const double* r = c->h_results + 4 * i;
    const int k = (int)r[0];
    if (!(std::isfinite(r[2])) || r[2] <= 0.0 || k < 1) {
      char buf[160];
      std::snprintf(buf, sizeof(buf), "bond %d: theta is zero or not finite (||theta||^2=%g)", links[i] + 1, r[2]);
      return fail(c, TEBD_ERR_NAN, buf);
    }
    c->chi[links[i] + 1] = k;
    tsum += r[1];
  }
  c->stats[0] += nb;
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

void free_all(tebd_ctx* c) {
  hipFree(c->gam);
  hipFree(c->lam);
  hipFree(c->gates);
  hipFree(c->W);
  hipFree(c->V);
  hipFree(c->s);
  hipFree(c->perm);
  hipFree(c->counters);
  hipFree(c->results);
  hipFree(c->d_tasks);
  hipHostFree(c->h_tasks);
  hipHostFree(c->h_results);
  hipHostFree(c->h_counters);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->st) hipStreamDestroy(c->st);
}

}  // namespace

extern "C" {

int tebd_create(tebd_ctx** out, int L, int d, int chi_max, int order, char bc, unsigned mode) {
  if (!out) return TEBD_ERR_ARG;
  *out = nullptr;
  if (L < 2 || d < 2 || chi_max < 1) return TEBD_ERR_ARG;
  if (order != 2 || (bc != 'O' && bc != 'o')) return TEBD_ERR_ARG;
  if (!(d == 2 || d == 3 || d == 4 || d == 5 || d == 6 || d == 8)) return TEBD_ERR_ARG;
  if (d * chi_max > 4096) return TEBD_ERR_ARG;
  tebd_ctx* c = new tebd_ctx();
  c->L = L;
  c->d = d;
  c->chi_max = chi_max;
  c->mode = mode;
  c->chi.assign(L + 1, 1);
  c->Rmax = d * chi_max;
  c->Cpmax = round_up(d * chi_max, JC);
  c->ntask_max = (mode & TEBD_MODE_BATCHED) ? (L / 2) : 1;
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
  if (bad(hipMalloc(&c->gates, sizeof(zc) * c->d4() * L))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->W, sizeof(zc) * (size_t)c->Rmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->V, sizeof(zc) * (size_t)c->Cpmax * c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->s, sizeof(double) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->perm, sizeof(int) * (size_t)c->Cpmax * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->counters, sizeof(int) * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->results, sizeof(double) * 4 * nt))) return TEBD_ERR_HIP;
  if (bad(hipMalloc(&c->d_tasks, sizeof(BondTask) * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_tasks, sizeof(BondTask) * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_results, sizeof(double) * 4 * nt))) return TEBD_ERR_HIP;
  if (bad(hipHostMalloc(&c->h_counters, sizeof(int) * nt))) return TEBD_ERR_HIP;
  *out = c;
  return TEBD_OK;
}

void tebd_destroy(tebd_ctx* c) {
  if (!c) return;
  if (c->st) hipStreamSynchronize(c->st);
  free_all(c);
  delete c;
}

int tebd_set_gates(tebd_ctx* c, const double* U, int nb) {
  if (!c) return TEBD_ERR_ARG;
  if (!U || nb != c->L - 1) return fail(c, TEBD_ERR_ARG, "set_gates: need L-1 gates");
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

int tebd_trotter_step(tebd_ctx* c, double* truncerr) {
  int e = ensure_state(c, true);
  if (e) return e;
  const int L = c->L;
  double total = 0.0;
  HIPCHK(c, hipEventRecord(c->ev0, c->st));
  const int starts[3] = {0, 1, 0};  // reference bonds 1,3,5.. / 2,4,.. / 1,3,5..
  std::vector<int> links;
  std::vector<const zc*> gp;
  for (int half = 0; half < 3; ++half) {
    links.clear();
    gp.clear();
    for (int b = starts[half]; b < L - 1; b += 2) {
      links.push_back(b);
      gp.push_back(c->gates + (size_t)b * c->d4());
    }
    if (c->mode & TEBD_MODE_BATCHED) {
      double t = 0.0;
      e = run_bonds(c, links.data(), (int)links.size(), gp.data(), &t);
      if (e) return e;
      total += t;
    } else {
      for (size_t i = 0; i < links.size(); ++i) {
        double t = 0.0;
        e = run_bonds(c, &links[i], 1, &gp[i], &t);
        if (e) return e;
        total += t;
      }
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
  zc* slot = c->gates + (size_t)(c->L - 1) * c->d4();
  HIPCHK(c, hipMemcpyAsync(slot, op2, sizeof(zc) * c->d4(), hipMemcpyHostToDevice, c->st));
  const int l0 = link - 1;
  const zc* gp = slot;
  double t = 0.0;
  e = run_bonds(c, &l0, 1, &gp, &t);
  if (e) return e;
  if (truncerr) *truncerr = t;
  return TEBD_OK;
}

int tebd_canonical_form_all(tebd_ctx* c) {
  int e = ensure_state(c, false);
  if (e) return e;
  return fail(c, TEBD_ERR_UNSUPPORTED, "canonical_form_all: not available");
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
  for (int i = 0; i < n && i < 4; ++i) stats[i] = c->stats[i];
  return TEBD_OK;
}

int tebd_sync(tebd_ctx* c) {
  if (!c) return TEBD_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->st));
  return TEBD_OK;
}

const char* tebd_last_error(const tebd_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"
