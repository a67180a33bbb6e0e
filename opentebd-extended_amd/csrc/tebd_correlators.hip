// Kernels of the all-pairs two-site reduced density matrices (TwoSiteExpVal and
// EvaluateMeasures, observables_module.f90:836-883,2286-2526, built from ThetaKernal,
// ThetaNext and TwoSiteRho, :704-832).  The host (tebd_capi.cpp) carries, for a right
// site l2, the environment Theta(a, i, j, b) of l2 with the sites between absorbed, and
// steps it one site l1 to the left with two GEMMs; these kernels do the parts that are
// not GEMMs: the per-row scales lambda(a) phase(k) and the contraction of the left site
// into the two-site density matrix.
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

// s[a + chi k] = lambda[a] * phase[k] (phase null: 1), a < chi, k < d
__global__ void k_scale_vec(const double* __restrict__ lambda, const double* __restrict__ phase, int chi, int d,
                            double* __restrict__ s) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < chi * d; e += gridDim.x * blockDim.x) {
    const int a = e % chi, k = e / chi;
    s[e] = lambda[a] * (phase ? phase[k] : 1.0);
  }
}

// The bra factor of ThetaKernal (observables_module.f90:704-730) as a chi_r x (d chi_l) matrix
// with columns j + d b: out[e + chi_r (j + d b)] = conj(G(b, j, e)) lambda1[b] lambda2[e], so
// that Theta = K K^H-like product (ket factor) x out has Theta(a, i, j, b) at
// a + chi_l (i + d (j + d b)).
__global__ void k_bra_kernal(const zc* __restrict__ G, const double* __restrict__ lam1, const double* __restrict__ lam2,
                             int chi_l, int d, int chi_r, zc* __restrict__ out) {
  const size_t n = (size_t)chi_r * d * chi_l;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int eta = (int)(e % chi_r);
    const size_t rest = e / chi_r;
    const int j = (int)(rest % d), b = (int)(rest / d);
    const zc g = G[b + (size_t)chi_l * (j + (size_t)d * eta)];
    const double w = lam1[b] * lam2[eta];
    out[e] = zmk(g.x * w, -g.y * w);
  }
}

// TwoSiteRho's final contraction (observables_module.f90:815-826):
//   rho[(i d + j) + d^2 (l d + k)] = sum_{a < chi_a, b < chi_b} s[a + chi_a i] G(a, i, b) T(b, j, k; a + chi_a l)
// G: the left site's Gamma, element a + chi_a (i + d b); T: (chi_b d^2) x (chi_a d), element
// b + chi_b (j + d k) + chi_b d^2 (a + chi_a l) (the environment with the left site's bra
// Gamma absorbed).  One workgroup per entry of rho.
__global__ __launch_bounds__(256) void k_rho_reduce(const zc* __restrict__ G, const double* __restrict__ s,
                                                    const zc* __restrict__ T, int chi_a, int chi_b, int d,
                                                    zc* __restrict__ rho) {
  const int e = blockIdx.x;
  const int d2 = d * d;
  const int row = e % d2, col = e / d2;
  const int i = row / d, j = row % d, l = col / d, k = col % d;
  const size_t tcol0 = (size_t)chi_b * d2 * ((size_t)chi_a * l);
  double re = 0.0, im = 0.0;
  const int n = chi_a * chi_b;
  for (int q = threadIdx.x; q < n; q += 256) {
    const int a = q % chi_a, b = q / chi_a;
    const zc g = G[a + (size_t)chi_a * (i + (size_t)d * b)];
    const zc t = T[b + (size_t)chi_b * (j + (size_t)d * k) + tcol0 + (size_t)chi_b * d2 * a];
    const double w = s[a + (size_t)chi_a * i];
    re += w * (g.x * t.x - g.y * t.y);
    im += w * (g.x * t.y + g.y * t.x);
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      sr[threadIdx.x] += sr[threadIdx.x + st];
      si[threadIdx.x] += si[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) rho[e] = zmk(sr[0], si[0]);
}

// The one-site density matrix as the trace of ThetaKernal over its bond index
// (FormSingleSiteRho, observables_module.f90:95-117): rho[i + d j] = sum_a Theta(a, i, j, a),
// Theta element a + chi (i + d (j + d b)).
__global__ __launch_bounds__(256) void k_rho1_trace(const zc* __restrict__ Theta, int chi, int d, zc* __restrict__ rho) {
  const int i = blockIdx.x % d, j = blockIdx.x / d;
  double re = 0.0, im = 0.0;
  for (int a = threadIdx.x; a < chi; a += 256) {
    const zc t = Theta[a + (size_t)chi * (i + (size_t)d * (j + (size_t)d * a))];
    re += t.x;
    im += t.y;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      sr[threadIdx.x] += sr[threadIdx.x + st];
      si[threadIdx.x] += si[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) rho[blockIdx.x] = zmk(sr[0], si[0]);
}

}  // namespace

hipError_t launch_scale_vec(const double* lambda, const double* phase, int chi, int d, double* s, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_vec, dim3((chi * d + 255) / 256), dim3(256), 0, st, lambda, phase, chi, d, s);
  return hipGetLastError();
}

hipError_t launch_bra_kernal(const zc* G, const double* lam1, const double* lam2, int chi_l, int d, int chi_r, zc* out,
                             hipStream_t st) {
  const size_t n = (size_t)chi_r * d * chi_l;
  const int blocks = (int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(k_bra_kernal, dim3(blocks), dim3(256), 0, st, G, lam1, lam2, chi_l, d, chi_r, out);
  return hipGetLastError();
}

hipError_t launch_rho_reduce(const zc* G, const double* s, const zc* T, int chi_a, int chi_b, int d, zc* rho,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_rho_reduce, dim3(d * d * d * d), dim3(256), 0, st, G, s, T, chi_a, chi_b, d, rho);
  return hipGetLastError();
}

hipError_t launch_rho1_trace(const zc* Theta, int chi, int d, zc* rho, hipStream_t st) {
  hipLaunchKernelGGL(k_rho1_trace, dim3(d * d), dim3(256), 0, st, Theta, chi, d, rho);
  return hipGetLastError();
}
