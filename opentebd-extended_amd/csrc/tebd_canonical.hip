// Small kernels of the device CanonicalForm (local_operations_module.f90:496-620):
// scaled copies, column permutations of SVD factors, and the X / Y factors.
// The heavy lifting reuses the batched GEMM and the one-sided Jacobi SVD.
#include "tebd_device.h"
#include "tebd_kernels.h"

#include <algorithm>

namespace {

// Out[r + ldo c] = rs(r) cs(c) In[r + ldi c]            (conj_t = 0)
// Out[c + ldo r] = rs(r) cs(c) conj(In[r + ldi c])      (conj_t = 1)
// rs(r) = rowscale[r % row_mod] (1 if null); cs(c) = colscale[c / col_div] (1 if null)
__global__ void k_scaled_copy(const zc* __restrict__ in, int ldi, zc* __restrict__ out, int ldo, int rows, int cols,
                              const double* __restrict__ rowscale, int row_mod, const double* __restrict__ colscale,
                              int col_div, int conj_t) {
  const size_t total = (size_t)rows * cols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % rows), c = (int)(e / rows);
    double f = 1.0;
    if (rowscale) f *= rowscale[r % row_mod];
    if (colscale) f *= colscale[c / col_div];
    const zc v = zscale(in[r + (size_t)ldi * c], f);
    if (!conj_t) out[r + (size_t)ldo * c] = v;
    else out[c + (size_t)ldo * r] = zconj(v);
  }
}

// Dst[:, b] = Src[:, perm[b]] * g(s[b]) for b < ncols, with
//   mode 0: g = 1                 (right factor of a Jacobi SVD)
//   mode 1: g = 1/s               (left factor W/s of a Jacobi SVD)
//   mode 2: g = 1/s if s^2 > thr, else 0   (U diag(1/sqrt(S)) with S = s^2, masked)
__global__ void k_perm_cols(const zc* __restrict__ src, int lds, zc* __restrict__ dst, int ldd, int rows, int ncols,
                            const int* __restrict__ perm, const double* __restrict__ s, int mode, double thr) {
  const size_t total = (size_t)rows * ncols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % rows), b = (int)(e / rows);
    double g = 1.0;
    if (mode == 1) g = s[b] > 0.0 ? 1.0 / s[b] : 0.0;
    if (mode == 2) g = (s[b] * s[b] > thr) ? 1.0 / s[b] : 0.0;
    dst[r + (size_t)ldd * b] = zscale(src[r + (size_t)lds * perm[b]], g);
  }
}

// X = diag(s1) U1^dagger ; Y = diag(lam1) U2 diag(s2)   (all n x n)
__global__ void k_cf_xy(const zc* __restrict__ U1, const double* __restrict__ s1, const zc* __restrict__ U2,
                        const double* __restrict__ s2, const double* __restrict__ lam1, zc* __restrict__ X,
                        zc* __restrict__ Y, int n) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(e % n), b = (int)(e / n);
    X[e] = zscale(zconj(U1[b + (size_t)n * a]), s1[a]);
    Y[e] = zscale(U2[e], lam1[a] * s2[b]);
  }
}

// lambda[a] = s[a] / sqrt(sum s^2) for a < n, 0 for n <= a < chi_max
__global__ void k_cf_lambda(const double* __restrict__ s, int n, double* __restrict__ lam, int chi_max) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int a = threadIdx.x; a < n; a += 256) acc += s[a] * s[a];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  const double inv = red[0] > 0.0 ? 1.0 / sqrt(red[0]) : 0.0;
  for (int a = threadIdx.x; a < chi_max; a += 256) lam[a] = a < n ? s[a] * inv : 0.0;
}

// Gamma(a, i, b) of a site slab cut to a < ncl, b < ncr (CopyGamLam,
// TEBDtools_module.f90:523-553): src holds the chi_l x d x chi_r slab, dst receives the
// ncl x d x ncr one, both packed with their own leading dimensions.
__global__ void k_shrink_slab(const zc* __restrict__ src, int cl, int d, zc* __restrict__ dst, int ncl, int ncr) {
  const size_t n = (size_t)ncl * d * ncr;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int a = (int)(e % ncl);
    const size_t rest = e / ncl;
    const int i = (int)(rest % d);
    const int b = (int)(rest / d);
    dst[e] = src[a + (size_t)cl * (i + (size_t)d * b)];
  }
}

// lambda[l][a] = 0 for chi <= a < chi_max, every bond l < nbonds
__global__ void k_zero_lambda_tail(double* __restrict__ lam, int nbonds, int chi, int chi_max) {
  const int w = chi_max - chi;
  const size_t n = (size_t)nbonds * w;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    lam[(e / w) * chi_max + chi + (e % w)] = 0.0;
}

}  // namespace

hipError_t launch_scaled_copy(const zc* in, int ldi, zc* out, int ldo, int rows, int cols, const double* rowscale,
                              int row_mod, const double* colscale, int col_div, int conj_t, hipStream_t st) {
  hipLaunchKernelGGL(k_scaled_copy, dim3(256), dim3(256), 0, st, in, ldi, out, ldo, rows, cols, rowscale,
                     row_mod > 0 ? row_mod : 1, colscale, col_div > 0 ? col_div : 1, conj_t);
  return hipGetLastError();
}

hipError_t launch_perm_cols(const zc* src, int lds, zc* dst, int ldd, int rows, int ncols, const int* perm,
                            const double* s, int mode, double thr, hipStream_t st) {
  hipLaunchKernelGGL(k_perm_cols, dim3(256), dim3(256), 0, st, src, lds, dst, ldd, rows, ncols, perm, s, mode, thr);
  return hipGetLastError();
}

hipError_t launch_cf_xy(const zc* U1, const double* s1, const zc* U2, const double* s2, const double* lam1, zc* X,
                        zc* Y, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_cf_xy, dim3(256), dim3(256), 0, st, U1, s1, U2, s2, lam1, X, Y, n);
  return hipGetLastError();
}

hipError_t launch_cf_lambda(const double* s, int n, double* lam, int chi_max, hipStream_t st) {
  hipLaunchKernelGGL(k_cf_lambda, dim3(1), dim3(256), 0, st, s, n, lam, chi_max);
  return hipGetLastError();
}

hipError_t launch_shrink_slab(const zc* src, int cl, int d, zc* dst, int ncl, int ncr, hipStream_t st) {
  const size_t n = (size_t)ncl * d * ncr;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_shrink_slab, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, src, cl, d, dst, ncl, ncr);
  return hipGetLastError();
}

hipError_t launch_zero_lambda_tail(double* lam, int nbonds, int chi, int chi_max, hipStream_t st) {
  if (chi >= chi_max) return hipSuccess;
  const size_t n = (size_t)nbonds * (chi_max - chi);
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero_lambda_tail, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, lam, nbonds, chi, chi_max);
  return hipGetLastError();
}
