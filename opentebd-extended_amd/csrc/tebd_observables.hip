// Device expectation values of two-site operators.
//
// For a state in canonical Vidal form the reduced density matrix of sites (l, l+1)
// is rho = theta theta^dagger with theta = ll Gamma_l lc Gamma_r lr, so
//   <psi| O_{l,l+1} |psi> = sum conj(theta) (O theta)
// (TwoSiteRho / TotalEnergy, observables_module.f90:704-832,1432-1466; one-site
// expectation values are the two-site ones of op x 1).  theta and O theta are
// formed by the theta+gate kernel with the gate replaced by 1 and by O; k_zdotc
// reduces their inner product in a fixed order.
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

// out[2 i], out[2 i + 1] = sum_e conj(A_i.W[e]) B_i.W[e] over the R x C elements of
// task i (both column-major with leading dimension R).  One workgroup per task.
__global__ __launch_bounds__(256) void k_zdotc(const BondTask* __restrict__ ta, const BondTask* __restrict__ tb,
                                               double* __restrict__ out) {
  const BondTask a = ta[blockIdx.x], b = tb[blockIdx.x];
  __shared__ double rre[256], rim[256];
  const size_t n = (size_t)a.R * a.C;
  double sr = 0.0, si = 0.0;
  for (size_t e = threadIdx.x; e < n; e += 256) {
    const zc x = a.W[e], y = b.W[e];
    sr += x.x * y.x + x.y * y.y;
    si += x.x * y.y - x.y * y.x;
  }
  rre[threadIdx.x] = sr;
  rim[threadIdx.x] = si;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      rre[threadIdx.x] += rre[threadIdx.x + st];
      rim[threadIdx.x] += rim[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = rre[0];
    out[2 * blockIdx.x + 1] = rim[0];
  }
}

}  // namespace

hipError_t launch_zdotc(const BondTask* d_a, const BondTask* d_b, int ntasks, double* d_out, hipStream_t st) {
  hipLaunchKernelGGL(k_zdotc, dim3(ntasks), dim3(256), 0, st, d_a, d_b, d_out);
  return hipGetLastError();
}
