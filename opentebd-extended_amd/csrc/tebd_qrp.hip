// Column-pivoted Householder QR preconditioning of the bond SVD.
//
// The bond matrix W (Rb x Cb, Rb >= Cb: theta, or theta^dagger) is factorised as
// W P = Q R with column pivoting (LAPACK xGEQP3's pivot rule: the column of largest
// remaining norm first) and stopped at the numerical rank r, where every remaining
// column norm is below tol_rel * ||W||_F.  The one-sided Jacobi then runs on
// X = R(0:r, :)^dagger (Cb x r), whose columns are graded and nearly orthogonal:
// on late-time Fermi-Hubbard bonds (1024 x 1024, rank ~830) it converges in 4
// sweeps instead of ~22 for W itself, and the sweeps cover r instead of Cb columns.
// With X V = U_x S:  W = (Q V) S (P U_x)^dagger, so the Rb-side singular vectors
// are Q V (k_qrp_vectors applies the reflectors) and the Cb-side ones P U_x.
//
//   k_qrp_begin    column norms, ||W||_F, identity pivot, rank := running
//   k_qrp_pivot    step j: argmax of the remaining norms, column swap, Householder
//                  reflector of column j (xLARFG convention, real diagonal R_jj)
//   k_qrp_update   step j: apply H_j^dagger to the trailing columns, exact new norms
//   k_qrp_build_x  X = R(0:r, :)^dagger, zero padding columns
//   k_qrp_vectors  Rb-side vectors Q V(:, kept) and Cb-side vectors P X(:, kept)/s
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Fixed-order sum over the 256 threads of a workgroup (bitwise reproducible).
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_qrp_begin(const BondTask* __restrict__ tasks, double tol_rel) {
  const BondTask t = tasks[blockIdx.x];
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = wave; k < t.Cb; k += 4) {
    const zc* col = t.Wq + (size_t)t.Rb * k;
    double s = 0.0;
    for (int i = lane; i < t.Rb; i += 64) s += zabs2(col[i]);
    s = wave_sum(s);
    if (lane == 0) t.qnrm[k] = s;
  }
  for (int k = threadIdx.x; k < t.Cb; k += 256) t.qpiv[k] = k;
  __syncthreads();
  double part = 0.0;
  for (int k = threadIdx.x; k < t.Cb; k += 256) part += t.qnrm[k];
  const double tot = block_sum(part, red);
  if (threadIdx.x == 0) {
    t.qstate[0] = -1.0;                  // rank: running
    t.qstate[1] = tol_rel * tol_rel * tot;  // stop when max remaining |col|^2 <= this
  }
}

__global__ __launch_bounds__(256) void k_qrp_pivot(const BondTask* __restrict__ tasks, int j) {
  const BondTask t = tasks[blockIdx.x];
  if (t.qstate[0] >= 0.0) return;
  const int kmax = t.Rb < t.Cb ? t.Rb : t.Cb;
  if (j >= kmax) {
    if (threadIdx.x == 0) t.qstate[0] = (double)kmax;
    return;
  }
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ double red[4];
  // argmax of the remaining norms (ties: lowest index)
  double bv = -1.0;
  int bi = j;
  for (int k = j + threadIdx.x; k < t.Cb; k += 256) {
    const double v = t.qnrm[k];
    if (v > bv) { bv = v; bi = k; }
  }
  rv[threadIdx.x] = bv;
  ri[threadIdx.x] = bi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      const double o = rv[threadIdx.x + st];
      const int oi = ri[threadIdx.x + st];
      if (o > rv[threadIdx.x] || (o == rv[threadIdx.x] && oi < ri[threadIdx.x])) {
        rv[threadIdx.x] = o;
        ri[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const double vmax = rv[0];
  const int p = ri[0];
  if (!(vmax > t.qstate[1])) {
    if (threadIdx.x == 0) t.qstate[0] = (double)j;  // numerical rank reached
    return;
  }
  zc* cj = t.Wq + (size_t)t.Rb * j;
  if (p != j) {
    zc* cp = t.Wq + (size_t)t.Rb * p;
    for (int i = threadIdx.x; i < t.Rb; i += 256) {
      const zc a = cj[i];
      cj[i] = cp[i];
      cp[i] = a;
    }
    if (threadIdx.x == 0) {
      const double tn = t.qnrm[j]; t.qnrm[j] = t.qnrm[p]; t.qnrm[p] = tn;
      const int tp = t.qpiv[j]; t.qpiv[j] = t.qpiv[p]; t.qpiv[p] = tp;
    }
  }
  __syncthreads();
  // reflector H = I - tau v v^H with v_j = 1 mapping x = W(j:, j) to (beta, 0, ..)
  double part = 0.0;
  for (int i = j + 1 + threadIdx.x; i < t.Rb; i += 256) part += zabs2(cj[i]);
  const double xn2 = block_sum(part, red);
  const zc alpha = cj[j];
  if (xn2 == 0.0 && alpha.y == 0.0) {
    if (threadIdx.x == 0) t.tau[j] = zmk(0.0, 0.0);
    return;
  }
  const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
  const double beta = alpha.x >= 0.0 ? -nrm : nrm;
  const zc tau = zmk((beta - alpha.x) / beta, -alpha.y / beta);
  // scale = 1 / (alpha - beta)
  const zc am = zmk(alpha.x - beta, alpha.y);
  const double den = zabs2(am);
  const zc scale = zmk(am.x / den, -am.y / den);
  __syncthreads();  // every thread has read alpha before it is overwritten
  for (int i = j + 1 + threadIdx.x; i < t.Rb; i += 256) cj[i] = zmul(cj[i], scale);
  if (threadIdx.x == 0) {
    cj[j] = zmk(beta, 0.0);
    t.tau[j] = tau;
  }
}

constexpr int kQrCols = 16;  // columns per workgroup in the trailing update

// W(j:, k) <- (I - conj(tau) v v^H) W(j:, k) for the workgroup's columns k > j; one
// wavefront per column; the exact partial norm |W(j+1:, k)|^2 replaces qnrm[k].
__global__ __launch_bounds__(256) void k_qrp_update(const BondTask* __restrict__ tasks, int j) {
  const BondTask t = tasks[blockIdx.y];
  if (t.qstate[0] >= 0.0) return;
  const int k0 = blockIdx.x * kQrCols;
  if (k0 + kQrCols <= j + 1 || k0 >= t.Cb) return;
  const zc tau = t.tau[j];
  if (tau.x == 0.0 && tau.y == 0.0) {
    // H = I: only the norms move down one row
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = k0 + wave; k < k0 + kQrCols && k < t.Cb; k += 4) {
      if (k <= j) continue;
      if (lane == 0) {
        const zc a = t.Wq[(size_t)t.Rb * k + j];
        t.qnrm[k] = fmax(t.qnrm[k] - zabs2(a), 0.0);
      }
    }
    return;
  }
  const zc ctau = zconj(tau);
  const zc* v = t.Wq + (size_t)t.Rb * j;  // v_j = 1 (the stored beta is not read)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = k0 + wave; k < k0 + kQrCols && k < t.Cb; k += 4) {
    if (k <= j) continue;
    zc* a = t.Wq + (size_t)t.Rb * k;
    zc w = zmk(0.0, 0.0);
    for (int i = j + lane; i < t.Rb; i += 64) {
      const zc vi = (i == j) ? zmk(1.0, 0.0) : v[i];
      zfmac(w, vi, a[i]);
    }
    w.x = wave_sum(w.x);
    w.y = wave_sum(w.y);
    const zc f = zmul(ctau, w);
    double s = 0.0;
    for (int i = j + lane; i < t.Rb; i += 64) {
      const zc vi = (i == j) ? zmk(1.0, 0.0) : v[i];
      const zc ai = zsub(a[i], zmul(f, vi));
      a[i] = ai;
      if (i > j) s += zabs2(ai);
    }
    s = wave_sum(s);
    if (lane == 0) t.qnrm[k] = s;
  }
}

// X (Cb x Cpad, ld Cb) = R(0:r, :)^dagger; columns r..Cpad-1 zero.  R is the upper
// trapezoid of Wq.
__global__ void k_qrp_build_x(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const int r = t.C;  // the Jacobi task's column count is the numerical rank
  const size_t n = (size_t)t.Cb * t.Cpad;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(e % t.Cb), col = (int)(e / t.Cb);  // X(row, col) = conj(R(col, row))
    zc x = zmk(0.0, 0.0);
    if (col < r && col <= row) x = zconj(t.Wq[col + (size_t)t.Rb * row]);
    t.W[e] = x;
  }
}

// Rb-side vectors Y(:, b) = Q [V(:, perm[b]); 0] and Cb-side vectors
// Z(qpiv[i], b) = X(i, perm[b]) / s[b] for the kept b < k.  One workgroup per
// NB columns b; thread t owns rows t + 256 q (q < QR) of its columns in registers
// while the reflectors H_{r-1} .. H_0 are applied.
template <int QR, int NB>
__global__ __launch_bounds__(256) void k_qrp_vectors(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const int k = (int)t.result[0];
  const int b0 = blockIdx.x * NB;
  if (b0 >= k || t.Rb > 256 * QR) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = t.C;
  // Cb-side vectors
  for (int bb = 0; bb < NB && b0 + bb < k; ++bb) {
    const int b = b0 + bb, col = t.perm[b];
    const double inv = t.s[b] > 0.0 ? 1.0 / t.s[b] : 0.0;
    for (int i = tid; i < t.Cb; i += 256)
      t.Zc[t.qpiv[i] + (size_t)t.Cb * b] = zscale(t.W[i + (size_t)t.Cb * col], inv);
  }
  // Rb-side vectors
  zc y[QR][NB];
#pragma unroll
  for (int q = 0; q < QR; ++q)
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const int i = tid + 256 * q, b = b0 + bb;
      y[q][bb] = (i < r && b < k) ? t.V[i + (size_t)t.NV * t.perm[b]] : zmk(0.0, 0.0);
    }
  __shared__ double part[2][4][NB][2];
  for (int jj = r - 1; jj >= 0; --jj) {
    const zc tau = t.tau[jj];
    const zc* v = t.Wq + (size_t)t.Rb * jj;
    zc w[NB];
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) w[bb] = zmk(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < QR; ++q) {
      const int i = tid + 256 * q;
      if (i >= jj && i < t.Rb) {
        const zc vi = (i == jj) ? zmk(1.0, 0.0) : v[i];
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) zfmac(w[bb], vi, y[q][bb]);
      }
    }
    const int buf = jj & 1;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const double wr = wave_sum(w[bb].x), wi = wave_sum(w[bb].y);
      if (lane == 0) {
        part[buf][wave][bb][0] = wr;
        part[buf][wave][bb][1] = wi;
      }
    }
    __syncthreads();
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const zc ws = zmk(part[buf][0][bb][0] + part[buf][1][bb][0] + part[buf][2][bb][0] + part[buf][3][bb][0],
                        part[buf][0][bb][1] + part[buf][1][bb][1] + part[buf][2][bb][1] + part[buf][3][bb][1]);
      w[bb] = zmul(tau, ws);
    }
#pragma unroll
    for (int q = 0; q < QR; ++q) {
      const int i = tid + 256 * q;
      if (i >= jj && i < t.Rb) {
        const zc vi = (i == jj) ? zmk(1.0, 0.0) : v[i];
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) y[q][bb] = zsub(y[q][bb], zmul(w[bb], vi));
      }
    }
    // the other buffer is rewritten two reflectors later, after the next barrier
  }
#pragma unroll
  for (int q = 0; q < QR; ++q)
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const int i = tid + 256 * q, b = b0 + bb;
      if (i < t.Rb && b < k) t.Yr[i + (size_t)t.Rb * b] = y[q][bb];
    }
}

}  // namespace

hipError_t launch_qrp_begin(const BondTask* d_tasks, int ntasks, double tol_rel, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_begin, dim3(ntasks), dim3(256), 0, st, d_tasks, tol_rel);
  return hipGetLastError();
}

hipError_t launch_qrp_step(const BondTask* d_tasks, int ntasks, int max_cb, int j, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_pivot, dim3(ntasks), dim3(256), 0, st, d_tasks, j);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || j >= max_cb - 1) return e;
  hipLaunchKernelGGL(k_qrp_update, dim3((max_cb + kQrCols - 1) / kQrCols, ntasks), dim3(256), 0, st, d_tasks, j);
  return hipGetLastError();
}

hipError_t launch_qrp_build_x(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_build_x, dim3(256, ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

hipError_t launch_qrp_vectors(const BondTask* d_tasks, int ntasks, int max_rb, int max_k, hipStream_t st) {
  auto go = [&](auto kern, int nb) {
    hipLaunchKernelGGL(kern, dim3((max_k + nb - 1) / nb, ntasks), dim3(256), 0, st, d_tasks);
    return hipGetLastError();
  };
  if (max_rb <= 256) return go(k_qrp_vectors<1, 16>, 16);
  if (max_rb <= 512) return go(k_qrp_vectors<2, 16>, 16);
  if (max_rb <= 1024) return go(k_qrp_vectors<4, 8>, 8);
  if (max_rb <= 2048) return go(k_qrp_vectors<8, 4>, 4);
  if (max_rb <= 4096) return go(k_qrp_vectors<16, 2>, 2);
  return hipErrorInvalidValue;
}
