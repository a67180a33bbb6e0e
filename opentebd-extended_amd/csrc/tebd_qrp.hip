// Column-pivoted Householder QR preconditioning of the bond SVD.
//
// The bond matrix W (Rb x Cb, Rb >= Cb: theta, or theta^dagger) is factorised as
// W P = Q R with block column pivoting and stopped at the numerical rank r, where
// every remaining column norm is below tol_rel * ||W||_F.  Each panel of QB columns
// takes the QB columns of largest remaining norm (exact norms, recomputed by every
// trailing update) in descending order; the panel itself is factorised without
// further pivoting and applied to the rest as one compact-WY block (BLAS3).  On
// late-time Fermi-Hubbard bonds this ordering gives the Jacobi the same few sweeps
// as column-by-column pivoting (xGEQP3) at a 32-fold smaller trailing traffic.  The one-sided Jacobi then runs on
// X = R(0:r, :)^dagger (Cb x r), whose columns are graded and nearly orthogonal:
// on late-time Fermi-Hubbard bonds (1024 x 1024, rank ~830) it converges in 4
// sweeps instead of ~22 for W itself, and the sweeps cover r instead of Cb columns.
// With X V = U_x S:  W = (Q V) S (P U_x)^dagger, so the Rb-side singular vectors
// are Q V (k_qrp_vectors applies the reflectors) and the Cb-side ones P U_x.
//
//   k_qrp_begin    column norms, ||W||_F, identity pivot, rank := running
//   k_qrp_panel    panel j0: the QB largest remaining columns, descending, swapped into
//                  place; their unpivoted Householder QR; the compact-WY factor T
//   k_qrp_trail    A2 <- Q_panel^H A2 for the columns right of the panel (MFMA), exact
//                  remaining norms for the next panel's choice
//   k_qrp_build_x  X = R(0:r, :)^dagger, zero padding columns
//   k_qrp_vectors  Rb-side vectors Q V(:, kept) and Cb-side vectors P X(:, kept)/s, or the
//                  explicit Q of a factorisation (the sketched SVD's range basis)
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

constexpr int QB = 32;          // panel width
constexpr int kMaxCols = 4096;  // largest Cb (d chi_max <= 4096)

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

typedef double d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16).  Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// element g of the result in lane l is D[l/16 + 4g][l%16].
__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Per-wave partial sums of D = L^H R (32 x 32 as four 16x16 blocks) over rows
// r < K: wave w takes rows 8 (w + 4 it) .. +7 in two MFMA k-steps of 4 rows; lload(i, c)
// and rload(i, c) return element (i, c) of L and R (zero outside).
template <class LL, class RL>
__device__ __forceinline__ void gram32_partial(LL lload, RL rload, int K, int wave, int lane, d4 (&dre)[2][2],
                                               d4 (&dim)[2][2]) {
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      dre[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      dim[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  for (int r0 = 8 * wave; r0 < K; r0 += 32) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = r0 + 4 * h + lk;
      const zc l[2] = {lload(i, lr), lload(i, 16 + lr)};
      const zc r[2] = {rload(i, lr), rload(i, 16 + lr)};
#pragma unroll
      for (int pb = 0; pb < 2; ++pb)
#pragma unroll
        for (int qb = 0; qb < 2; ++qb) {
          dre[pb][qb] = mfma4(l[pb].x, r[qb].x, dre[pb][qb]);
          dre[pb][qb] = mfma4(l[pb].y, r[qb].y, dre[pb][qb]);
          dim[pb][qb] = mfma4(l[pb].x, r[qb].y, dim[pb][qb]);
          dim[pb][qb] = mfma4(-l[pb].y, r[qb].x, dim[pb][qb]);
        }
    }
  }
}

// Fixed-order sum of the four waves' partials into S[32][33] (bitwise reproducible).
__device__ __forceinline__ void gram32_reduce(const d4 (&dre)[2][2], const d4 (&dim)[2][2], double (*red)[2][256],
                                              zc (*S)[QB + 1], int wave, int lane) {
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int pb = 0; pb < 2; ++pb)
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
      if (wave > 0) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          red[wave - 1][0][g * 64 + lane] = dre[pb][qb][g];
          red[wave - 1][1][g * 64 + lane] = dim[pb][qb][g];
        }
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          double xr = dre[pb][qb][g], xi = dim[pb][qb][g];
          for (int w = 0; w < 3; ++w) {
            xr += red[w][0][g * 64 + lane];
            xi += red[w][1][g * 64 + lane];
          }
          S[16 * pb + lk + 4 * g][16 * qb + lr] = zmk(xr, xi);
        }
      }
      __syncthreads();
    }
}

// Panel j0 (one workgroup per task):
//  1. the nbp = min(QB, kmax - j0) columns of largest remaining norm, in descending
//     order, are swapped into columns j0 .. j0+nbp-1 (if even the largest is below the
//     stop threshold the numerical rank is j0 and the factorisation ends);
//  2. unpivoted Householder QR of the panel rows j0.. (xLARFG convention);
//  3. T (nbp x nbp, upper) of the compact WY form H_0 .. H_{nbp-1} = I - V T V^H.
__global__ __launch_bounds__(256) void k_qrp_panel(const BondTask* __restrict__ tasks, int j0) {
  const BondTask t = tasks[blockIdx.x];
  if (t.qstate[0] >= 0.0) return;
  const int kmax = t.Rb < t.Cb ? t.Rb : t.Cb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (j0 >= kmax) {
    if (tid == 0) t.qstate[0] = (double)kmax;
    return;
  }
  const int nbp = (kmax - j0) < QB ? (kmax - j0) : QB;
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ double red[4];
  __shared__ int sel[QB];
  __shared__ unsigned char taken[kMaxCols];
  __shared__ zc G[QB][QB + 1];
  __shared__ zc T[QB][QB + 1];
  __shared__ zc stau[QB];
  __shared__ double gred[3][2][256];
  for (int k = tid; k < t.Cb; k += 256) taken[k] = 0;
  __syncthreads();
  // 1. selection: nbp rounds of argmax (ties: lowest index)
  for (int q = 0; q < nbp; ++q) {
    double bv = -1.0;
    int bi = -1;
    for (int k = j0 + tid; k < t.Cb; k += 256) {
      const double v = t.qnrm[k];
      if (!taken[k] && v > bv) { bv = v; bi = k; }
    }
    rv[tid] = bv;
    ri[tid] = bi;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st) {
        const double o = rv[tid + st];
        const int oi = ri[tid + st];
        if (oi >= 0 && (o > rv[tid] || ri[tid] < 0 || (o == rv[tid] && oi < ri[tid]))) {
          rv[tid] = o;
          ri[tid] = oi;
        }
      }
      __syncthreads();
    }
    if (q == 0 && !(rv[0] > t.qstate[1])) {
      if (tid == 0) t.qstate[0] = (double)j0;  // numerical rank reached
      return;
    }
    if (tid == 0) {
      sel[q] = ri[0];
      taken[ri[0]] = 1;
    }
    __syncthreads();
  }
  // swap the selected columns into place, in order
  for (int q = 0; q < nbp; ++q) {
    const int p = sel[q], dst = j0 + q;
    if (p != dst) {
      zc* a = t.Wq + (size_t)t.Rb * dst;
      zc* b = t.Wq + (size_t)t.Rb * p;
      for (int i = tid; i < t.Rb; i += 256) {
        const zc x = a[i];
        a[i] = b[i];
        b[i] = x;
      }
    }
    __syncthreads();
    if (tid == 0 && p != dst) {
      const double tn = t.qnrm[dst]; t.qnrm[dst] = t.qnrm[p]; t.qnrm[p] = tn;
      const int tp = t.qpiv[dst]; t.qpiv[dst] = t.qpiv[p]; t.qpiv[p] = tp;
      for (int q2 = q + 1; q2 < nbp; ++q2)
        if (sel[q2] == dst) sel[q2] = p;
    }
    __syncthreads();
  }
  // 2. Householder QR of the panel
  for (int q = 0; q < nbp; ++q) {
    const int j = j0 + q;
    zc* cj = t.Wq + (size_t)t.Rb * j;
    double part = 0.0;
    for (int i = j + 1 + tid; i < t.Rb; i += 256) part += zabs2(cj[i]);
    const double xn2 = block_sum(part, red);
    const zc alpha = cj[j];
    zc tau = zmk(0.0, 0.0);
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      const double beta = alpha.x >= 0.0 ? -nrm : nrm;
      tau = zmk((beta - alpha.x) / beta, -alpha.y / beta);
      const zc am = zmk(alpha.x - beta, alpha.y);
      const double den = zabs2(am);
      const zc scale = zmk(am.x / den, -am.y / den);
      __syncthreads();  // every thread has read alpha
      for (int i = j + 1 + tid; i < t.Rb; i += 256) cj[i] = zmul(cj[i], scale);
      if (tid == 0) cj[j] = zmk(beta, 0.0);
    }
    if (tid == 0) {
      t.tau[j] = tau;
      stau[q] = tau;
    }
    __syncthreads();
    if (tau.x != 0.0 || tau.y != 0.0) {
      const zc ctau = zconj(tau);
      for (int k = j + 1 + wave; k < j0 + nbp; k += 4) {
        zc* a = t.Wq + (size_t)t.Rb * k;
        zc w = zmk(0.0, 0.0);
        for (int i = j + lane; i < t.Rb; i += 64) zfmac(w, (i == j) ? zmk(1.0, 0.0) : cj[i], a[i]);
        w.x = wave_sum(w.x);
        w.y = wave_sum(w.y);
        const zc f = zmul(ctau, w);
        for (int i = j + lane; i < t.Rb; i += 64) a[i] = zsub(a[i], zmul(f, (i == j) ? zmk(1.0, 0.0) : cj[i]));
      }
    }
    __syncthreads();
  }
  // 3. T: T[q][q] = tau_q, T[0:q, q] = -tau_q T[0:q, 0:q] (V^H v_q)[0:q]
  {
    const int K = t.Rb - j0;
    const zc* P = t.Wq + (size_t)t.Rb * j0 + j0;
    auto vload = [&](int i, int q) -> zc {
      if (q >= nbp || i < q || i >= K) return zmk(0.0, 0.0);
      if (i == q) return zmk(1.0, 0.0);
      return P[i + (size_t)t.Rb * q];
    };
    d4 gre[2][2], gim[2][2];
    gram32_partial(vload, vload, K, wave, lane, gre, gim);
    gram32_reduce(gre, gim, gred, G, wave, lane);
  }
  if (tid == 0) {
    for (int q = 0; q < nbp; ++q) {
      for (int pp = 0; pp < q; ++pp) {
        zc acc = zmk(0.0, 0.0);
        for (int kk = pp; kk < q; ++kk) zfma(acc, T[pp][kk], G[kk][q]);
        T[pp][q] = zmul(zmk(-stau[q].x, -stau[q].y), acc);
      }
      T[q][q] = stau[q];
      for (int pp = q + 1; pp < QB; ++pp) T[pp][q] = zmk(0.0, 0.0);
    }
  }
  __syncthreads();
  for (int e = tid; e < QB * QB; e += 256) {
    const int pp = e % QB, qq = e / QB;
    t.Tq[e] = (pp < nbp && qq < nbp) ? T[pp][qq] : zmk(0.0, 0.0);
  }
  if (tid == 0 && j0 + nbp >= kmax) t.qstate[0] = (double)kmax;
}

// Trailing update of panel j0 on 32 columns per workgroup:
//   A2 <- Q_panel^H A2 = A2 - V (T^H (V^H A2))
// with Y = V^H A2 and A2 -= V Z on FP64 MFMA (V: rows j0.., implicit unit diagonal),
// and the exact remaining norms |A2(j0+nbp:, c)|^2 into qnrm.
__global__ __launch_bounds__(256) void k_qrp_trail(const BondTask* __restrict__ tasks, int j0) {
  const BondTask t = tasks[blockIdx.y];
  const double rank = t.qstate[0];
  if (rank >= 0.0 && rank <= (double)j0) return;
  const int kmax = t.Rb < t.Cb ? t.Rb : t.Cb;
  if (j0 >= kmax) return;
  const int nbp = (kmax - j0) < QB ? (kmax - j0) : QB;
  const int kb = j0 + nbp + 32 * blockIdx.x;
  if (kb >= t.Cb) return;
  const int nc = (t.Cb - kb) < 32 ? (t.Cb - kb) : 32;
  const int K = t.Rb - j0;
  const size_t ld = t.Rb;
  const zc* P = t.Wq + ld * j0 + j0;  // P[i + ld q] = A(j0 + i, j0 + q)
  zc* A2 = t.Wq + ld * kb + j0;       // A2[i + ld c] = A(j0 + i, kb + c)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  auto vload = [&](int i, int q) -> zc {
    if (q >= nbp || i < q || i >= K) return zmk(0.0, 0.0);
    if (i == q) return zmk(1.0, 0.0);
    return P[i + ld * q];
  };
  auto aload = [&](int i, int c) -> zc {
    if (c >= nc || i >= K) return zmk(0.0, 0.0);
    return A2[i + ld * c];
  };
  __shared__ double red[3][2][256];
  __shared__ zc Y[QB][QB + 1];
  __shared__ zc Ts[QB][QB + 1];
  __shared__ double nsum[4][32];
  for (int e = tid; e < QB * QB; e += 256) Ts[e % QB][e / QB] = t.Tq[e];
  // ---- Y = V^H A2 (32 x 32), rows split over the waves
  {
    d4 yre[2][2], yim[2][2];
    gram32_partial(vload, aload, K, wave, lane, yre, yim);
    gram32_reduce(yre, yim, red, Y, wave, lane);
  }
  // ---- Z = T^H Y, held as MFMA A operands Z[4ks + lk][16ct + lr]
  double zre[2][8], zim[2][8];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int pr = 4 * ks + lk, qc = 16 * ct + lr;
      zc acc = zmk(0.0, 0.0);
      for (int kk = 0; kk <= pr; ++kk) zfmac(acc, Ts[kk][pr], Y[kk][qc]);
      zre[ct][ks] = acc.x;
      zim[ct][ks] = acc.y;
    }
  // ---- A2 -= V Z, as (V Z)^T = Z^T V^T: A operand Z^T, B operand V^T
  double ns[2][4];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int g = 0; g < 4; ++g) ns[ct][g] = 0.0;
  for (int r0 = 16 * wave; r0 < K; r0 += 64) {
    const int i = r0 + lr;
    zc x[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) x[ks] = vload(i, 4 * ks + lk);
    d4 ore[2], oim[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      ore[ct] = d4{0.0, 0.0, 0.0, 0.0};
      oim[ct] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        ore[ct] = mfma4(zre[ct][ks], x[ks].x, ore[ct]);
        ore[ct] = mfma4(-zim[ct][ks], x[ks].y, ore[ct]);
        oim[ct] = mfma4(zre[ct][ks], x[ks].y, oim[ct]);
        oim[ct] = mfma4(zim[ct][ks], x[ks].x, oim[ct]);
      }
    }
    if (i < K) {
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int c = 16 * ct + lk + 4 * g;
          if (c < nc) {
            zc* ap = A2 + i + ld * c;
            const zc a = zsub(*ap, zmk(ore[ct][g], oim[ct][g]));
            *ap = a;
            if (i >= nbp) ns[ct][g] += zabs2(a);
          }
        }
    }
  }
  // column norms: reduce over the 16 lanes of equal lk, then over the waves
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double v = ns[ct][g];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      if (lr == 0) nsum[wave][16 * ct + lk + 4 * g] = v;
    }
  __syncthreads();
  if (tid < nc) t.qnrm[kb + tid] = nsum[0][tid] + nsum[1][tid] + nsum[2][tid] + nsum[3][tid];
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
// Z(qpiv[i], b) = X(i, perm[b]) / s[b] for the kept b < k (FORM_Q = false), or the
// explicit orthonormal factor Qs(:, b) = Q e_b, b < r (FORM_Q = true).  One workgroup
// per NB columns b; thread t owns rows t + 256 q (q < QR) of its columns in
// registers while the reflectors H_{r-1} .. H_0 are applied.
template <int QR, int NB, bool FORM_Q>
__global__ __launch_bounds__(256) void k_qrp_vectors(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const int r = t.C;
  const int k = FORM_Q ? r : (int)t.result[0];
  const int b0 = blockIdx.x * NB;
  if (b0 >= k || t.Rb > 256 * QR) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (!FORM_Q) {
    // Cb-side vectors
    for (int bb = 0; bb < NB && b0 + bb < k; ++bb) {
      const int b = b0 + bb, col = t.perm[b];
      const double inv = t.s[b] > 0.0 ? 1.0 / t.s[b] : 0.0;
      for (int i = tid; i < t.Cb; i += 256)
        t.Zc[t.qpiv[i] + (size_t)t.Cb * b] = zscale(t.W[i + (size_t)t.Cb * col], inv);
    }
  }
  // Rb-side vectors
  zc y[QR][NB];
#pragma unroll
  for (int q = 0; q < QR; ++q)
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      const int i = tid + 256 * q, b = b0 + bb;
      if (FORM_Q) y[q][bb] = zmk((i == b && b < k) ? 1.0 : 0.0, 0.0);
      else y[q][bb] = (i < r && b < k) ? t.V[i + (size_t)t.NV * t.perm[b]] : zmk(0.0, 0.0);
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
      if (i < t.Rb && b < k) (FORM_Q ? t.Qs : t.Yr)[i + (size_t)t.Rb * b] = y[q][bb];
    }
}

}  // namespace

hipError_t launch_qrp_begin(const BondTask* d_tasks, int ntasks, double tol_rel, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_begin, dim3(ntasks), dim3(256), 0, st, d_tasks, tol_rel);
  return hipGetLastError();
}

hipError_t launch_qrp_panel(const BondTask* d_tasks, int ntasks, int max_cb, int j0, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_panel, dim3(ntasks), dim3(256), 0, st, d_tasks, j0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || j0 + QB >= max_cb) return e;
  hipLaunchKernelGGL(k_qrp_trail, dim3((max_cb - j0 - 1) / 32 + 1, ntasks), dim3(256), 0, st, d_tasks, j0);
  return hipGetLastError();
}

hipError_t launch_qrp_build_x(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_build_x, dim3(256, ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

template <bool FORM_Q>
hipError_t launch_vectors(const BondTask* d_tasks, int ntasks, int max_rb, int max_k, hipStream_t st) {
  auto go = [&](auto kern, int nb) {
    hipLaunchKernelGGL(kern, dim3((max_k + nb - 1) / nb, ntasks), dim3(256), 0, st, d_tasks);
    return hipGetLastError();
  };
  if (max_rb <= 256) return go(k_qrp_vectors<1, 16, FORM_Q>, 16);
  if (max_rb <= 512) return go(k_qrp_vectors<2, 16, FORM_Q>, 16);
  if (max_rb <= 1024) return go(k_qrp_vectors<4, 8, FORM_Q>, 8);
  if (max_rb <= 2048) return go(k_qrp_vectors<8, 4, FORM_Q>, 4);
  if (max_rb <= 4096) return go(k_qrp_vectors<16, 2, FORM_Q>, 2);
  return hipErrorInvalidValue;
}

hipError_t launch_qrp_vectors(const BondTask* d_tasks, int ntasks, int max_rb, int max_k, hipStream_t st) {
  return launch_vectors<false>(d_tasks, ntasks, max_rb, max_k, st);
}

hipError_t launch_qrp_form_q(const BondTask* d_tasks, int ntasks, int max_rb, int max_r, hipStream_t st) {
  return launch_vectors<true>(d_tasks, ntasks, max_rb, max_r, st);
}
