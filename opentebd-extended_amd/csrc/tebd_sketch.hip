// gfx950 kernels of the sketched (randomised) SVD used by the learned-rank path, and
// the batched complex GEMM it and CanonicalForm use.
//
// The reference always takes the full SVD of theta (ZGESVD,
// local_operations_module.f90:322-345) and keeps chi values.  The sketched path
// computes only the leading singular triplets from the range of theta Omega, with
//   Omega = [ (Gamma_r lambda_r)^dagger | Gaussian ]   the right Schmidt vectors of
//           the pre-gate theta plus random directions for rank growth,
// whose width is the predicted rank plus oversampling.
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

// ---------------------------------------------------------------------------
// batched complex GEMM, 64 x 64 tiles, 4 x 4 complex register micro-tiles
// ---------------------------------------------------------------------------
constexpr int GT = 64;
constexpr int GK = 16;

template <bool CONJ_A>
__global__ __launch_bounds__(256) void k_zgemm(const GemmTask* __restrict__ tasks) {
  const GemmTask t = tasks[blockIdx.z];
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  if (m0 >= t.M || n0 >= t.N) return;
  __shared__ zc As[GK][GT + 1];
  __shared__ zc Bs[GK][GT + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  zc acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = zmk(0.0, 0.0);
  for (int k0 = 0; k0 < t.K; k0 += GK) {
    for (int e = tid; e < GK * GT; e += 256) {
      int kk, mm;
      zc v = zmk(0.0, 0.0);
      if (!CONJ_A) {  // A is M x K: contiguous along m
        mm = e % GT; kk = e / GT;
        const int m = m0 + mm, k = k0 + kk;
        if (m < t.M && k < t.K) v = t.A[m + (size_t)t.lda * k];
      } else {        // A is K x M: contiguous along k
        kk = e % GK; mm = e / GK;
        const int m = m0 + mm, k = k0 + kk;
        if (m < t.M && k < t.K) v = zconj(t.A[k + (size_t)t.lda * m]);
      }
      As[kk][mm] = v;
    }
    for (int e = tid; e < GK * GT; e += 256) {
      const int kk = e % GK, nn = e / GK;
      const int k = k0 + kk, n = n0 + nn;
      Bs[kk][nn] = (k < t.K && n < t.N) ? t.B[k + (size_t)t.ldb * n] : zmk(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      zc a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) zfma(acc[i][j], a[i], b[j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
      if (m < t.M && n < t.N) {
        zc v = zscale(acc[i][j], t.alpha);
        if (t.accumulate) v = zadd(v, t.C[m + (size_t)t.ldc * n]);
        t.C[m + (size_t)t.ldc * n] = v;
      }
    }
}

// ---------------------------------------------------------------------------
// batched complex GEMM on the FP64 matrix cores: TM x TN tiles, 2 NWR waves as NWR x 2
// (NWR = 2, or 4 for 128 x 96 tiles), each wave (TM/NWR) x (TN/2) in blocks of 16 x 16,
// complex products by the three-multiplication form (three real v_mfma_f64_16x16x4f64
// per k-step instead of four): T1 = sum Ar Br, T2 = sum Ai Bi, T3 = sum (Ar + Ai)(Br + Bi),
// Re C = T1 - T2, Im C = T3 - T1 - T2.  It is normwise as accurate as the four-product
// form (its bound on Im C is (|Ar| + |Ai|)(|Br| + |Bi|) eps per term instead of
// (|Ar||Bi| + |Ai||Br|) eps, at most twice as large; Higham, Accuracy and Stability of
// Numerical Algorithms, 2nd ed., sec. 23.2.4).  Operands staged through LDS in chunks of
// KCH = 16 (or 32: half the barriers per product and twice the time for the next
// chunk's loads to arrive, for twice the LDS).  Every tile shape and chunk length takes
// the same operations in the same order per element.
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16): lane l supplies A[l%16][l/16] and B[l/16][l%16];
// element g of the result in lane l is D[l/16 + 4g][l%16].
__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int MK = 16;  // k chunk staged in LDS

template <bool CONJ_A, int TM, int TN, int KCH, int NWR>
__global__ __launch_bounds__(128 * NWR) void k_zgemm_mfma(const GemmTask* __restrict__ tasks) {
  constexpr int NT = 128 * NWR;                  // NWR x 2 waves
  constexpr int RB = TM / (16 * NWR), CB = TN / 32;
  static_assert(TM % (16 * NWR) == 0 && TN % 32 == 0 && (KCH * TM) % NT == 0 && (KCH * TN) % NT == 0,
                "tile shape and thread count must match");
  const GemmTask t = tasks[blockIdx.z];
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  if (m0 >= t.M || n0 >= t.N) return;
  __shared__ zc As[KCH][TM + 1];
  __shared__ zc Bs[KCH][TN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;  // wave tile (TM / NWR) x (TN / 2)
  d4 t1[RB][CB], t2[RB][CB], t3[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      t1[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      t2[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      t3[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
  // chunk k0 of op(A) and B is loaded into registers while the previous one is
  // multiplied from LDS: element e = tid + NT u of the chunk
  constexpr int NA = KCH * TM / NT, NB = KCH * TN / NT;
  zc pa[NA], pb[NB];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = tid + NT * u;
      zc v = zmk(0.0, 0.0);
      if (!CONJ_A) {  // A is M x K: contiguous along m
        const int m = m0 + e % TM, k = k0 + e / TM;
        if (m < t.M && k < t.K) v = t.A[m + (size_t)t.lda * k];
      } else {        // A is K x M: contiguous along k
        const int m = m0 + e / KCH, k = k0 + e % KCH;
        if (m < t.M && k < t.K) v = zconj(t.A[k + (size_t)t.lda * m]);
      }
      pa[u] = v;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int e = tid + NT * u;
      const int k = k0 + e % KCH, n = n0 + e / KCH;
      pb[u] = (k < t.K && n < t.N) ? t.B[k + (size_t)t.ldb * n] : zmk(0.0, 0.0);
    }
  };
  load_chunk(0);
  for (int k0 = 0; k0 < t.K; k0 += KCH) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = tid + NT * u;
      if (!CONJ_A) As[e / TM][e % TM] = pa[u];
      else As[e % KCH][e / KCH] = pa[u];
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int e = tid + NT * u;
      Bs[e % KCH][e / KCH] = pb[u];
    }
    __syncthreads();
    if (k0 + KCH < t.K) load_chunk(k0 + KCH);
#pragma unroll 4
    for (int ks = 0; ks < KCH / 4; ++ks) {
      const int kk = 4 * ks + lk;
      zc a[RB], b[CB];
      double sa[RB], sb[CB];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        a[i] = As[kk][wr * (TM / NWR) + 16 * i + lr];
        sa[i] = a[i].x + a[i].y;
      }
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        b[j] = Bs[kk][wc * (TN / 2) + 16 * j + lr];
        sb[j] = b[j].x + b[j].y;
      }
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          t1[i][j] = mfma4(a[i].x, b[j].x, t1[i][j]);
          t2[i][j] = mfma4(a[i].y, b[j].y, t2[i][j]);
          t3[i][j] = mfma4(sa[i], sb[j], t3[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < CB; ++j)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int m = m0 + wr * (TM / NWR) + 16 * i + lk + 4 * g, n = n0 + wc * (TN / 2) + 16 * j + lr;
        if (m < t.M && n < t.N) {
          const double re = t1[i][j][g] - t2[i][j][g], im = t3[i][j][g] - t1[i][j][g] - t2[i][j][g];
          zc v = zmk(re * t.alpha, im * t.alpha);
          if (t.accumulate) v = zadd(v, t.C[m + (size_t)t.ldc * n]);
          t.C[m + (size_t)t.ldc * n] = v;
        }
      }
}

// ---------------------------------------------------------------------------
// sketch construction
// ---------------------------------------------------------------------------
// 64-bit mix (splitmix64 finaliser) -> uniform (0,1]
__device__ __forceinline__ double u01(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return ((double)(x >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

// Z0[(j, b), c] = conj(Gamma_r(c, j, b) lambda_r(b)) for c < chi_c, Gaussian for
// chi_c <= c < k', zero for k' <= c < kpad.  Z0 has n = d chi_r rows (theta's
// column index (j, b) -> j chi_r + b).
// The Gaussian columns are scaled by kSketchNoiseScale (a power of two, so exactly):
// the range of Y = theta Z0 is unchanged, but the pivoted QR of Y now takes the columns
// of the old Schmidt vectors first, in the order of their weights, and the random ones
// last.  Q then starts with the old left Schmidt directions and B^H = theta^H Q has nearly
// orthogonal columns for the Jacobi SVD that follows.  (Unscaled, a random column is
// heavier than any Schmidt column and the pivoting took those first, mixing every
// direction into every column of Q.)  Householder QR errs column by column relative to
// each column's own norm, so the scale costs no accuracy; the scaled columns stay far
// above the QR's rank threshold of 1e-17 ||Y||_F unless theta has no weight outside the
// old Schmidt directions, where they are rightly dropped.
constexpr double kSketchNoiseScale = 1.0 / 65536.0;

__global__ void k_build_sketch(const BondTask* __restrict__ tasks, const int* __restrict__ kprime,
                               unsigned long long seed) {
  const BondTask t = tasks[blockIdx.y];
  const int n = t.d * t.chir;
  const int kp = kprime[blockIdx.y];
  const size_t total = (size_t)n * t.Cpad;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(e % n), c = (int)(e / n);
    const int j = row / t.chir, b = row % t.chir;
    zc v = zmk(0.0, 0.0);
    if (c < t.chic && c < kp) {
      v = zconj(zscale(t.gr[c + (size_t)t.chic * (j + (size_t)t.d * b)], t.lr[b]));
    } else if (c < kp) {
      const unsigned long long key = seed ^ ((unsigned long long)t.rng_key << 48) ^ ((unsigned long long)e << 1);
      const double u1 = u01(key), u2 = u01(key + 0x632BE59BD9B4E019ull);
      const double r = kSketchNoiseScale * sqrt(-2.0 * log(u1));
      v = zmk(r * cos(6.283185307179586 * u2), r * sin(6.283185307179586 * u2));
    }
    t.V[e] = v;
  }
}

// Standard complex Gaussian (E|g|^2 = 1) number e of a probe stream.
__device__ __forceinline__ zc probe_gauss(unsigned long long seed, int key, unsigned long long e) {
  const unsigned long long x = seed ^ 0xA5A5F00DCAFEull ^ ((unsigned long long)key << 48) ^ (e << 1);
  const double u1 = u01(x), u2 = u01(x + 0x632BE59BD9B4E019ull);
  const double r = sqrt(-log(u1));  // |g|^2 ~ Exp(1)
  return zmk(r * cos(6.283185307179586 * u2), r * sin(6.283185307179586 * u2));
}

// Probes of the sketch check: C1 = G1, C2 = S G1 (k x p) and G2 into columns p..2p-1 of
// Om (n x p), from the bond's own stream, so that the check of a bond does not depend on
// its batch slot.
__global__ void k_probe_coeffs(const ProbeTask* __restrict__ tasks, unsigned long long seed) {
  const ProbeTask t = tasks[blockIdx.y];
  const int n1 = t.k * kProbes, n2 = t.n * kProbes;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n1 + n2; e += gridDim.x * blockDim.x) {
    const zc g = probe_gauss(seed, t.rng_key, (unsigned long long)e);
    if (e < n1) {
      const double sj = t.s[e % t.k];
      t.C1[e] = g;
      t.C2[e] = zmk(sj * g.x, sj * g.y);
    } else {
      t.Om[(size_t)t.n * kProbes + (e - n1)] = g;
    }
  }
}

// out[0] = ||Z1||_F^2, out[1] = ||Z2||_F^2 (Z = [Z1 | Z2], m x 2p), one workgroup per task,
// fixed-order reductions.
__global__ __launch_bounds__(256) void k_probe_norm(const ProbeTask* __restrict__ tasks) {
  const ProbeTask t = tasks[blockIdx.x];
  const size_t n = (size_t)t.m * kProbes;
  __shared__ double red[2][256];
  double a1 = 0.0, a2 = 0.0;
  for (size_t e = threadIdx.x; e < n; e += 256) {
    const zc z1 = t.Z[e], z2 = t.Z[n + e];
    a1 += z1.x * z1.x + z1.y * z1.y;
    a2 += z2.x * z2.x + z2.y * z2.y;
  }
  red[0][threadIdx.x] = a1;
  red[1][threadIdx.x] = a2;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      red[0][threadIdx.x] += red[0][threadIdx.x + st];
      red[1][threadIdx.x] += red[1][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    t.out[0] = red[0][0];
    t.out[1] = red[1][0];
  }
}

}  // namespace

hipError_t launch_probe_coeffs(const ProbeTask* d_tasks, int ntasks, int max_kn, unsigned long long seed,
                               hipStream_t st) {
  const int blocks = (max_kn * kProbes + 255) / 256;
  hipLaunchKernelGGL(k_probe_coeffs, dim3(blocks, ntasks), dim3(256), 0, st, d_tasks, seed);
  return hipGetLastError();
}

hipError_t launch_probe_norm(const ProbeTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_norm, dim3(ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

hipError_t launch_zgemm(const GemmTask* d_tasks, int ntasks, int max_m, int max_n, bool conj_a, hipStream_t st,
                        int variant) {
  auto go = [&](auto kern, int tm, int tn, int threads = 256) {
    dim3 grid((max_m + tm - 1) / tm, (max_n + tn - 1) / tn, ntasks);
    hipLaunchKernelGGL(kern, grid, dim3(threads), 0, st, d_tasks);
    return hipGetLastError();
  };
  switch (variant) {
    case ZGEMM_VALU_64x64:
      return conj_a ? go(k_zgemm<true>, GT, GT) : go(k_zgemm<false>, GT, GT);
    case ZGEMM_MFMA_64x64:
      return conj_a ? go(k_zgemm_mfma<true, 64, 64, MK, 2>, 64, 64) : go(k_zgemm_mfma<false, 64, 64, MK, 2>, 64, 64);
    case ZGEMM_MFMA_128x64:
      return conj_a ? go(k_zgemm_mfma<true, 128, 64, MK, 2>, 128, 64)
                    : go(k_zgemm_mfma<false, 128, 64, MK, 2>, 128, 64);
    case ZGEMM_MFMA_64x96:
      return conj_a ? go(k_zgemm_mfma<true, 64, 96, MK, 2>, 64, 96) : go(k_zgemm_mfma<false, 64, 96, MK, 2>, 64, 96);
    case ZGEMM_MFMA_64x64_K32:
      return conj_a ? go(k_zgemm_mfma<true, 64, 64, 2 * MK, 2>, 64, 64)
                    : go(k_zgemm_mfma<false, 64, 64, 2 * MK, 2>, 64, 64);
    case ZGEMM_MFMA_128x96_W8:
      return conj_a ? go(k_zgemm_mfma<true, 128, 96, MK, 4>, 128, 96, 512)
                    : go(k_zgemm_mfma<false, 128, 96, MK, 4>, 128, 96, 512);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_build_sketch(const BondTask* d_tasks, const int* d_kprime, int ntasks, unsigned long long seed,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_build_sketch, dim3(256, ntasks), dim3(256), 0, st, d_tasks, d_kprime, seed);
  return hipGetLastError();
}


