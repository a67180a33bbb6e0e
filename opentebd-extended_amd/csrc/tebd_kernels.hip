// gfx950 kernels of the two-site update (FP64).
//
//   k_theta_gate   FormTheta + ThetaOperation_c (local_operations_module.f90:184-211,248-280)
//                  fused: theta'(a,i,j,b) = ll(a) lr(b) sum_kl U(ij,kl) sum_g Gl(a,k,g) lc(g) Gr(g,l,b)
//                  written straight into the SVD work matrix W (theta, or theta^dagger when
//                  the bond has more columns than rows).
//   k_col_norms    squared column norms of W.
//   k_truncate     sort descending, choose the kept rank, lambda and truncerr
//                  (FormLambda1, local_operations_module.f90:347-379).
//   k_write_gammas FormGamma1 / FormGamma2 (local_operations_module.f90:381-428).
#include "tebd_device.h"
#include "tebd_kernels.h"

#include <cfloat>

namespace {

// ---------------------------------------------------------------------------
// theta + gate
// ---------------------------------------------------------------------------
constexpr int KC = 16;  // gamma (contraction) chunk staged in LDS

template <int D>
struct ThetaCfg {
  static constexpr int BA = (D <= 5) ? 16 : 8;  // alpha block == beta block
  static constexpr int TR = BA * D;             // tile rows == tile cols: (k, alpha) pairs
  static constexpr int MT = TR / 16;            // per-thread micro tile edge
};

template <int D>
__global__ __launch_bounds__(256) void k_theta_gate(const BondTask* __restrict__ tasks) {
  using Cfg = ThetaCfg<D>;
  constexpr int BA = Cfg::BA, TR = Cfg::TR, MT = Cfg::MT, D2 = D * D;
  const BondTask t = tasks[blockIdx.z];
  const int a0 = blockIdx.x * BA, b0 = blockIdx.y * BA;
  if (a0 >= t.chil || b0 >= t.chir) return;

  constexpr int AB_ELEMS = 2 * KC * (TR + 1);
  constexpr int T_ELEMS = TR * (TR + 1);
  constexpr int SM_ELEMS = AB_ELEMS > T_ELEMS ? AB_ELEMS : T_ELEMS;
  __shared__ zc sm[SM_ELEMS];
  __shared__ zc sgate[D2 * D2];
  zc* As = sm;                     // [KC][TR+1]  A(r, g) = Gl(a,k,g) lc(g), r = k*BA + a_local
  zc* Bs = sm + KC * (TR + 1);     // [KC][TR+1]  B(g, c) = Gr(g,l,b),       c = l*BA + b_local
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;

  for (int e = tid; e < D2 * D2; e += 256) sgate[e] = t.gate[e];

  zc acc[MT][MT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) acc[i][j] = zmk(0.0, 0.0);

  const int chil = t.chil, chic = t.chic, chir = t.chir;
  for (int g0 = 0; g0 < chic; g0 += KC) {
    // A chunk: r fastest (contiguous alpha within one k)
    for (int e = tid; e < KC * TR; e += 256) {
      const int r = e % TR, g = e / TR;
      const int k = r / BA, a = a0 + (r % BA), gg = g0 + g;
      zc v = zmk(0.0, 0.0);
      if (a < chil && gg < chic) v = zscale(t.gl[a + (size_t)chil * (k + (size_t)D * gg)], t.lc[gg]);
      As[g * (TR + 1) + r] = v;
    }
    // B chunk: g fastest (contiguous gamma)
    for (int e = tid; e < KC * TR; e += 256) {
      const int g = e % KC, c = e / KC;
      const int l = c / BA, b = b0 + (c % BA), gg = g0 + g;
      zc v = zmk(0.0, 0.0);
      if (b < chir && gg < chic) v = t.gr[gg + (size_t)chic * (l + (size_t)D * b)];
      Bs[g * (TR + 1) + c] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int g = 0; g < KC; ++g) {
      zc av[MT], bv[MT];
#pragma unroll
      for (int i = 0; i < MT; ++i) av[i] = As[g * (TR + 1) + ty + 16 * i];
#pragma unroll
      for (int j = 0; j < MT; ++j) bv[j] = Bs[g * (TR + 1) + tx + 16 * j];
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) zfma(acc[i][j], av[i], bv[j]);
    }
    __syncthreads();
  }
  // stage T(r, c) in LDS for the gate
  zc* T = sm;
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) T[(ty + 16 * i) * (TR + 1) + tx + 16 * j] = acc[i][j];
  __syncthreads();

  // gate: one (alpha, beta) pair per thread
  for (int pr = tid; pr < BA * BA; pr += 256) {
    int al, bl;
    if (!t.transposed) { al = pr % BA; bl = pr / BA; } else { bl = pr % BA; al = pr / BA; }
    const int a = a0 + al, b = b0 + bl;
    if (a >= chil || b >= chir) continue;
    zc x[D2];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int l = 0; l < D; ++l) x[k * D + l] = T[(k * BA + al) * (TR + 1) + l * BA + bl];
    const double sc = t.ll[a] * t.lr[b];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        zc o = zmk(0.0, 0.0);
#pragma unroll
        for (int kl = 0; kl < D2; ++kl) zfma(o, sgate[(i * D + j) + D2 * kl], x[kl]);
        o = zscale(o, sc);
        const size_t row = (size_t)i * chil + a;  // theta row (i, alpha)
        const size_t col = (size_t)j * chir + b;  // theta col (j, beta)
        if (!t.transposed) t.W[row + (size_t)t.R * col] = o;
        else t.W[col + (size_t)t.R * row] = zconj(o);
      }
  }
}

// ---------------------------------------------------------------------------
// theta + gate on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  Same tile as
// k_theta_gate: BA alphas x BA betas, rows (k, alpha) and columns (l, beta) of the
// pre-gate theta, TR = BA d.  Stage 1 accumulates the (TR/16)^2 blocks of 16 x 16 over
// the middle bond (4 waves, blocks dealt round-robin); stage 2 applies the gate to the
// BA^2 (alpha, beta) pairs as one (d^2 x d^2) x (d^2 x BA^2) complex product.
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16): lane l supplies A[l%16][l/16] and B[l/16][l%16];
// element g of the result in lane l is D[l/16 + 4g][l%16].
__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void zmfma(zc a, zc b, d4& re, d4& im) {
  re = mfma4(a.x, b.x, re);
  re = mfma4(-a.y, b.y, re);
  im = mfma4(a.x, b.y, im);
  im = mfma4(a.y, b.x, im);
}

template <int D>
struct MfmaThetaCfg {
  static constexpr int BA = (D <= 5) ? 16 : 8;
  static constexpr int TR = BA * D;
  static constexpr int NB1 = TR / 16;
  static constexpr int NBW = (NB1 * NB1 + 3) / 4;
  static constexpr int D2 = D * D;
  static constexpr int MP = (D2 + 15) / 16 * 16;
  static constexpr int KP = (D2 + 3) / 4 * 4;
  static constexpr int NP = BA * BA;
  static constexpr int LDT = TR + 1;
  // middle-bond chunk staged in LDS: 32 (twice the staging of the VALU kernel) halves
  // the barriers per MFMA and gives the next chunk's loads twice the time; the staging
  // shares its LDS with the pre-gate tile, which is as large for d = 4 and 8
  static constexpr int KM = 32;
  static constexpr int AB = 2 * KM * LDT;
  static constexpr int SM = AB > TR * LDT ? AB : TR * LDT;
  static_assert((KM * TR) % 256 == 0, "theta chunks are staged 256 elements at a time");
};

template <int D>
__global__ __launch_bounds__(256) void k_theta_gate_mfma(const BondTask* __restrict__ tasks) {
  using C = MfmaThetaCfg<D>;
  constexpr int BA = C::BA, TR = C::TR, NB1 = C::NB1, NBW = C::NBW, D2 = C::D2;
  constexpr int MP = C::MP, KP = C::KP, NP = C::NP, LDT = C::LDT, KM = C::KM;
  const BondTask t = tasks[blockIdx.z];
  const int a0 = blockIdx.x * BA, b0 = blockIdx.y * BA;
  if (a0 >= t.chil || b0 >= t.chir) return;

  __shared__ zc sm[C::SM];     // As, Bs during stage 1; the pre-gate tile T after it
  __shared__ zc sU[MP * KP];   // gate, row-major [ij][kl], zero padded
  zc* As = sm;                 // [KM][LDT]
  zc* Bs = sm + KM * LDT;      // [KM][LDT]
  zc* T = sm;                  // [TR][LDT]

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;

  for (int e = tid; e < MP * KP; e += 256) {
    const int ij = e / KP, kl = e % KP;
    sU[e] = (ij < D2 && kl < D2) ? t.gate[ij + D2 * kl] : zmk(0.0, 0.0);
  }

  d4 re[NBW], im[NBW];
#pragma unroll
  for (int s = 0; s < NBW; ++s) {
    re[s] = d4{0.0, 0.0, 0.0, 0.0};
    im[s] = d4{0.0, 0.0, 0.0, 0.0};
  }

  const int chil = t.chil, chic = t.chic, chir = t.chir;
  // chunk g0 is loaded into registers while the previous one is multiplied from LDS:
  // element e = tid + 256 u of the A chunk (r = e % TR, g = e / TR) and of the B chunk
  // (g = e % KM, c = e / KM)
  constexpr int NL = KM * TR / 256;
  zc stage_a[NL], stage_b[NL];
  auto load_chunk = [&](int g0) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = tid + 256 * u;
      {
        const int r = e % TR, g = e / TR;
        const int k = r / BA, a = a0 + (r % BA), gg = g0 + g;
        stage_a[u] = (a < chil && gg < chic) ? zscale(t.gl[a + (size_t)chil * (k + (size_t)D * gg)], t.lc[gg])
                                        : zmk(0.0, 0.0);
      }
      {
        const int g = e % KM, c = e / KM;
        const int l = c / BA, b = b0 + (c % BA), gg = g0 + g;
        stage_b[u] = (b < chir && gg < chic) ? t.gr[gg + (size_t)chic * (l + (size_t)D * b)] : zmk(0.0, 0.0);
      }
    }
  };
  load_chunk(0);
  for (int g0 = 0; g0 < chic; g0 += KM) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = tid + 256 * u;
      As[(e / TR) * LDT + e % TR] = stage_a[u];
      Bs[(e % KM) * LDT + e / KM] = stage_b[u];
    }
    __syncthreads();
    if (g0 + KM < chic) load_chunk(g0 + KM);
#pragma unroll
    for (int ks = 0; ks < KM / 4; ++ks) {
      const int kk = ks * 4 + lk;
#pragma unroll
      for (int s = 0; s < NBW; ++s) {
        const int q = wave + 4 * s;
        if (q < NB1 * NB1) {
          const int rb = q / NB1, cb = q % NB1;
          zmfma(As[kk * LDT + rb * 16 + lr], Bs[kk * LDT + cb * 16 + lr], re[s], im[s]);
        }
      }
    }
    __syncthreads();
  }
  // pre-gate tile into LDS (over As / Bs, which are no longer read)
#pragma unroll
  for (int s = 0; s < NBW; ++s) {
    const int q = wave + 4 * s;
    if (q < NB1 * NB1) {
      const int rb = q / NB1, cb = q % NB1;
#pragma unroll
      for (int g = 0; g < 4; ++g) T[(rb * 16 + lk + 4 * g) * LDT + cb * 16 + lr] = zmk(re[s][g], im[s][g]);
    }
  }
  __syncthreads();

  const bool tr = t.transposed != 0;
  for (int pb = wave; pb < NP / 16; pb += 4) {
    const int p = pb * 16 + lr;
    const int al = tr ? p / BA : p % BA;
    const int bl = tr ? p % BA : p / BA;
    const int a = a0 + al, b = b0 + bl;
    const bool in = a < chil && b < chir;
    const double sc = in ? t.ll[a] * t.lr[b] : 0.0;
#pragma unroll
    for (int mb = 0; mb < MP / 16; ++mb) {
      d4 ore = {0.0, 0.0, 0.0, 0.0}, oim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < KP / 4; ++ks) {
        const int kl = ks * 4 + lk;
        const zc x = kl < D2 ? T[((kl / D) * BA + al) * LDT + (kl % D) * BA + bl] : zmk(0.0, 0.0);
        zmfma(sU[(mb * 16 + lr) * KP + kl], x, ore, oim);
      }
      if (in) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int ij = mb * 16 + lk + 4 * g;
          if (ij < D2) {
            const int i = ij / D, j = ij % D;
            const size_t row = (size_t)i * chil + a;
            const size_t col = (size_t)j * chir + b;
            if (!tr) t.W[row + (size_t)t.R * col] = zmk(ore[g] * sc, oim[g] * sc);
            else t.W[col + (size_t)t.R * row] = zmk(ore[g] * sc, -(oim[g] * sc));
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// W padding columns := 0, V := identity, the reference total weight (||theta||^2 or none)
// ---------------------------------------------------------------------------
__global__ void k_prepare_svd(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const size_t nV = (size_t)t.Cpad * t.Cpad;
  const size_t nWp = (size_t)t.R * (t.Cpad - t.C);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nV + nWp; e += (size_t)gridDim.x * blockDim.x) {
    if (e < nV) {
      const size_t r = e % t.Cpad, c = e / t.Cpad;
      t.V[e] = zmk(r == c ? 1.0 : 0.0, 0.0);
    } else {
      t.W[(size_t)t.R * t.C + (e - nV)] = zmk(0.0, 0.0);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) t.result[3] = t.wtot ? *t.wtot : -1.0;
}

// ---------------------------------------------------------------------------
// column norms, truncation, Gamma rebuild
// ---------------------------------------------------------------------------
// one wavefront per column: s[c] = ||W[:, c]||^2 (not yet sorted)
__global__ __launch_bounds__(256) void k_col_norms(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + wave;
  if (c >= t.Cpad) return;
  double acc = 0.0;
  if (c < t.C)
    for (int r = lane; r < t.R; r += 64) acc += zabs2(t.W[r + (size_t)t.R * c]);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) t.s[c] = acc;
}

// Column-norm reductions over s (one workgroup per task, fixed order so repeated
// runs are bitwise identical): result[2] = max_c |w_c|^2 (the Jacobi rounding
// floor's scale), and with keep_sum also result[3] = sum_c |w_c|^2 = ||W||_F^2.
__global__ __launch_bounds__(256) void k_frob(const BondTask* __restrict__ tasks, int keep_sum) {
  const BondTask t = tasks[blockIdx.x];
  __shared__ double red[256];
  __shared__ double rmx[256];
  double acc = 0.0, mx = 0.0;
  for (int c = threadIdx.x; c < t.C; c += 256) { acc += t.s[c]; mx = fmax(mx, t.s[c]); }
  red[threadIdx.x] = acc;
  rmx[threadIdx.x] = mx;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      red[threadIdx.x] += red[threadIdx.x + st];
      rmx[threadIdx.x] = fmax(rmx[threadIdx.x], rmx[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    t.result[2] = rmx[0];
    if (keep_sum) t.result[3] = red[0];
  }
}

constexpr int SORT_N = 4096;

// one workgroup per task: bitonic sort of the squared norms (descending), kept rank,
// lambda_c and truncerr.  Rule (FormLambda1): keep index < chi_max with
// s^2/||theta||^2 > cutoff (cutoff >= 1e-30, i.e. |s| > 1e-15 after normalisation).
__global__ __launch_bounds__(1024) void k_truncate(const BondTask* __restrict__ tasks, double cutoff) {
  const BondTask t = tasks[blockIdx.x];
  __shared__ double key[SORT_N];
  __shared__ int idx[SORT_N];
  __shared__ double red[1024];
  __shared__ int s_k;
  const int tid = threadIdx.x;
  int n2 = 1;
  while (n2 < t.Cpad) n2 <<= 1;
  for (int e = tid; e < n2; e += 1024) {
    key[e] = (e < t.C) ? t.s[e] : -1.0;
    idx[e] = e;
  }
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid; e < n2; e += 1024) {
        const int partner = e ^ stride;
        if (partner > e) {
          const bool desc = ((e & size) == 0);
          const double ke = key[e], kp = key[partner];
          const bool swap = desc ? (ke < kp) : (ke > kp);
          if (swap) {
            key[e] = kp; key[partner] = ke;
            const int ti = idx[e]; idx[e] = idx[partner]; idx[partner] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  // weight held by W; when W is a sketch of theta the reference total is ||theta||^2
  double part = 0.0;
  for (int e = tid; e < t.C; e += 1024) part += key[e];
  red[tid] = part;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) { if (tid < st) red[tid] += red[tid + st]; __syncthreads(); }
  const double captured = red[0];
  const double total = t.result[3] > 0.0 ? t.result[3] : captured;
  __syncthreads();
  if (tid == 0) s_k = 0;
  __syncthreads();
  const int kmax = t.C < t.chi_max ? t.C : t.chi_max;
  const double thr = (cutoff > 1e-30 ? cutoff : 1e-30) * total;
  int cnt = 0;
  for (int e = tid; e < kmax; e += 1024) cnt += (key[e] > thr) ? 1 : 0;
  if (cnt) atomicAdd(&s_k, cnt);
  __syncthreads();
  const int k = s_k;
  part = 0.0;
  for (int e = tid; e < k; e += 1024) part += key[e];
  red[tid] = part;
  __syncthreads();
  for (int st = 512; st > 0; st >>= 1) { if (tid < st) red[tid] += red[tid + st]; __syncthreads(); }
  const double kept = red[0];
  const double inv = kept > 0.0 ? 1.0 / sqrt(kept) : 0.0;
  for (int e = tid; e < t.Cpad; e += 1024) {
    const double v = (e < t.C && key[e] > 0.0) ? key[e] : 0.0;
    t.s[e] = sqrt(v);
    t.perm[e] = idx[e];
  }
  for (int e = tid; e < t.chi_max; e += 1024) t.lc_out[e] = (e < k) ? sqrt(key[e]) * inv : 0.0;
  if (tid == 0) {
    t.result[0] = (double)k;
    t.result[1] = total > 0.0 ? fabs(1.0 - kept / total) : 1.0;
    t.result[2] = total;
    t.result[4] = captured;
    t.result[5] = k > 0 ? key[k - 1] : 0.0;
  }
}

// Gamma_l(a,i,b) = U((i,a), b)/ll(a);  Gamma_r(g,j,b) = conj(V((j,b), g))/lr(b), with U and V
// the left / right singular vectors of theta: from W / s and the accumulated V of the
// Jacobi, or with QR preconditioning from Yr and Zc (already sorted and normalised)
__global__ void k_write_gammas(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const int k = (int)t.result[0];
  const int dd = t.d;
  if (t.lc_commit && blockIdx.x == 0)
    for (int e = threadIdx.x; e < t.chi_max; e += blockDim.x) t.lc_commit[e] = t.lc_out[e];
  const size_t nl = (size_t)t.chil * dd * k;
  const size_t nr = (size_t)k * dd * t.chir;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nl + nr; e += (size_t)gridDim.x * blockDim.x) {
    if (e < nl) {
      const int a = (int)(e % t.chil);
      const size_t rest = e / t.chil;
      const int i = (int)(rest % dd);
      const int b = (int)(rest / dd);
      const size_t row = (size_t)i * t.chil + a;
      const int col = t.perm[b];
      zc u;
      if (t.qr) u = !t.transposed ? t.Yr[row + (size_t)t.Rb * b] : t.Zc[row + (size_t)t.Cb * b];
      else if (!t.transposed) u = zscale(t.W[row + (size_t)t.R * col], t.s[b] > 0.0 ? 1.0 / t.s[b] : 0.0);
      else u = t.V[row + (size_t)t.NV * col];
      const double la = t.ll[a];
      t.gl_out[e] = fabs(la) > 1e-15 ? zscale(u, 1.0 / la) : zmk(0.0, 0.0);
    } else {
      const size_t f = e - nl;
      const int g = (int)(f % k);
      const size_t rest = f / k;
      const int j = (int)(rest % dd);
      const int b = (int)(rest / dd);
      const size_t row = (size_t)j * t.chir + b;
      const int col = t.perm[g];
      zc v;
      if (t.qr) v = !t.transposed ? t.Zc[row + (size_t)t.Cb * g] : t.Yr[row + (size_t)t.Rb * g];
      else if (!t.transposed) v = t.V[row + (size_t)t.NV * col];
      else v = zscale(t.W[row + (size_t)t.R * col], t.s[g] > 0.0 ? 1.0 / t.s[g] : 0.0);
      const double lb = t.lr[b];
      t.gr_out[f] = fabs(lb) > 1e-15 ? zscale(zconj(v), 1.0 / lb) : zmk(0.0, 0.0);
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_theta_gate(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                             hipStream_t st) {
  auto go = [&](auto kern, int BA) {
    dim3 grid((max_chil + BA - 1) / BA, (max_chir + BA - 1) / BA, ntasks);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, d_tasks);
    return hipGetLastError();
  };
  switch (d) {
    case 2: return go(k_theta_gate<2>, ThetaCfg<2>::BA);
    case 3: return go(k_theta_gate<3>, ThetaCfg<3>::BA);
    case 4: return go(k_theta_gate<4>, ThetaCfg<4>::BA);
    case 5: return go(k_theta_gate<5>, ThetaCfg<5>::BA);
    case 6: return go(k_theta_gate<6>, ThetaCfg<6>::BA);
    case 8: return go(k_theta_gate<8>, ThetaCfg<8>::BA);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_theta_gate_mfma(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                                  hipStream_t st) {
  auto go = [&](auto kern, int BA) {
    dim3 grid((max_chil + BA - 1) / BA, (max_chir + BA - 1) / BA, ntasks);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, d_tasks);
    return hipGetLastError();
  };
  switch (d) {
    case 2: return go(k_theta_gate_mfma<2>, MfmaThetaCfg<2>::BA);
    case 3: return go(k_theta_gate_mfma<3>, MfmaThetaCfg<3>::BA);
    case 4: return go(k_theta_gate_mfma<4>, MfmaThetaCfg<4>::BA);
    case 5: return go(k_theta_gate_mfma<5>, MfmaThetaCfg<5>::BA);
    case 6: return go(k_theta_gate_mfma<6>, MfmaThetaCfg<6>::BA);
    case 8: return go(k_theta_gate_mfma<8>, MfmaThetaCfg<8>::BA);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_prepare_svd(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_prepare_svd, dim3(256, ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

hipError_t launch_frobenius(const BondTask* d_tasks, int ntasks, int max_cpad, hipStream_t st, int keep_sum) {
  hipLaunchKernelGGL(k_col_norms, dim3((max_cpad + 3) / 4, ntasks), dim3(256), 0, st, d_tasks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_frob, dim3(ntasks), dim3(256), 0, st, d_tasks, keep_sum);
  return hipGetLastError();
}

hipError_t launch_finalize(const BondTask* d_tasks, int ntasks, int max_cpad, double cutoff, hipStream_t st) {
  hipLaunchKernelGGL(k_col_norms, dim3((max_cpad + 3) / 4, ntasks), dim3(256), 0, st, d_tasks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_truncate, dim3(ntasks), dim3(1024), 0, st, d_tasks, cutoff);
  return hipGetLastError();
}

hipError_t launch_col_norms(const BondTask* d_tasks, int ntasks, int max_cpad, hipStream_t st) {
  hipLaunchKernelGGL(k_col_norms, dim3((max_cpad + 3) / 4, ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

hipError_t launch_truncate(const BondTask* d_tasks, int ntasks, double cutoff, hipStream_t st) {
  hipLaunchKernelGGL(k_truncate, dim3(ntasks), dim3(1024), 0, st, d_tasks, cutoff);
  return hipGetLastError();
}

hipError_t launch_write_gammas(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_write_gammas, dim3(512, ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}
