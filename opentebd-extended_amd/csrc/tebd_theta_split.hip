// gfx950 kernels of the split-operand two-site update: FormTheta + ThetaOperation_c
// (local_operations_module.f90:184-211,248-280) with the middle-bond contraction on the
// int8 matrix cores and FP64 everywhere else.
//
//   theta'(a,i,j,b) = ll(a) lr(b) sum_kl U(ij,kl) sum_g A((k,a), g) B(g, (l,b)),
//   A((k,a), g) = Gl(a,k,g) lc(g),  B(g, (l,b)) = Gr(g,l,b).
//
// In a Vidal state Gamma_r(g, l, b) carries 1/lambda_c(g) and A carries lambda_c(g), so
// each middle index g is first balanced by a power of two: A(:, g) 2^eg, B(g, :) 2^-eg,
// eg the binary exponent of B's largest entry in row g (exact, and the product is
// unchanged).  Then every row of A and every column of B is scaled by a power of two,
// 2^e with e the binary exponent of its largest real or imaginary part, and cut by
// truncation into S signed 7-bit slices: x = 2^e sum_s q_s 2^-7(s+1) + r, q_s in
// [-127, 127], |r| < 2^(e-7S).  The truncations and the subtractions are exact in FP64.
// The contraction keeps the slice products of level L = s + t < S:
//   sum_g A B = 2^(e_r + e_c) sum_L 2^-7(L+2) sum_(s+t=L) sum_g q^A_s q^B_t,
// and each inner sum is an integer that v_mfma_i32_16x16x64_i8 accumulates exactly in
// int32 over a 64-wide chunk of g (at most S slice pairs x 2 x 64 products of magnitude
// <= 127^2 per level: far inside int32).  Each level's integer is exact in FP64 and
// scaled by an exact power of two, so the only roundings are the FP64 sums over levels
// and chunks; the dropped slices and levels leave ~(S + 3) 2^-7S |A_r|max |B_c|max per
// product term (S = 7: 2e-14).  The gate (d^2 x d^2) and the lambda scaling are FP64,
// and so is everything after theta.
//
// Slices are stored as plane p = 2 s + part (real, imaginary), group h = g / 16, row r:
// 16 bytes (g % 16) at ((p G + h) rows + r) 16, G = pitch / 16 groups, so that both the
// slicing and the contraction's chunk loads move whole runs of rows.  Rows of A are
// r = k round_up(chi_l, 16) + a, columns of B c = l round_up(chi_r, 16) + b.
//
//   k_split_clear    exponents := "no nonzero value"
//   k_split_gexp     the balance exponent eg of each middle index (atomic max)
//   k_split_exp      the exponent of each row of A / column of B (atomic max)
//   k_split_cut      the slices of A / B, one thread per (row, group of 16 g)
//   k_theta_split    per BA x BA block of (alpha, beta): the slice products on the int8
//                    matrix cores (next chunk loaded while this one is multiplied), the
//                    FP64 combination, the gate and the write into W
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int S = kSplitSlices;
constexpr int KC = 64;                 // chunk of the middle bond: one i8 MFMA k-extent
constexpr int kBadExp = 1 << 20;       // a row or column holds a value that is not finite
constexpr int kNoExp = -(1 << 20);     // no nonzero value (yet)

// D(16x16) += A(16x64) B(64x16) in int32: lane l supplies row l%16 of A and column
// l%16 of B over the same 16 values of k (bytes of a v4i, k block l/16); element v of
// the result in lane l is D[4 (l/16) + v][l%16].
__device__ __forceinline__ v4i mfma_i8(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Slices of y = x 2^-e (|y| < 1) into q[0..S): q_s = trunc(128 y_s), y_(s+1) = 128 y_s - q_s.
__device__ __forceinline__ void cut(double y, signed char* q) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    y *= 128.0;
    const double t = trunc(y);
    q[s] = (signed char)(int)t;
    y -= t;
  }
}

// Exponent of one complex value: the binary exponent E of its larger part (|part| <
// 2^E), kNoExp for zero, kBadExp when it is not finite.
__device__ __forceinline__ int zexp(zc x) {
  if (!isfinite(x.x) || !isfinite(x.y)) return kBadExp;
  const double m = fmax(fabs(x.x), fabs(x.y));
  if (m == 0.0) return kNoExp;
  int e;
  (void)frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
  return e;
}

__device__ __forceinline__ int eg_of(const int* eg, int g) {
  const int e = eg[g];
  return e == kNoExp ? 0 : e;
}

__global__ __launch_bounds__(256) void k_split_clear(SplitWs ws, int ntasks) {
  const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
  if (i < ws.rows * ntasks) {
    ws.ea[i] = kNoExp;
    ws.eb[i] = kNoExp;
  }
  if (i < (size_t)ws.pitch * ntasks) ws.eg[i] = kNoExp;
}

// eg[g] = max over the columns c of the exponent of B(g, c) = Gamma_r(g, c) (values that
// are not finite are left to the column exponents).  Block: 64 middle indices x 4
// threads, each over 16 columns of the block's 64 (grid.z).
__global__ __launch_bounds__(256) void k_split_gexp(const BondTask* __restrict__ tasks, SplitWs ws) {
  const BondTask t = tasks[blockIdx.y];
  const int g = blockIdx.x * 64 + (threadIdx.x & 63);
  if (g >= t.chic) return;
  const int nc = t.d * t.chir;
  const int c0 = blockIdx.z * 64 + (threadIdx.x >> 6) * 16;
  int m = kNoExp;
  for (int j = 0; j < 16 && c0 + j < nc; ++j) {
    const int e = zexp(t.gr[g + (size_t)t.chic * (c0 + j)]);
    if (e != kBadExp) m = max(m, e);
  }
  if (m != kNoExp) atomicMax(ws.eg + (size_t)blockIdx.y * ws.pitch + g, m);
}

// Value g of row r of A (PART 0) or column r of B (PART 1), balanced.
template <int PART>
__device__ __forceinline__ zc split_value(const BondTask& t, const int* eg, int r, int g) {
  if (g >= t.chic) return zmk(0.0, 0.0);
  if (PART == 0) {
    const int chilp = (t.chil + 15) / 16 * 16;
    const int k = r / chilp, a = r % chilp;
    if (a >= t.chil) return zmk(0.0, 0.0);
    const zc x = zscale(t.gl[a + (size_t)t.chil * (k + (size_t)t.d * g)], t.lc[g]);
    return zmk(ldexp(x.x, eg_of(eg, g)), ldexp(x.y, eg_of(eg, g)));
  }
  const int chirp = (t.chir + 15) / 16 * 16;
  const int l = r / chirp, b = r % chirp;
  if (b >= t.chir) return zmk(0.0, 0.0);
  const zc x = t.gr[g + (size_t)t.chic * (l + (size_t)t.d * b)];
  return zmk(ldexp(x.x, -eg_of(eg, g)), ldexp(x.y, -eg_of(eg, g)));
}

// Rows (PART 0) or columns (PART 1) of one operand: grid (rows / 256, tasks, groups).
template <int PART>
__device__ __forceinline__ int split_rows_of(const BondTask& t) {
  return t.d * (((PART == 0 ? t.chil : t.chir) + 15) / 16 * 16);
}

template <int PART>
__global__ __launch_bounds__(256) void k_split_exp(const BondTask* __restrict__ tasks, SplitWs ws) {
  const BondTask t = tasks[blockIdx.y];
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int g0 = 16 * blockIdx.z;
  if (r >= split_rows_of<PART>(t) || g0 >= t.chic) return;
  const int* eg = ws.eg + (size_t)blockIdx.y * ws.pitch;
  int m = kNoExp;
#pragma unroll 4
  for (int u = 0; u < 16; ++u) m = max(m, zexp(split_value<PART>(t, eg, r, g0 + u)));
  if (m != kNoExp) atomicMax((PART == 0 ? ws.ea : ws.eb) + (size_t)blockIdx.y * ws.rows + r, m);
}

template <int PART>
__global__ __launch_bounds__(256) void k_split_cut(const BondTask* __restrict__ tasks, SplitWs ws) {
  const BondTask t = tasks[blockIdx.y];
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int h = blockIdx.z;
  if (r >= split_rows_of<PART>(t) || 16 * h >= ws.pitch) return;
  const int* eg = ws.eg + (size_t)blockIdx.y * ws.pitch;
  const int er = (PART == 0 ? ws.ea : ws.eb)[(size_t)blockIdx.y * ws.rows + r];
  const bool zero = er == kNoExp || er == kBadExp;  // all zero, or flagged (theta gets NaN)
  unsigned wr[S][4], wi[S][4];
#pragma unroll
  for (int s = 0; s < S; ++s)
#pragma unroll
    for (int w = 0; w < 4; ++w) wr[s][w] = wi[s][w] = 0u;
  if (!zero) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const zc x = split_value<PART>(t, eg, r, 16 * h + u);
      signed char tr[S], ti[S];
      cut(ldexp(x.x, -er), tr);
      cut(ldexp(x.y, -er), ti);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        wr[s][u / 4] |= ((unsigned)(unsigned char)tr[s]) << (8 * (u % 4));
        wi[s][u / 4] |= ((unsigned)(unsigned char)ti[s]) << (8 * (u % 4));
      }
    }
  }
  v4i* out = reinterpret_cast<v4i*>(PART == 0 ? ws.a : ws.b) + (size_t)blockIdx.y * (ws.stride / 16);
  const size_t G = ws.pitch / 16;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    out[((2 * s) * G + h) * ws.rows + r] = v4i{(int)wr[s][0], (int)wr[s][1], (int)wr[s][2], (int)wr[s][3]};
    out[((2 * s + 1) * G + h) * ws.rows + r] = v4i{(int)wi[s][0], (int)wi[s][1], (int)wi[s][2], (int)wi[s][3]};
  }
}

template <int D>
struct SplitCfg {
  static constexpr int BA = (D <= 5) ? 16 : 8;  // alpha block == beta block
  static constexpr int TR = BA * D;              // tile edge (rows (k, a), columns (l, b))
  static constexpr int NB1 = TR / 16;            // 16 x 16 blocks per edge
  static constexpr int NW = NB1 * NB1 >= 8 ? 8 : NB1 * NB1;  // waves
  static constexpr int NT = 64 * NW;
  static constexpr int NBW = (NB1 * NB1 + NW - 1) / NW;      // blocks per wave
  static constexpr int D2 = D * D;
  static constexpr int NP = BA * BA;             // (alpha, beta) pairs of the tile
  static constexpr int LDT = TR + 1;
  static constexpr int NV = 2 * S * 4 * TR;      // 16-byte groups of one operand's chunk
  static constexpr int CHUNK = NV * 16;
  static constexpr int NPF = (NV + NT - 1) / NT; // groups each thread prefetches per operand
  static constexpr int TILE = TR * LDT * 16;     // bytes of the FP64 pre-gate tile
  static constexpr int GATE = D2 * D2 * 16;      // bytes of the gate
  static constexpr int LDS = (2 * CHUNK > TILE + GATE ? 2 * CHUNK : TILE + GATE);
};

template <int D>
__global__ __launch_bounds__(SplitCfg<D>::NT) void k_theta_split(const BondTask* __restrict__ tasks, SplitWs ws) {
  using C = SplitCfg<D>;
  constexpr int BA = C::BA, TR = C::TR, NB1 = C::NB1, NW = C::NW, NT = C::NT, NBW = C::NBW;
  constexpr int D2 = C::D2, NP = C::NP, LDT = C::LDT, NV = C::NV, NPF = C::NPF;
  const BondTask t = tasks[blockIdx.z];
  const int a0 = blockIdx.x * BA, b0 = blockIdx.y * BA;
  if (a0 >= t.chil || b0 >= t.chir) return;
  const int chilp = (t.chil + 15) / 16 * 16, chirp = (t.chir + 15) / 16 * 16;

  // operand chunks: [plane 2 S][group 4][row TR] 16-byte groups (tile row k BA + al); after
  // the contraction the same memory holds the FP64 pre-gate tile and the gate, [ij][kl]
  __shared__ __align__(16) unsigned char lds[C::LDS];
  v4i* As = reinterpret_cast<v4i*>(lds);
  v4i* Bs = As + NV;
  zc* T = reinterpret_cast<zc*>(lds);
  zc* sU = reinterpret_cast<zc*>(lds + C::TILE);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const size_t task = blockIdx.z;
  const size_t G = ws.pitch / 16;
  const v4i* ga = reinterpret_cast<const v4i*>(ws.a) + task * (ws.stride / 16);
  const v4i* gb = reinterpret_cast<const v4i*>(ws.b) + task * (ws.stride / 16);

  v4i pa[NPF], pb[NPF];  // the next chunk, in flight while this one is multiplied
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
      const int e = tid + NT * u;
      if (e < NV) {
        const int pl = e / (4 * TR), rem = e % (4 * TR), hl = rem / TR, row = rem % TR;
        const int k = row / BA, al = row % BA;
        const size_t gidx = (pl * G + k0 / 16 + hl) * ws.rows;
        pa[u] = ga[gidx + (size_t)k * chilp + a0 + al];
        pb[u] = gb[gidx + (size_t)k * chirp + b0 + al];
      }
    }
  };

  double accr[NBW][4], acci[NBW][4];
#pragma unroll
  for (int q = 0; q < NBW; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) accr[q][v] = acci[q][v] = 0.0;

  const int kend = (t.chic + KC - 1) / KC * KC;
  load_chunk(0);
  for (int k0 = 0; k0 < kend; k0 += KC) {
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
      const int e = tid + NT * u;
      if (e < NV) {
        As[e] = pa[u];
        Bs[e] = pb[u];
      }
    }
    __syncthreads();
    if (k0 + KC < kend) load_chunk(k0 + KC);
#pragma unroll 1
    for (int q = 0; q < NBW; ++q) {  // one block at a time: its slices stay in registers
      const int blk = wave + NW * q;
      if (blk < NB1 * NB1) {
        const int rb = blk / NB1, cb = blk % NB1;
        v4i ar[S], ai[S], br[S], bi[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
          ar[s] = As[((2 * s) * 4 + lk) * TR + rb * 16 + lr];
          ai[s] = As[((2 * s + 1) * 4 + lk) * TR + rb * 16 + lr];
          br[s] = Bs[((2 * s) * 4 + lk) * TR + cb * 16 + lr];
          bi[s] = Bs[((2 * s + 1) * 4 + lk) * TR + cb * 16 + lr];
        }
        double xr[4] = {0.0, 0.0, 0.0, 0.0}, xi[4] = {0.0, 0.0, 0.0, 0.0};
        double sc = 6.103515625e-05;  // 2^-7(L+2), L = 0
#pragma unroll
        for (int L = 0; L < S; ++L) {  // level by level: three live integer sums
          v4i pre = {0, 0, 0, 0}, pii = {0, 0, 0, 0}, pim = {0, 0, 0, 0};
#pragma unroll
          for (int s = 0; s <= L; ++s) {
            pre = mfma_i8(ar[s], br[L - s], pre);
            pii = mfma_i8(ai[s], bi[L - s], pii);
            pim = mfma_i8(ar[s], bi[L - s], pim);
            pim = mfma_i8(ai[s], br[L - s], pim);
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            xr[v] = fma((double)(pre[v] - pii[v]), sc, xr[v]);
            xi[v] = fma((double)pim[v], sc, xi[v]);
          }
          sc *= 0.0078125;
        }
#pragma unroll
        for (int qq = 0; qq < NBW; ++qq)  // static register index of block q's sums
          if (qq == q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              accr[qq][v] += xr[v];
              acci[qq][v] += xi[v];
            }
          }
      }
    }
    __syncthreads();
  }

  for (int e = tid; e < D2 * D2; e += NT) sU[e] = t.gate[(e / D2) + D2 * (e % D2)];
  // pre-gate tile in LDS (over the chunk buffers): T[row][col], row = k BA + al,
  // col = l BA + bl, scaled by 2^(e_r + e_c)
  const int* ea = ws.ea + task * ws.rows;
  const int* eb = ws.eb + task * ws.rows;
#pragma unroll
  for (int q = 0; q < NBW; ++q) {
    const int blk = wave + NW * q;
    if (blk < NB1 * NB1) {
      const int rb = blk / NB1, cb = blk % NB1;
      const int col = cb * 16 + lr;
      const int ec = eb[(size_t)(col / BA) * chirp + b0 + col % BA];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rb * 16 + 4 * lk + v;
        const int er = ea[(size_t)(row / BA) * chilp + a0 + row % BA];
        zc o;
        if (er == kBadExp || ec == kBadExp) {
          o = zmk(__builtin_nan(""), __builtin_nan(""));
        } else if (er == kNoExp || ec == kNoExp) {
          o = zmk(0.0, 0.0);
        } else {
          o = zmk(ldexp(accr[q][v], er + ec), ldexp(acci[q][v], er + ec));
        }
        T[row * LDT + col] = o;
      }
    }
  }
  __syncthreads();

  // gate, FP64: O(ij, p) = sum_kl U(ij, kl) X(kl, p), X(kl, p) = T(k BA + al, l BA + bl).
  // p runs over alpha fastest (theta stored as is) or beta fastest (theta^dagger), so
  // consecutive threads write consecutive elements of W.
  const bool tr = t.transposed != 0;
  constexpr int NG = NT / NP > 0 ? NT / NP : 1;  // thread groups splitting the ij range
  for (int e = tid; e < NP * NG; e += NT) {
    const int p = e % NP, grp = e / NP;
    const int al = tr ? p / BA : p % BA;
    const int bl = tr ? p % BA : p / BA;
    const int a = a0 + al, b = b0 + bl;
    if (a >= t.chil || b >= t.chir) continue;
    const double sc = t.ll[a] * t.lr[b];
    constexpr int NX = D2 <= 16 ? D2 : 1;  // X(:, p) held in registers for small d
    zc x[NX];
    if (D2 <= 16) {
#pragma unroll
      for (int kl = 0; kl < NX; ++kl) x[kl] = T[((kl / D) * BA + al) * LDT + (kl % D) * BA + bl];
    }
    for (int ij = grp; ij < D2; ij += NG) {
      zc o = zmk(0.0, 0.0);
      if (D2 <= 16) {
#pragma unroll
        for (int kl = 0; kl < NX; ++kl) zfma(o, sU[ij * D2 + kl], x[kl]);
      } else {
#pragma unroll 4
        for (int kl = 0; kl < D2; ++kl) zfma(o, sU[ij * D2 + kl], T[((kl / D) * BA + al) * LDT + (kl % D) * BA + bl]);
      }
      const int i = ij / D, j = ij % D;
      const size_t row = (size_t)i * t.chil + a;  // theta row (i, alpha)
      const size_t col = (size_t)j * t.chir + b;  // theta col (j, beta)
      const zc w = zscale(o, sc);
      if (!tr) t.W[row + (size_t)t.R * col] = w;
      else t.W[col + (size_t)t.R * row] = zconj(w);
    }
  }
}

}  // namespace

hipError_t launch_theta_gate_split(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                                   const SplitWs& ws, hipStream_t st) {
  const int rows_l = (max_chil + 15) / 16 * 16 * d, rows_r = (max_chir + 15) / 16 * 16 * d;
  if ((size_t)rows_l > ws.rows || (size_t)rows_r > ws.rows) return hipErrorInvalidValue;
  const int groups = ws.pitch / 16;
  const size_t nclear = std::max(ws.rows, (size_t)ws.pitch) * ntasks;
  hipLaunchKernelGGL(k_split_clear, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, st, ws, ntasks);
  hipLaunchKernelGGL(k_split_gexp, dim3((ws.pitch + 63) / 64, ntasks, (rows_r + 63) / 64), dim3(256), 0, st, d_tasks,
                     ws);
  hipLaunchKernelGGL(k_split_exp<0>, dim3((rows_l + 255) / 256, ntasks, groups), dim3(256), 0, st, d_tasks, ws);
  hipLaunchKernelGGL(k_split_exp<1>, dim3((rows_r + 255) / 256, ntasks, groups), dim3(256), 0, st, d_tasks, ws);
  hipLaunchKernelGGL(k_split_cut<0>, dim3((rows_l + 255) / 256, ntasks, groups), dim3(256), 0, st, d_tasks, ws);
  hipLaunchKernelGGL(k_split_cut<1>, dim3((rows_r + 255) / 256, ntasks, groups), dim3(256), 0, st, d_tasks, ws);
  auto go = [&](auto kern, int BA, int nt) {
    dim3 grid((max_chil + BA - 1) / BA, (max_chir + BA - 1) / BA, ntasks);
    hipLaunchKernelGGL(kern, grid, dim3(nt), 0, st, d_tasks, ws);
    return hipGetLastError();
  };
  switch (d) {
    case 2: return go(k_theta_split<2>, SplitCfg<2>::BA, SplitCfg<2>::NT);
    case 3: return go(k_theta_split<3>, SplitCfg<3>::BA, SplitCfg<3>::NT);
    case 4: return go(k_theta_split<4>, SplitCfg<4>::BA, SplitCfg<4>::NT);
    case 5: return go(k_theta_split<5>, SplitCfg<5>::BA, SplitCfg<5>::NT);
    case 6: return go(k_theta_split<6>, SplitCfg<6>::BA, SplitCfg<6>::NT);
    case 8: return go(k_theta_split<8>, SplitCfg<8>::BA, SplitCfg<8>::NT);
    default: return hipErrorInvalidValue;
  }
}
