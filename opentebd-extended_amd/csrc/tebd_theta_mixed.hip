// gfx950 kernel of the mixed-precision two-site update: FormTheta + ThetaOperation_c
// (local_operations_module.f90:184-211,248-280) on the FP32 matrix cores.
//
//   theta'(a,i,j,b) = ll(a) lr(b) sum_kl U(ij,kl) sum_g Gl(a,k,g) lc(g) Gr(g,l,b)
//
// Both contractions run as complex products of real v_mfma_f32_16x16x4f32 tiles with
// FP32 accumulation (four real products per complex one).  The operands are rounded
// from the FP64 state once, when they are staged in LDS; the lambda_l lambda_r scaling
// and the write into the SVD work matrix W are FP64, so everything after theta (Gram /
// orthogonalisation, cutoff, lambda renormalisation) stays FP64.
//
// Workgroup tile: BA alphas x BA betas, i.e. TR x TR entries of the pre-gate theta with
// rows (k, alpha) and columns (l, beta), TR = BA d.  Stage 1 accumulates it over the
// middle bond in chunks of KC; stage 2 applies the d^2 x d^2 gate to the BA^2 (alpha,
// beta) pairs of the tile as one (d^2 x d^2) x (d^2 x BA^2) product.
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

struct __align__(8) cf {
  float x, y;
};

__device__ __forceinline__ f4 mfma_f32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// acc_re/acc_im += (a) (b), complex, one 16x16x4 step
__device__ __forceinline__ void cmfma(cf a, cf b, f4& re, f4& im) {
  re = mfma_f32(a.x, b.x, re);
  re = mfma_f32(-a.y, b.y, re);
  im = mfma_f32(a.x, b.y, im);
  im = mfma_f32(a.y, b.x, im);
}

constexpr int KC = 16;  // middle-bond chunk staged in LDS (4 MFMA k-steps)

template <int D>
struct MixCfg {
  static constexpr int BA = (D <= 5) ? 16 : 8;        // alpha block == beta block
  static constexpr int TR = BA * D;                    // tile edge, a multiple of 16
  static constexpr int NB1 = TR / 16;                  // 16 x 16 MFMA blocks per edge
  static constexpr int NBW = (NB1 * NB1 + 3) / 4;      // blocks per wave (4 waves)
  static constexpr int D2 = D * D;
  static constexpr int MP = (D2 + 15) / 16 * 16;       // gate rows, padded
  static constexpr int KP = (D2 + 3) / 4 * 4;          // gate columns (contraction), padded
  static constexpr int NP = BA * BA;                   // (alpha, beta) pairs of the tile
  static constexpr int LDT = TR + 1;
};

template <int D>
__global__ __launch_bounds__(256) void k_theta_gate_mixed(const BondTask* __restrict__ tasks) {
  using C = MixCfg<D>;
  constexpr int BA = C::BA, TR = C::TR, NB1 = C::NB1, NBW = C::NBW, D2 = C::D2;
  constexpr int MP = C::MP, KP = C::KP, NP = C::NP, LDT = C::LDT;
  const BondTask t = tasks[blockIdx.z];
  const int a0 = blockIdx.x * BA, b0 = blockIdx.y * BA;
  if (a0 >= t.chil || b0 >= t.chir) return;

  __shared__ cf As[KC][LDT];
  __shared__ cf Bs[KC][LDT];
  __shared__ cf T[TR * LDT];
  __shared__ cf sU[MP * KP];  // gate, row-major [ij][kl], zero padded

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr16 = lane & 15, lk = lane >> 4;

  for (int e = tid; e < MP * KP; e += 256) {
    const int ij = e / KP, kl = e % KP;
    cf u = {0.f, 0.f};
    if (ij < D2 && kl < D2) {
      const zc g = t.gate[ij + D2 * kl];
      u.x = (float)g.x;
      u.y = (float)g.y;
    }
    sU[e] = u;
  }

  f4 re[NBW], im[NBW];
#pragma unroll
  for (int s = 0; s < NBW; ++s) {
    re[s] = f4{0.f, 0.f, 0.f, 0.f};
    im[s] = f4{0.f, 0.f, 0.f, 0.f};
  }

  const int chil = t.chil, chic = t.chic, chir = t.chir;
  // chunk g0 is loaded (and rounded) into registers while the previous one is
  // multiplied from LDS: element e = tid + 256 u of each chunk
  constexpr int NL = KC * TR / 256;
  static_assert((KC * TR) % 256 == 0, "theta chunks are staged 256 elements at a time");
  cf stage_a[NL], stage_b[NL];
  auto load_chunk = [&](int g0) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = tid + 256 * u;
      {  // A(r, g) = Gl(a, k, g) lc(g), r = k BA + a_local (contiguous alpha within one k)
        const int r = e % TR, g = e / TR;
        const int k = r / BA, a = a0 + (r % BA), gg = g0 + g;
        cf v = {0.f, 0.f};
        if (a < chil && gg < chic) {
          const zc x = t.gl[a + (size_t)chil * (k + (size_t)D * gg)];
          const double l = t.lc[gg];
          v.x = (float)(x.x * l);
          v.y = (float)(x.y * l);
        }
        stage_a[u] = v;
      }
      {  // B(g, c) = Gr(g, l, b), c = l BA + b_local (contiguous gamma)
        const int g = e % KC, c = e / KC;
        const int l = c / BA, b = b0 + (c % BA), gg = g0 + g;
        cf v = {0.f, 0.f};
        if (b < chir && gg < chic) {
          const zc x = t.gr[gg + (size_t)chic * (l + (size_t)D * b)];
          v.x = (float)x.x;
          v.y = (float)x.y;
        }
        stage_b[u] = v;
      }
    }
  };
  load_chunk(0);
  for (int g0 = 0; g0 < chic; g0 += KC) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int e = tid + 256 * u;
      As[e / TR][e % TR] = stage_a[u];
      Bs[e % KC][e / KC] = stage_b[u];
    }
    __syncthreads();
    if (g0 + KC < chic) load_chunk(g0 + KC);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int kk = ks * 4 + lk;
#pragma unroll
      for (int s = 0; s < NBW; ++s) {
        const int q = wave + 4 * s;
        if (q < NB1 * NB1) {
          const int rb = q / NB1, cb = q % NB1;
          cmfma(As[kk][rb * 16 + lr16], Bs[kk][cb * 16 + lr16], re[s], im[s]);
        }
      }
    }
    __syncthreads();
  }
  // pre-gate tile into LDS: lane holds rows 4 lk + v, column lr16 of its 16 x 16 block
#pragma unroll
  for (int s = 0; s < NBW; ++s) {
    const int q = wave + 4 * s;
    if (q < NB1 * NB1) {
      const int rb = q / NB1, cb = q % NB1;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        cf o;
        o.x = re[s][v];
        o.y = im[s][v];
        T[(rb * 16 + 4 * lk + v) * LDT + cb * 16 + lr16] = o;
      }
    }
  }
  __syncthreads();

  // gate: O(ij, p) = sum_kl U(ij, kl) X(kl, p), X(kl, p) = T(k BA + al, l BA + bl).
  // p runs over alpha fastest (theta stored as is) or beta fastest (theta^dagger), so
  // consecutive lanes write consecutive elements of W.
  const bool tr = t.transposed != 0;
  for (int pb = wave; pb < NP / 16; pb += 4) {
    const int p = pb * 16 + lr16;
    const int al = tr ? p / BA : p % BA;
    const int bl = tr ? p % BA : p / BA;
#pragma unroll
    for (int mb = 0; mb < MP / 16; ++mb) {
      f4 ore = {0.f, 0.f, 0.f, 0.f}, oim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KP / 4; ++ks) {
        const int kl = ks * 4 + lk;
        cf x = {0.f, 0.f};
        if (kl < D2) x = T[((kl / D) * BA + al) * LDT + (kl % D) * BA + bl];
        const cf u = sU[(mb * 16 + lr16) * KP + kl];
        cmfma(u, x, ore, oim);
      }
      // lane holds O(ij = mb 16 + 4 lk + v, p)
      const int a = a0 + al, b = b0 + bl;
      if (a < chil && b < chir) {
        const double sc = t.ll[a] * t.lr[b];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ij = mb * 16 + 4 * lk + v;
          if (ij < D2) {
            const int i = ij / D, j = ij % D;
            const size_t row = (size_t)i * chil + a;  // theta row (i, alpha)
            const size_t col = (size_t)j * chir + b;  // theta col (j, beta)
            const double xr = (double)ore[v] * sc, xi = (double)oim[v] * sc;
            if (!tr) t.W[row + (size_t)t.R * col] = zmk(xr, xi);
            else t.W[col + (size_t)t.R * row] = zmk(xr, -xi);
          }
        }
      }
    }
  }
}

}  // namespace

hipError_t launch_theta_gate_mixed(const BondTask* d_tasks, int ntasks, int d, int max_chil, int max_chir,
                                   hipStream_t st) {
  auto go = [&](auto kern, int BA) {
    dim3 grid((max_chil + BA - 1) / BA, (max_chir + BA - 1) / BA, ntasks);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, d_tasks);
    return hipGetLastError();
  };
  switch (d) {
    case 2: return go(k_theta_gate_mixed<2>, MixCfg<2>::BA);
    case 3: return go(k_theta_gate_mixed<3>, MixCfg<3>::BA);
    case 4: return go(k_theta_gate_mixed<4>, MixCfg<4>::BA);
    case 5: return go(k_theta_gate_mixed<5>, MixCfg<5>::BA);
    case 6: return go(k_theta_gate_mixed<6>, MixCfg<6>::BA);
    case 8: return go(k_theta_gate_mixed<8>, MixCfg<8>::BA);
    default: return hipErrorInvalidValue;
  }
}
