// Column-pivoted Householder QR preconditioning of the bond SVD.
//
// The bond matrix W (Rb x Cb, Rb >= Cb: theta, or theta^dagger) is factorised as
// W P = Q R with block column pivoting and stopped at the numerical rank r, where
// every remaining column norm is below tol_rel * ||W||_F.  Each panel of QB = 16
// columns takes the QB columns of largest remaining norm (exact norms, recomputed by
// every trailing update) in descending order; the panel itself is factorised without
// further pivoting and applied to the rest as one compact-WY block (BLAS3).  On
// late-time Fermi-Hubbard bonds this ordering gives the Jacobi the same few sweeps
// as column-by-column pivoting (xGEQP3).  The one-sided Jacobi then runs on
// X = R(0:r, :)^dagger (Cb x r), whose columns are graded and nearly orthogonal:
// on late-time Fermi-Hubbard bonds (1024 x 1024, rank ~830) it converges in 4
// sweeps instead of ~22 for W itself, and the sweeps cover r instead of Cb columns.
// With X V = U_x S:  W = (Q V) S (P U_x)^dagger, so the Rb-side singular vectors
// are Q V and the Cb-side ones P U_x.
//
//   k_qrp_colnorms     column norms (one wave per column)
//   k_qrp_begin        ||W||_F, identity pivot, rank := running
//   k_qrp_panel_reg    panel j0 for Rb <= 1024: the QB largest remaining columns are
//                      gathered into registers (512 threads, one or two rows each),
//                      factorised there (one 32-value reduce-scatter per reflector
//                      gives the reflector's products with the later panel columns
//                      and the earlier reflectors at once) and stored in place, with
//                      the displaced columns moved to the vacated slots; the
//                      compact-WY factor T of the panel is kept (panel p at Tq + p QB^2)
//   k_qrp_panel        the same for any Rb, through global memory (256 threads)
//   k_qrp_trailing_reg trailing update A2 <- Q_panel^H A2 on FP64 MFMA and the exact
//                      remaining norms for the next panel's choice, 16 columns per
//                      workgroup held in registers between the two products (Rb <= 1024)
//   k_qrp_reflect<0>   the same for any Rb, reading the trailing block twice
//   k_qrp_build_x      X = R(0:r, :)^dagger, zero padding columns
//   k_qrp_vectors_init Cb-side vectors P X(:, kept)/s and the start [V(:, kept); 0] of the
//                      Rb-side ones, or the identity for the explicit Q
//   k_qrp_reflect<1,2> Q applied to those columns as the stored panels' block
//                      reflectors, last panel first: Y <- Y - V_p (T_p (V_p^H Y));
//                      k_qrp_reflect_pair<1,2> two panels per pass
#include <type_traits>

#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

constexpr int QB = 16;          // panel width
constexpr int kMaxCols = 4096;  // largest Cb (d chi_max <= 4096)
constexpr int PT = 512;         // threads of the register-resident panel kernel
constexpr int PW = PT / 64;     // its waves

// Calls f(std::integral_constant<int, i>) for i = 0 .. N-1 in order (a loop whose index is
// a compile-time constant in every iteration).
template <int I, int N, class F>
__device__ __forceinline__ void static_for_impl(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for_impl<I + 1, N>(f);
  }
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl<0, N>(f);
}

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

// Squared column norms of the bond matrix: one wave per column, 4 columns per workgroup.
__global__ __launch_bounds__(256) void k_qrp_colnorms(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = 4 * blockIdx.x + wave;
  if (k >= t.Cb) return;
  const zc* col = t.Wq + (size_t)t.Rb * k;
  double s = 0.0;
  for (int i = lane; i < t.Rb; i += 64) s += zabs2(col[i]);
  s = wave_sum(s);
  if (lane == 0) t.qnrm[k] = s;
}

// After k_qrp_colnorms: identity pivot, ||W||_F^2 (fixed-order sum), stop threshold.
__global__ __launch_bounds__(256) void k_qrp_begin(const BondTask* __restrict__ tasks, double tol_rel) {
  const BondTask t = tasks[blockIdx.x];
  __shared__ double red[4];
  for (int k = threadIdx.x; k < t.Cb; k += 256) t.qpiv[k] = k;
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

// Per-wave partial sums of D = L^H R ((16 NP) x (16 NQ) as NP x NQ blocks of 16x16)
// over rows r < K: wave w of nw takes rows 8 (w + nw it) .. +7 in two MFMA k-steps of
// 4 rows; lload(i, c) and rload(i, c) return element (i, c) of L and R (zero outside).
template <int NP, int NQ, class LL, class RL>
__device__ __forceinline__ void gram_partial(LL lload, RL rload, int K, int wave, int nw, int lane,
                                             d4 (&dre)[NP][NQ], d4 (&dim)[NP][NQ]) {
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < NP; ++a)
#pragma unroll
    for (int b = 0; b < NQ; ++b) {
      dre[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      dim[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  for (int r0 = 8 * wave; r0 < K; r0 += 8 * nw) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = r0 + 4 * h + lk;
      zc l[NP], r[NQ];
#pragma unroll
      for (int pb = 0; pb < NP; ++pb) l[pb] = lload(i, 16 * pb + lr);
#pragma unroll
      for (int qb = 0; qb < NQ; ++qb) r[qb] = rload(i, 16 * qb + lr);
#pragma unroll
      for (int pb = 0; pb < NP; ++pb)
#pragma unroll
        for (int qb = 0; qb < NQ; ++qb) {
          dre[pb][qb] = mfma4(l[pb].x, r[qb].x, dre[pb][qb]);
          dre[pb][qb] = mfma4(l[pb].y, r[qb].y, dre[pb][qb]);
          dim[pb][qb] = mfma4(l[pb].x, r[qb].y, dim[pb][qb]);
          dim[pb][qb] = mfma4(-l[pb].y, r[qb].x, dim[pb][qb]);
        }
    }
  }
}

// Fixed-order sum of the NW waves' partials into S (row stride lds; bitwise reproducible).
template <int NP, int NQ, int NW = 4>
__device__ __forceinline__ void gram_reduce(const d4 (&dre)[NP][NQ], const d4 (&dim)[NP][NQ],
                                            double (*red)[2][256], zc* S, int lds, int wave, int lane) {
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int pb = 0; pb < NP; ++pb)
#pragma unroll
    for (int qb = 0; qb < NQ; ++qb) {
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
          for (int w = 0; w < NW - 1; ++w) {
            xr += red[w][0][g * 64 + lane];
            xi += red[w][1][g * 64 + lane];
          }
          S[(16 * pb + lk + 4 * g) * lds + 16 * qb + lr] = zmk(xr, xi);
        }
      }
      __syncthreads();
    }
}

// T (nbp x nbp, upper) of the compact WY form H_0 .. H_{nbp-1} = I - V T V^H from the
// reflector scalars and G = V^H V:  T[q][q] = tau_q, T[0:q, q] = -tau_q T[0:q, 0:q] G[0:q, q].
// Row p of T only involves row p's own earlier entries (T is upper triangular), so
// thread p < QB forms row p in registers; written to the panel's slot Tp (column-major,
// zero-padded to QB x QB).  Called by every thread of the workgroup.
__device__ void form_t(const zc (*G)[QB + 1], const zc* stau, int nbp, zc* Tp) {
  __syncthreads();  // G and stau are complete
  if (threadIdx.x >= QB) return;
  const int pp = threadIdx.x;
  zc row[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) row[q] = zmk(0.0, 0.0);
  if (pp < nbp) {
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      if (q == pp) row[q] = stau[q];
      if (q > pp && q < nbp) {
        zc acc = zmk(0.0, 0.0);
#pragma unroll
        for (int kk = 0; kk < QB; ++kk)
          if (kk >= pp && kk < q) zfma(acc, row[kk], G[kk][q]);
        row[q] = zmul(zmk(-stau[q].x, -stau[q].y), acc);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < QB; ++q) Tp[pp + QB * q] = row[q];
}

// Householder scalars of the column (alpha; x) with |x|^2 = xn2 (xLARFG convention):
// H^H (alpha; x) = (beta; 0) with H = I - tau v v^H, v = (1; x * scale).  Returns false
// (tau = 0, H = I) when the column is already (real alpha; 0).
__device__ __forceinline__ bool householder(zc alpha, double xn2, zc& tau, double& beta, zc& scale) {
  tau = zmk(0.0, 0.0);
  if (xn2 == 0.0 && alpha.y == 0.0) return false;
  const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
  beta = alpha.x >= 0.0 ? -nrm : nrm;
  tau = zmk((beta - alpha.x) / beta, -alpha.y / beta);
  const zc am = zmk(alpha.x - beta, alpha.y);
  const double den = zabs2(am);
  scale = zmk(am.x / den, -am.y / den);
  return true;
}

// Selection of panel j0 by the NT threads of a workgroup: sel[0..nbp) := the nbp
// candidates (columns j0 + c, c < ncand, squared norms snrm[c] in LDS) of largest norm
// in descending order, ties to the lowest index; NaN norms are never taken.  *sstop :=
// 1 when even the largest is not above thr (the numerical rank is j0), 2 when fewer than
// nbp candidates have a norm, else 0.  Up to NT candidates, every thread ranks one
// candidate against all the others (one pass, broadcast LDS reads); beyond that, wave 0
// takes nbp argmax passes (snrm entries taken are overwritten).  snrm has kMaxCols + 1
// entries: the last receives the largest norm.  Ends with a barrier.
template <int NT>
__device__ void select_panel(double* snrm, int ncand, int nbp, double thr, int j0, int* sel, int* sstop, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
  if (tid < QB) sel[tid] = -1;
  __syncthreads();
  if (ncand <= NT) {
    if (tid < ncand) {
      const double v = snrm[tid];
      if (v == v) {
        int rank = 0;
        for (int k = 0; k < ncand; ++k) {
          const double u = snrm[k];
          rank += (u > v || (u == v && k < tid)) ? 1 : 0;
        }
        if (rank < nbp) sel[rank] = j0 + tid;
        if (rank == 0) snrm[kMaxCols] = v;
      }
    }
  } else if (wave == 0) {
    for (int q = 0; q < nbp; ++q) {
      double bv = -1.0;
      int bi = -1;
      for (int k = lane; k < ncand; k += 64) {
        const double v = snrm[k];
        if (v > bv) { bv = v; bi = k; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) { bv = ov; bi = oi; }
      }
      if (bi < 0) break;  // every lane holds the same winner
      if (lane == 0) sel[q] = j0 + bi;
      if (lane == 0 && q == 0) snrm[kMaxCols] = bv;
      if (lane == (bi & 63)) snrm[bi] = -1.0;  // a lane only ever reads the entries it clears
    }
  }
  __syncthreads();
  if (tid == 0) {
    int st = 0;
    if (sel[0] < 0 || !(snrm[kMaxCols] > thr)) st = 1;
    for (int q = 1; q < nbp && !st; ++q)
      if (sel[q] < 0) st = 2;
    *sstop = st;
  }
  __syncthreads();
}

// Column moves of a panel, by one wave (lane q < QB handles panel slot j0 + q and the
// q-th selected column): sel[0..nbp) (distinct columns >= j0) go to j0..j0+nbp-1 in
// that order; the columns of j0..j0+nbp-1 not selected go to the selected slots beyond
// the panel (both lists ascending), so every column lands in exactly one place.
// Updates qpiv and qnrm, writes the moves to mv_src/mv_dst and returns their number
// (the same in every lane).  Every global read is issued before any global write, and
// the reads are independent of each other.
__device__ int panel_moves(const BondTask& t, int j0, int nbp, const int* sel, int* mv_src, int* mv_dst,
                           int lane) {
  int sr[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) sr[q] = q < nbp ? sel[q] : -1;
  const int d = j0 + lane;                       // this lane's panel slot
  const int mine = lane < nbp ? sel[lane] : -1;  // this lane's selected column
  bool in = false;
  int rank_dst = 0;
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    in = in || sr[q] == d;
    if (sr[q] >= j0 + nbp && sr[q] < mine) ++rank_dst;
  }
  const bool is_src = lane < nbp && !in;           // an unselected panel column: it moves out
  const bool is_dst = lane < nbp && mine >= j0 + nbp;  // a selected column beyond the panel: it is vacated
  const unsigned long long bsrc = __ballot(is_src);
  const int nmv = __popcll(bsrc);
  const int rank_src = __popcll(bsrc & ((1ull << lane) - 1ull));
  const int piv_sel = lane < nbp ? t.qpiv[mine] : 0;
  const int piv_src = is_src ? t.qpiv[d] : 0;
  const double nrm_src = is_src ? t.qnrm[d] : 0.0;
  // the moved-out column of rank m fills the vacated slot of rank m
  int from = 0;
  if (is_dst) {
    unsigned long long b = bsrc;
    for (int m = 0; m < rank_dst; ++m) b &= b - 1ull;
    from = __builtin_ctzll(b);
  }
  const int piv_mv = __shfl(piv_src, from, 64);
  const double nrm_mv = __shfl(nrm_src, from, 64);
  if (lane < nbp) t.qpiv[d] = piv_sel;
  if (is_dst) {
    t.qpiv[mine] = piv_mv;
    t.qnrm[mine] = nrm_mv;
    mv_dst[rank_dst] = mine;
  }
  if (is_src) mv_src[rank_src] = d;
  return nmv;
}

// Rows [i] of the displaced columns moved to the vacated slots, four loads in flight
// (sources and destinations are disjoint column sets).
__device__ __forceinline__ void move_rows(zc* Wq, size_t ld, int i, int nmv, const int* mv_src, const int* mv_dst) {
  for (int m0 = 0; m0 < nmv; m0 += 4) {
    zc v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (m0 + u < nmv) v[u] = Wq[i + ld * mv_src[m0 + u]];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (m0 + u < nmv) Wq[i + ld * mv_dst[m0 + u]] = v[u];
  }
}

// Panel j0 with the panel in registers (Rb <= RPT * PT; one workgroup of PT threads per
// task, thread t owns rows t + PT s, s < RPT):
//  1. wave 0 picks the nbp = min(QB, kmax - j0) columns of largest remaining norm in
//     descending order (ties: lowest index); if even the largest is below the stop
//     threshold the numerical rank is j0 and the factorisation ends;
//  2. the picked columns are loaded into registers; the displaced ones are moved;
//  3. unpivoted Householder QR of the panel rows j0.. (xLARFG convention).  For
//     reflector q one fixed-order reduction yields w_k = v_q^H a_k (k > q, for the
//     update) and G[p][q] = v_p^H v_q (p < q, for T);
//  4. the panel is stored back and T formed.
template <int RPT>
__global__ __launch_bounds__(PT) void k_qrp_panel_reg(const BondTask* __restrict__ tasks, int j0) {
  const BondTask t = tasks[blockIdx.x];
  if (t.qstate[0] >= 0.0) return;
  const int Rb = t.Rb, Cb = t.Cb;
  const int kmax = Rb < Cb ? Rb : Cb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (j0 >= kmax) {
    if (tid == 0) t.qstate[0] = (double)kmax;
    return;
  }
  const int nbp = (kmax - j0) < QB ? (kmax - j0) : QB;
  const int ncand = Cb - j0;
  const size_t ld = Rb;
  __shared__ double snrm[kMaxCols + 1];
  __shared__ int sel[QB], mv_src[QB], mv_dst[QB];
  __shared__ int snmv, sstop;
  __shared__ double red[PW][32];
  __shared__ double wsum[32];
  __shared__ double rnorm[PW];
  __shared__ zc salpha;
  __shared__ zc G[QB][QB + 1];
  __shared__ zc stau[QB];

  // 1. selection
  for (int k = tid; k < ncand; k += PT) snrm[k] = t.qnrm[j0 + k];
  select_panel<PT>(snrm, ncand, nbp, t.qstate[1], j0, sel, &sstop, tid);
  if (sstop) {
    if (tid == 0) t.qstate[0] = sstop == 1 ? (double)j0 : 0.0;
    return;
  }
  if (wave == 0) {
    const int n = panel_moves(t, j0, nbp, sel, mv_src, mv_dst, lane);
    if (lane == 0) snmv = n;
  }
  __syncthreads();

  // 2. gather the panel; each thread moves its own rows of the displaced columns
  zc a[RPT][QB];
  const int nmv = snmv;
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int i = tid + PT * s;
#pragma unroll
    for (int q = 0; q < QB; ++q) a[s][q] = (i < Rb && q < nbp) ? t.Wq[i + ld * sel[q]] : zmk(0.0, 0.0);
    if (i < Rb) move_rows(t.Wq, ld, i, nmv, mv_src, mv_dst);
  }

  // 3. Householder QR of the panel, the column loop unrolled so that every register
  // index is static: column q sits in a[.][q], first as data, then as reflector v_q
  // (rows > j0 + q) below R(:, q).
  static_for<QB>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if (q >= nbp) return;
    const int j = j0 + q;
    double part = 0.0;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int i = tid + PT * s;
      if (i > j && i < Rb) part += zabs2(a[s][q]);
      if (i == j) salpha = a[s][q];
    }
    part = wave_sum(part);
    if (lane == 0) rnorm[wave] = part;
    __syncthreads();
    double xn2 = 0.0;
#pragma unroll
    for (int w = 0; w < PW; ++w) xn2 += rnorm[w];
    zc tau, scale = zmk(1.0, 0.0);
    double beta = 0.0;
    const bool refl = householder(salpha, xn2, tau, beta, scale);
    // v_q (rows > j; v_j = 1 implicitly) and R(j, j)
    if (refl) {
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int i = tid + PT * s;
        if (i > j && i < Rb) a[s][q] = zmul(a[s][q], scale);
        if (i == j) a[s][q] = zmk(beta, 0.0);
      }
    }
    // slot m = 1 .. 15 - q: conj(v_q) a_(q+m) (the later columns, for the update);
    // slot m > 15 - q: conj(v_p) v_q, p = m - 16 + q (the earlier reflectors, for T)
    double x[2 * QB];
#pragma unroll
    for (int e = 0; e < 2 * QB; ++e) x[e] = 0.0;
    // v_q on this thread's rows: 1 at row j, zero outside rows j .. Rb - 1, so that the
    // products and the update below need no row test
    zc vq[RPT];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int i = tid + PT * s;
      vq[s] = (i == j) ? zmk(1.0, 0.0) : ((i > j && i < Rb) ? a[s][q] : zmk(0.0, 0.0));
    }
#pragma unroll
    for (int s = 0; s < RPT; ++s)
#pragma unroll
      for (int m = 1; m < QB; ++m) {
        const bool later = m <= QB - 1 - q;  // earlier reflectors: rows >= j > j0 + p
        const zc other = a[s][later ? q + m : m - QB + q];
        const zc l = later ? vq[s] : other;
        const zc r = later ? other : vq[s];
        x[2 * m] = fma(l.x, r.x, fma(l.y, r.y, x[2 * m]));
        x[2 * m + 1] = fma(l.x, r.y, fma(-l.y, r.x, x[2 * m + 1]));
      }
    // reduce-scatter of the 32 values over the wave: lane l ends with value
    // 16 b5 + 8 b4 + 4 b3 + 2 b2 + b1 (b = bits of l), summed over all 64 lanes
    {
      double y16[16], y8[8], y4[4], y2[2];
      const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4, h1 = lane & 2;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const double lo = x[m], hi = x[m + 16];
        y16[m] = (h5 ? hi : lo) + __shfl_xor(h5 ? lo : hi, 32, 64);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const double lo = y16[m], hi = y16[m + 8];
        y8[m] = (h4 ? hi : lo) + __shfl_xor(h4 ? lo : hi, 16, 64);
      }
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const double lo = y8[m], hi = y8[m + 4];
        y4[m] = (h3 ? hi : lo) + __shfl_xor(h3 ? lo : hi, 8, 64);
      }
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const double lo = y4[m], hi = y4[m + 2];
        y2[m] = (h2 ? hi : lo) + __shfl_xor(h2 ? lo : hi, 4, 64);
      }
      double y1 = (h1 ? y2[1] : y2[0]) + __shfl_xor(h1 ? y2[0] : y2[1], 2, 64);
      y1 += __shfl_xor(y1, 1, 64);
      const int idx = 16 * ((lane >> 5) & 1) + 8 * ((lane >> 4) & 1) + 4 * ((lane >> 3) & 1) +
                      2 * ((lane >> 2) & 1) + ((lane >> 1) & 1);
      if ((lane & 1) == 0) red[wave][idx] = y1;
    }
    __syncthreads();
    if (tid < 32) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < PW; ++w) sum += red[w][tid];
      wsum[tid] = sum;
    }
    __syncthreads();
    if (tid == 0) {
      for (int m = QB - q; m < QB; ++m) G[m - QB + q][q] = zmk(wsum[2 * m], wsum[2 * m + 1]);
      stau[q] = tau;
      t.tau[j] = tau;
    }
    if (refl) {
      const zc ctau = zconj(tau);
#pragma unroll
      for (int m = 1; m <= QB - 1 - q; ++m) {
        const zc f = zmul(ctau, zmk(wsum[2 * m], wsum[2 * m + 1]));
#pragma unroll
        for (int s = 0; s < RPT; ++s) a[s][q + m] = zsub(a[s][q + m], zmul(f, vq[s]));
      }
    }
  });

  // 4. store the panel (R above and on the diagonal, reflectors below) and T
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int i = tid + PT * s;
    if (i < Rb) {
#pragma unroll
      for (int c = 0; c < QB; ++c)
        if (c < nbp) t.Wq[i + ld * (j0 + c)] = a[s][c];
    }
  }
  form_t(G, stau, nbp, t.Tq + (size_t)(j0 / QB) * QB * QB);
  if (tid == 0 && j0 + nbp >= kmax) t.qstate[0] = (double)kmax;
}

// Panel j0 through global memory (any Rb; one workgroup of 256 threads per task): the
// same selection, column moves, factorisation and T as k_qrp_panel_reg, with the
// reflector applied column by column (one wave per column).
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
  const int ncand = t.Cb - j0;
  const size_t ld = t.Rb;
  __shared__ double snrm[kMaxCols + 1];
  __shared__ int sel[QB], mv_src[QB], mv_dst[QB];
  __shared__ int snmv, sstop;
  __shared__ double red[4];
  __shared__ zc G[QB][QB + 1];
  __shared__ zc stau[QB];
  __shared__ double gred[3][2][256];
  for (int k = tid; k < ncand; k += 256) snrm[k] = t.qnrm[j0 + k];
  select_panel<256>(snrm, ncand, nbp, t.qstate[1], j0, sel, &sstop, tid);
  if (sstop) {
    if (tid == 0) t.qstate[0] = sstop == 1 ? (double)j0 : 0.0;
    return;
  }
  if (wave == 0) {
    const int n = panel_moves(t, j0, nbp, sel, mv_src, mv_dst, lane);
    if (lane == 0) snmv = n;
  }
  __syncthreads();
  // row by row: gather the selected columns, move the displaced ones to the vacated
  // slots, then write the selected ones into the panel
  for (int i = tid; i < t.Rb; i += 256) {
    zc v[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) v[q] = q < nbp ? t.Wq[i + ld * sel[q]] : zmk(0.0, 0.0);
    move_rows(t.Wq, ld, i, snmv, mv_src, mv_dst);
#pragma unroll
    for (int q = 0; q < QB; ++q)
      if (q < nbp) t.Wq[i + ld * (j0 + q)] = v[q];
  }
  __syncthreads();
  for (int q = 0; q < nbp; ++q) {
    const int j = j0 + q;
    zc* cj = t.Wq + ld * j;
    double part = 0.0;
    for (int i = j + 1 + tid; i < t.Rb; i += 256) part += zabs2(cj[i]);
    const double xn2 = block_sum(part, red);
    const zc alpha = cj[j];
    zc tau, scale = zmk(1.0, 0.0);
    double beta = 0.0;
    const bool refl = householder(alpha, xn2, tau, beta, scale);
    __syncthreads();  // every thread has read alpha
    if (refl) {
      for (int i = j + 1 + tid; i < t.Rb; i += 256) cj[i] = zmul(cj[i], scale);
      if (tid == 0) cj[j] = zmk(beta, 0.0);
    }
    if (tid == 0) {
      t.tau[j] = tau;
      stau[q] = tau;
    }
    __syncthreads();
    if (refl) {
      const zc ctau = zconj(tau);
      for (int k = j + 1 + wave; k < j0 + nbp; k += 4) {
        zc* a = t.Wq + ld * k;
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
  {
    const int K = t.Rb - j0;
    const zc* P = t.Wq + ld * j0 + j0;
    auto vload = [&](int i, int q) -> zc {
      if (q >= nbp || i < q || i >= K) return zmk(0.0, 0.0);
      if (i == q) return zmk(1.0, 0.0);
      return P[i + ld * q];
    };
    d4 gre[1][1], gim[1][1];
    gram_partial<1, 1>(vload, vload, K, wave, 4, lane, gre, gim);
    gram_reduce<1, 1>(gre, gim, gred, &G[0][0], QB + 1, wave, lane);
  }
  form_t(G, stau, nbp, t.Tq + (size_t)(j0 / QB) * QB * QB);
  if (tid == 0 && j0 + nbp >= kmax) t.qstate[0] = (double)kmax;
}

// Block reflector of one panel on 32 columns per workgroup (FP64 MFMA):
//   MODE 0  trailing update of panel j0 = arg: A2 <- Q_p^H A2 = A2 - V (T^H (V^H A2)) on the
//           columns right of the panel, and the exact remaining norms |A2(j0+nbp:, c)|^2
//           into qnrm;
//   MODE 1  panel p = arg of the factor applied to the Rb-side vectors Yr (k columns):
//           Y <- Q_p Y = Y - V (T (V^H Y));
//   MODE 2  the same on the explicit factor Qs (r columns, identity start: column
//           blocks entirely left of the panel are still unit vectors and are skipped).
// V has rows j0.. with an implicit unit diagonal; only rows j0.. of the target change.
// Eight waves: a launch has few workgroups (one per 32 target columns of each bond),
// so each workgroup's stream over the rows is what the launch waits for.
constexpr int kReflWaves = 8;
constexpr int kReflThreads = 64 * kReflWaves;

template <int MODE>
__global__ __launch_bounds__(kReflThreads) void k_qrp_reflect(const BondTask* __restrict__ tasks, int arg) {
  const BondTask t = tasks[blockIdx.y];
  int j0, nbp, kb, nc;
  const size_t ld = t.Rb;
  zc* A2;
  if (MODE == 0) {
    j0 = arg;
    const double rank = t.qstate[0];
    if (rank >= 0.0 && rank <= (double)j0) return;
    const int kmax = t.Rb < t.Cb ? t.Rb : t.Cb;
    if (j0 >= kmax) return;
    nbp = (kmax - j0) < QB ? (kmax - j0) : QB;
    kb = j0 + nbp + 32 * blockIdx.x;
    if (kb >= t.Cb) return;
    nc = (t.Cb - kb) < 32 ? (t.Cb - kb) : 32;
    A2 = t.Wq + ld * kb + j0;
  } else {
    if (MODE == 1 && !t.qr) return;
    const int r = t.C;
    j0 = QB * arg;
    if (j0 >= r) return;
    nbp = (r - j0) < QB ? (r - j0) : QB;
    const int ncols = MODE == 2 ? r : (int)t.result[0];
    kb = 32 * blockIdx.x;
    if (kb >= ncols || (MODE == 2 && kb + 31 < j0)) return;
    nc = (ncols - kb) < 32 ? (ncols - kb) : 32;
    A2 = (MODE == 2 ? t.Qs : t.Yr) + ld * kb + j0;
  }
  const int K = t.Rb - j0;
  const zc* P = t.Wq + ld * j0 + j0;  // P[i + ld q] = A(j0 + i, j0 + q)
  const zc* Tp = t.Tq + (size_t)(j0 / QB) * QB * QB;
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
  __shared__ double red[kReflWaves - 1][2][256];
  __shared__ zc Y[QB][33];
  __shared__ zc Ts[QB][QB + 1];
  __shared__ double nsum[kReflWaves][32];
  for (int e = tid; e < QB * QB; e += kReflThreads) Ts[e % QB][e / QB] = Tp[e];
  // ---- Y = V^H A2 (16 x 32), rows split over the waves
  {
    d4 yre[1][2], yim[1][2];
    gram_partial<1, 2>(vload, aload, K, wave, kReflWaves, lane, yre, yim);
    gram_reduce<1, 2, kReflWaves>(yre, yim, red, &Y[0][0], 33, wave, lane);
  }
  // ---- Z = T^H Y (MODE 0) or T Y, held as MFMA A operands Z[4ks + lk][16ct + lr]
  double zre[2][4], zim[2][4];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int pr = 4 * ks + lk, qc = 16 * ct + lr;
      zc acc = zmk(0.0, 0.0);
      if (MODE == 0) {
        for (int kk = 0; kk <= pr; ++kk) zfmac(acc, Ts[kk][pr], Y[kk][qc]);
      } else {
        for (int kk = pr; kk < QB; ++kk) zfma(acc, Ts[pr][kk], Y[kk][qc]);
      }
      zre[ct][ks] = acc.x;
      zim[ct][ks] = acc.y;
    }
  // ---- A2 -= V Z, as (V Z)^T = Z^T V^T: A operand Z^T, B operand V^T
  double ns[2][4];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int g = 0; g < 4; ++g) ns[ct][g] = 0.0;
  for (int r0 = 16 * wave; r0 < K; r0 += 16 * kReflWaves) {
    const int i = r0 + lr;
    zc x[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) x[ks] = vload(i, 4 * ks + lk);
    d4 ore[2], oim[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      ore[ct] = d4{0.0, 0.0, 0.0, 0.0};
      oim[ct] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
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
            if (MODE == 0 && i >= nbp) ns[ct][g] += zabs2(a);
          }
        }
    }
  }
  if (MODE != 0) return;
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
  if (tid < nc) {
    double v = 0.0;
    for (int w = 0; w < kReflWaves; ++w) v += nsum[w][tid];
    t.qnrm[kb + tid] = v;
  }
}

// Trailing update of panel j0 (as k_qrp_reflect<0>) on 16 columns per workgroup, with the
// block held in registers between its two products, so that it is read once and written
// once instead of read twice: each wave takes RB blocks of 16 rows, and lane (lr, lk) keeps
// A2(r0 + 4g + lk, lr) of every block r0.  That is both the B operand of Y = V^H A2 (rows
// as the MFMA's k) and the layout of the result of V Z (rows as its m), so the update
// A2 -= V (T^H Y) and the remaining column norms come out where the loads went.
// K = Rb - j0 <= 128 RB rows.
template <int RB>
__global__ __launch_bounds__(kReflThreads) void k_qrp_trailing_reg(const BondTask* __restrict__ tasks, int j0) {
  const BondTask t = tasks[blockIdx.y];
  const double rank = t.qstate[0];
  if (rank >= 0.0 && rank <= (double)j0) return;
  const int kmax = t.Rb < t.Cb ? t.Rb : t.Cb;
  if (j0 >= kmax) return;
  const int nbp = (kmax - j0) < QB ? (kmax - j0) : QB;
  const int kb = j0 + nbp + 16 * blockIdx.x;
  if (kb >= t.Cb) return;
  const int nc = (t.Cb - kb) < 16 ? (t.Cb - kb) : 16;
  const size_t ld = t.Rb;
  zc* A2 = t.Wq + ld * kb + j0;
  const int K = t.Rb - j0;
  const zc* P = t.Wq + ld * j0 + j0;  // P[i + ld q] = A(j0 + i, j0 + q)
  const zc* Tp = t.Tq + (size_t)(j0 / QB) * QB * QB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  auto vload = [&](int i, int q) -> zc {
    if (q >= nbp || i < q || i >= K) return zmk(0.0, 0.0);
    if (i == q) return zmk(1.0, 0.0);
    return P[i + ld * q];
  };
  __shared__ double red[kReflWaves - 1][2][256];
  __shared__ zc Y[QB][17];
  __shared__ zc Ts[QB][QB + 1];
  __shared__ double nsum[kReflWaves][16];
  for (int e = tid; e < QB * QB; e += kReflThreads) Ts[e % QB][e / QB] = Tp[e];
  // ---- Y = V^H A2 (16 x 16): A operand conj(V)(q = lr, i), B operand A2(i, c = lr)
  zc a[RB][4];
  d4 yre[1][1], yim[1][1];
  yre[0][0] = d4{0.0, 0.0, 0.0, 0.0};
  yim[0][0] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    const int r0 = 16 * (wave + kReflWaves * b);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int i = r0 + 4 * g + lk;
      const zc v = vload(i, lr);
      const zc x = (lr < nc && i < K) ? A2[i + ld * lr] : zmk(0.0, 0.0);
      a[b][g] = x;
      yre[0][0] = mfma4(v.x, x.x, yre[0][0]);
      yre[0][0] = mfma4(v.y, x.y, yre[0][0]);
      yim[0][0] = mfma4(v.x, x.y, yim[0][0]);
      yim[0][0] = mfma4(-v.y, x.x, yim[0][0]);
    }
  }
  gram_reduce<1, 1, kReflWaves>(yre, yim, red, &Y[0][0], 17, wave, lane);
  // ---- Z = T^H Y, as the B operands Z(4 ks + lk, lr) of the second product
  zc zb[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int pr = 4 * ks + lk;
    zc acc = zmk(0.0, 0.0);
    for (int kk = 0; kk <= pr; ++kk) zfmac(acc, Ts[kk][pr], Y[kk][lr]);
    zb[ks] = acc;
  }
  // ---- A2 -= V Z: A operand V(r0 + lr, q), B operand Z(q, c); element gg of the result
  // in lane (lr, lk) is row r0 + lk + 4 gg, column lr: the register a[b][gg]
  double ns = 0.0;
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    const int r0 = 16 * (wave + kReflWaves * b);
    d4 ore = d4{0.0, 0.0, 0.0, 0.0}, oim = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const zc va = vload(r0 + lr, 4 * ks + lk);
      ore = mfma4(va.x, zb[ks].x, ore);
      ore = mfma4(-va.y, zb[ks].y, ore);
      oim = mfma4(va.x, zb[ks].y, oim);
      oim = mfma4(va.y, zb[ks].x, oim);
    }
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const int i = r0 + lk + 4 * gg;
      if (lr < nc && i < K) {
        const zc x = zsub(a[b][gg], zmk(ore[gg], oim[gg]));
        A2[i + ld * lr] = x;
        if (i >= nbp) ns += zabs2(x);
      }
    }
  }
  // column norms: over the four lanes of equal lr, then over the waves
  ns += __shfl_xor(ns, 16, 64);
  ns += __shfl_xor(ns, 32, 64);
  if (lk == 0) nsum[wave][lr] = ns;
  __syncthreads();
  if (tid < nc) {
    double v = 0.0;
    for (int w = 0; w < kReflWaves; ++w) v += nsum[w][tid];
    t.qnrm[kb + tid] = v;
  }
}

// Two stored panels applied at once (MODE 1: to the Rb-side vectors Yr, MODE 2: to the
// explicit factor Qs): Y <- Q_{p-1} Q_p Y = Y - V (T (V^H Y)) with the 32-reflector
// compact-WY block V = [V_{p-1} V_p] (columns j0 .. j0 + 31 of the factor, j0 the first
// column of panel p - 1) and T = [[T_{p-1}, -T_{p-1} (V_{p-1}^H V_p) T_p], [0, T_p]].
// Each target column block is streamed once per two panels instead of once per panel;
// the arithmetic is that of k_qrp_reflect<MODE> on the aggregated block.  arg = p.
template <int MODE>
__global__ __launch_bounds__(kReflThreads) void k_qrp_reflect_pair(const BondTask* __restrict__ tasks, int arg) {
  static_assert(MODE == 1 || MODE == 2, "pairs of panels only for the vector passes");
  const BondTask t = tasks[blockIdx.y];
  if (MODE == 1 && !t.qr) return;
  const int r = t.C;
  const int j0 = QB * (arg - 1);  // first column of panel p - 1
  if (j0 >= r) return;
  const int nb1 = (r - j0) < QB ? (r - j0) : QB;                    // reflectors of panel p - 1
  const int nb2 = (r - j0 - QB) <= 0 ? 0 : ((r - j0 - QB) < QB ? (r - j0 - QB) : QB);  // of panel p
  const int ncols = MODE == 2 ? r : (int)t.result[0];
  const int kb = 32 * blockIdx.x;
  if (kb >= ncols || (MODE == 2 && kb + 31 < j0)) return;
  const int nc = (ncols - kb) < 32 ? (ncols - kb) : 32;
  const size_t ld = t.Rb;
  zc* A2 = (MODE == 2 ? t.Qs : t.Yr) + ld * kb + j0;
  const int K = t.Rb - j0;
  const zc* P = t.Wq + ld * j0 + j0;  // P[i + ld q] = A(j0 + i, j0 + q)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  auto valid = [&](int q) { return q < QB ? q < nb1 : (q - QB) < nb2; };
  auto vload = [&](int i, int q) -> zc {
    if (!valid(q) || i < q || i >= K) return zmk(0.0, 0.0);
    if (i == q) return zmk(1.0, 0.0);
    return P[i + ld * q];
  };
  auto aload = [&](int i, int c) -> zc {
    if (c >= nc || i >= K) return zmk(0.0, 0.0);
    return A2[i + ld * c];
  };
  __shared__ double red[kReflWaves - 1][2][256];
  __shared__ zc Y[2 * QB][33];
  __shared__ zc G[QB][QB + 1];
  __shared__ zc Wt[QB][QB + 1];
  __shared__ zc Ts[2 * QB][2 * QB + 1];
  const zc* T1 = t.Tq + (size_t)(arg - 1) * QB * QB;
  const zc* T2 = t.Tq + (size_t)arg * QB * QB;
  for (int e = tid; e < QB * QB; e += kReflThreads) {
    const int a = e % QB, b = e / QB;
    Ts[a][b] = T1[e];
    Ts[QB + a][QB + b] = nb2 > 0 ? T2[e] : zmk(0.0, 0.0);
    Ts[QB + a][b] = zmk(0.0, 0.0);
  }
  // ---- Y = V^H A2 (32 x 32) and G = V_{p-1}^H V_p (16 x 16), rows split over the waves;
  // G from the same loads of V as Y (its columns 0-15 and 16-31 are V_{p-1} and V_p)
  {
    d4 yre[2][2], yim[2][2], gre[1][1], gim[1][1];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        yre[a][b] = d4{0.0, 0.0, 0.0, 0.0};
        yim[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      }
    gre[0][0] = d4{0.0, 0.0, 0.0, 0.0};
    gim[0][0] = d4{0.0, 0.0, 0.0, 0.0};
    for (int r0 = 8 * wave; r0 < K; r0 += 8 * kReflWaves) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = r0 + 4 * h + lk;
        zc l[2], rr[2];
#pragma unroll
        for (int pb = 0; pb < 2; ++pb) l[pb] = vload(i, 16 * pb + lr);
#pragma unroll
        for (int qb = 0; qb < 2; ++qb) rr[qb] = aload(i, 16 * qb + lr);
#pragma unroll
        for (int pb = 0; pb < 2; ++pb)
#pragma unroll
          for (int qb = 0; qb < 2; ++qb) {
            yre[pb][qb] = mfma4(l[pb].x, rr[qb].x, yre[pb][qb]);
            yre[pb][qb] = mfma4(l[pb].y, rr[qb].y, yre[pb][qb]);
            yim[pb][qb] = mfma4(l[pb].x, rr[qb].y, yim[pb][qb]);
            yim[pb][qb] = mfma4(-l[pb].y, rr[qb].x, yim[pb][qb]);
          }
        gre[0][0] = mfma4(l[0].x, l[1].x, gre[0][0]);
        gre[0][0] = mfma4(l[0].y, l[1].y, gre[0][0]);
        gim[0][0] = mfma4(l[0].x, l[1].y, gim[0][0]);
        gim[0][0] = mfma4(-l[0].y, l[1].x, gim[0][0]);
      }
    }
    gram_reduce<2, 2, kReflWaves>(yre, yim, red, &Y[0][0], 33, wave, lane);
    gram_reduce<1, 1, kReflWaves>(gre, gim, red, &G[0][0], QB + 1, wave, lane);
  }
  // ---- T12 = -T_{p-1} (G T_p), one entry per thread of the first 256
  {
    const int a = tid % QB, b = tid / QB;
    if (tid < QB * QB) {
      zc acc = zmk(0.0, 0.0);
      for (int m = 0; m < QB; ++m) zfma(acc, G[a][m], Ts[QB + m][QB + b]);
      Wt[a][b] = acc;
    }
    __syncthreads();
    if (tid < QB * QB) {
      zc acc = zmk(0.0, 0.0);
      for (int k = a; k < QB; ++k) zfma(acc, Ts[a][k], Wt[k][b]);
      Ts[a][QB + b] = zmk(-acc.x, -acc.y);
    }
    __syncthreads();
  }
  // ---- Z = T Y (T upper triangular), held as MFMA A operands Z[4ks + lk][16ct + lr]
  double zre[2][8], zim[2][8];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const int pr = 4 * ks + lk, qc = 16 * ct + lr;
      zc acc = zmk(0.0, 0.0);
      for (int kk = pr; kk < 2 * QB; ++kk) zfma(acc, Ts[pr][kk], Y[kk][qc]);
      zre[ct][ks] = acc.x;
      zim[ct][ks] = acc.y;
    }
  // ---- A2 -= V Z, as (V Z)^T = Z^T V^T: A operand Z^T, B operand V^T
  for (int r0 = 16 * wave; r0 < K; r0 += 16 * kReflWaves) {
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
            *ap = zsub(*ap, zmk(ore[ct][g], oim[ct][g]));
          }
        }
    }
  }
}

// X (Cb x Cpad, ld Cb) = R(0:r, :)^dagger; columns r..Cpad-1 zero.  R is the upper
// trapezoid of Wq.
__global__ void k_qrp_build_x(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  if (!t.qr) return;  // a task of the batch that skips the QR preconditioning
  const int r = t.C;  // the Jacobi task's column count is the numerical rank
  const size_t n = (size_t)t.Cb * t.Cpad;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(e % t.Cb), col = (int)(e / t.Cb);  // X(row, col) = conj(R(col, row))
    zc x = zmk(0.0, 0.0);
    if (col < r && col <= row) x = zconj(t.Wq[col + (size_t)t.Rb * row]);
    t.W[e] = x;
  }
}

// Column b (one workgroup) of the start of the singular vectors: for the kept b < k the
// Cb-side vector Z(qpiv[i], b) = X(i, perm[b]) / s[b] and the Rb-side start
// Y(:, b) = [V(0:r, perm[b]); 0] (FORM_Q = false), or Qs(:, b) = e_b, b < r (FORM_Q = true).
template <bool FORM_Q>
__global__ __launch_bounds__(256) void k_qrp_vectors_init(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.y];
  if (!FORM_Q && !t.qr) return;  // its vectors come from the Jacobi's W and V
  const int r = t.C;
  const int k = FORM_Q ? r : (int)t.result[0];
  const int b = blockIdx.x;
  if (b >= k) return;
  if (FORM_Q) {
    for (int i = threadIdx.x; i < t.Rb; i += 256) t.Qs[i + (size_t)t.Rb * b] = zmk(i == b ? 1.0 : 0.0, 0.0);
    return;
  }
  const int col = t.perm[b];
  const double inv = t.s[b] > 0.0 ? 1.0 / t.s[b] : 0.0;
  for (int i = threadIdx.x; i < t.Cb; i += 256)
    t.Zc[t.qpiv[i] + (size_t)t.Cb * b] = zscale(t.W[i + (size_t)t.Cb * col], inv);
  for (int i = threadIdx.x; i < t.Rb; i += 256)
    t.Yr[i + (size_t)t.Rb * b] = i < r ? t.V[i + (size_t)t.NV * col] : zmk(0.0, 0.0);
}

}  // namespace

hipError_t launch_qrp_begin(const BondTask* d_tasks, int ntasks, int max_cb, double tol_rel, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_colnorms, dim3((max_cb + 3) / 4, ntasks), dim3(256), 0, st, d_tasks);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_qrp_begin, dim3(ntasks), dim3(256), 0, st, d_tasks, tol_rel);
  return hipGetLastError();
}

hipError_t launch_qrp_panel(const BondTask* d_tasks, int ntasks, int max_rb, int max_cb, int j0, hipStream_t st) {
  if (max_cb > kMaxCols) return hipErrorInvalidValue;
  if (max_rb <= PT) hipLaunchKernelGGL(k_qrp_panel_reg<1>, dim3(ntasks), dim3(PT), 0, st, d_tasks, j0);
  else if (max_rb <= 2 * PT) hipLaunchKernelGGL(k_qrp_panel_reg<2>, dim3(ntasks), dim3(PT), 0, st, d_tasks, j0);
  else hipLaunchKernelGGL(k_qrp_panel, dim3(ntasks), dim3(256), 0, st, d_tasks, j0);
  return hipGetLastError();
}

hipError_t launch_qrp_trailing(const BondTask* d_tasks, int ntasks, int max_rb, int max_cb, int j0, hipStream_t st) {
  if (j0 + QB >= max_cb) return hipSuccess;
  const dim3 g16((max_cb - j0 - 1) / 16 + 1, ntasks);
  if (max_rb <= 256) hipLaunchKernelGGL(k_qrp_trailing_reg<2>, g16, dim3(kReflThreads), 0, st, d_tasks, j0);
  else if (max_rb <= 512) hipLaunchKernelGGL(k_qrp_trailing_reg<4>, g16, dim3(kReflThreads), 0, st, d_tasks, j0);
  else if (max_rb <= 1024) hipLaunchKernelGGL(k_qrp_trailing_reg<8>, g16, dim3(kReflThreads), 0, st, d_tasks, j0);
  else
    hipLaunchKernelGGL(k_qrp_reflect<0>, dim3((max_cb - j0 - 1) / 32 + 1, ntasks), dim3(kReflThreads), 0, st,
                       d_tasks, j0);
  return hipGetLastError();
}

hipError_t launch_qrp_build_x(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_build_x, dim3(256, ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

// The stored panels applied last to first: a single panel first when their number is
// odd, then pairs of panels as 32-reflector blocks.
template <int MODE>
hipError_t apply_panels(const BondTask* d_tasks, int ntasks, int ncols, int max_r, hipStream_t st) {
  const dim3 grid((ncols + 31) / 32, ntasks);
  int p = (max_r + QB - 1) / QB - 1;
  hipError_t e = hipSuccess;
  if (p >= 0 && (p + 1) % 2 == 1) {
    hipLaunchKernelGGL(k_qrp_reflect<MODE>, grid, dim3(kReflThreads), 0, st, d_tasks, p);
    e = hipGetLastError();
    --p;
  }
  for (; p >= 1 && e == hipSuccess; p -= 2) {
    hipLaunchKernelGGL(k_qrp_reflect_pair<MODE>, grid, dim3(kReflThreads), 0, st, d_tasks, p);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_qrp_vectors(const BondTask* d_tasks, int ntasks, int max_k, int max_r, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_vectors_init<false>, dim3(max_k, ntasks), dim3(256), 0, st, d_tasks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return apply_panels<1>(d_tasks, ntasks, max_k, max_r, st);
}

hipError_t launch_qrp_form_q(const BondTask* d_tasks, int ntasks, int max_r, hipStream_t st) {
  hipLaunchKernelGGL(k_qrp_vectors_init<true>, dim3(max_r, ntasks), dim3(256), 0, st, d_tasks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return apply_panels<2>(d_tasks, ntasks, max_r, max_r, st);
}
