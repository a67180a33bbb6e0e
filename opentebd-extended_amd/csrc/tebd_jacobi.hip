// One-sided block-Jacobi SVD of the bond matrix (replaces ZGESVD in SVDTruncation,
// local_operations_module.f90:322-345).
//
// W (R x Cpad, column-major) is orthogonalised in place and V accumulates the right
// rotations.  The columns form 2P blocks of JB; a round pairs every block with one
// other (round-robin "circle" order, 2P-1 rounds per sweep).  A round is two kernels:
//   k_jacobi_round   one workgroup per pair: the pair's Gram matrix on FP64 MFMA
//                    (v_mfma_f64_16x16x4f64), the decision whether it needs a
//                    rotation, and the 32x32 Hermitian eigenproblem by cyclic Jacobi
//                    in LDS -> Q (32x32, to global memory); rotated pairs are queued;
//   k_jacobi_rotate  [W; V](:, pair) <- [W; V](:, pair) Q for every queued pair, on
//                    MFMA, spread over 64-row slabs so a rotation costs one memory
//                    round trip of the whole GPU instead of one workgroup's stream.
// After a block has been part of a rotated (or verified) pair its columns are
// mutually orthogonal, so after the first round of a sweep only the 16x16 cross
// block and the 32 column norms are formed (4x less Gram work than the full 32x32).
// The first round of every sweep forms the full Gram of each pair: with a strongly
// graded spectrum the rounding of rotations lets the small columns of a block drift
// away from orthogonality, and that round restores it (a cross-only iteration does
// not converge there).
//
// Work that cannot change the result is skipped exactly:
//   * a task whose previous sweep rotated nothing has converged: its workgroups exit;
//   * a pair whose two blocks have not been rotated since the pair was last found
//     orthogonal would produce the same Gram and the same decision: it is skipped.
// Per task, counter[0..1] count rotated pairs of the current / previous sweep (by
// sweep parity), counter[2] pairs whose Gram was formed and counter[3] rotated pairs
// over the whole SVD; blkver[b] counts rotations of block b and pairseen[lo*2P+hi]
// holds (blkver[lo], blkver[hi]) from the last time that pair was verified.
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// D(16x16) += A(16x4) B(4x16).  Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// element g of the result in lane l is D[l/16 + 4g][l%16].
__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// k_jacobi_round runs 8 waves: all of them form the Gram matrix; in the inner solve
// waves 0-3 update one 2x2 block of A per thread and waves 4-7 the rotation
// accumulator Q, so that no wave issues both (a round costs about the longest wave's
// instruction stream).
constexpr int kRoundThreads = 512;
constexpr int kGramWaves = kRoundThreads / 64;

// Round-robin ("circle") schedule over nb = 2P blocks: block 0 stays, the others
// rotate; over nb-1 rounds every pair of blocks meets exactly once.
__device__ __forceinline__ void rr_pair(int nb, int round, int p, int& bi, int& bj) {
  auto pos = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + round) % (nb - 1); };
  bi = pos(p);
  bj = pos(nb - 1 - p);
}

// 1/sqrt(x) and 1/x to full FP64 precision: the hardware estimate refined by two
// Newton-Raphson steps (y <- y (3 - x y^2) / 2, r <- r (2 - x r)).
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * fma(-x, r, 2.0);
  return r * fma(-x, r, 2.0);
}

// Hermitian 2x2 Jacobi rotation J = [[c, s w], [-s conj(w), c]], w = g/|g|, that
// zeroes g in [[a, g], [conj(g), b]] via J^dagger H J, with the inner angle
// (|theta| <= pi/4, Rutishauser's smaller root).  With e = sqrt((b - a)^2 + 4|g|^2),
// cos 2theta = |b - a| / e and sin 2theta = 2|g| / e, and the half-angle formulas give
// c = sqrt((1 + cos 2theta) / 2) = (1 + cos 2theta) r and |s| = sin 2theta / (2c) =
// sin 2theta r with r = 1/sqrt(2 (1 + cos 2theta)); s takes the sign of b - a.  Every
// step is well conditioned (1 + cos 2theta lies in [1, 2]) and nothing overflows for tiny
// |g|.  The two first reciprocal square roots are independent, so the rotation's
// critical path holds two refined estimates.  Returns c, s w and t|g|, t = s / c =
// sign(b - a) sin 2theta / (1 + cos 2theta) = sign(b - a) sin 2theta 2 r^2: the rotated
// diagonal is (a - t|g|, b + t|g|).
__device__ __forceinline__ void jrot(double a, double b, zc g, double& c, zc& sw, double& tg) {
  const double g2 = zabs2(g);
  if (g2 == 0.0) { c = 1.0; sw = zmk(0.0, 0.0); tg = 0.0; return; }
  const double dd = b - a;
  const double inv = rsqrt_nr(g2);
  const double re = rsqrt_nr(fma(dd, dd, 4.0 * g2));
  const double h = fma(fabs(dd), re, 1.0);  // 1 + cos 2theta
  const double sin2 = (2.0 * g2 * inv) * re;
  const double r = rsqrt_nr(2.0 * h);
  c = h * r;
  const double s = dd < 0.0 ? -sin2 * r : sin2 * r;
  sw = zmk(s * (g.x * inv), s * (g.y * inv));
  tg = (dd < 0.0 ? -sin2 : sin2) * (2.0 * r * r) * (g2 * inv);
}

// Gram block D = X^H Y (16 x 16) of two 16-column blocks over rows [0, R), as
// per-wave partial sums of kGramWaves waves: wave w takes rows 8 (w + kGramWaves m)
// .. +7, m = 0, 1, ..., each in two MFMA k-steps of 4 rows, with the loads of two
// consecutive m issued before their MFMAs (the pass is bound by memory latency).  Also
// accumulates the lane's partial squared norms of X and Y.
__device__ __forceinline__ void gram_block(const zc* __restrict__ X, const zc* __restrict__ Y, int R, int wave,
                                           int lane, d4& gre, d4& gim, double& nx, double& ny) {
  constexpr int U = 2;
  const int lr = lane & 15, lk = lane >> 4;
  gre = d4{0.0, 0.0, 0.0, 0.0};
  gim = d4{0.0, 0.0, 0.0, 0.0};
  nx = 0.0;
  ny = 0.0;
  const zc* xc = X + (size_t)R * lr;
  const zc* yc = Y + (size_t)R * lr;
  constexpr int PER = 8 * kGramWaves;  // rows per period: 8 of them for each wave
  for (int r0 = 8 * wave; r0 < R; r0 += PER * U) {
    zc a[2 * U], b[2 * U];
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) {
      const int r = r0 + PER * (u >> 1) + 4 * (u & 1) + lk;
      a[u] = r < R ? xc[r] : zmk(0.0, 0.0);
      b[u] = r < R ? yc[r] : zmk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 2 * U; ++u) {
      // A[p][k] = conj(X[k][p]), B[k][q] = Y[k][q]
      gre = mfma4(a[u].x, b[u].x, gre);
      gre = mfma4(a[u].y, b[u].y, gre);
      gim = mfma4(a[u].x, b[u].y, gim);
      gim = mfma4(-a[u].y, b[u].x, gim);
      nx += zabs2(a[u]);
      ny += zabs2(b[u]);
    }
  }
}

// Sums the Gram waves' partial 16x16 blocks in a fixed order (bitwise reproducible)
// and stores rows [ro, ro+16) x cols [co, co+16) of A.  With `diag` the block is a
// diagonal block of the Hermitian A: only the upper triangle is taken and mirrored.
// The real and the imaginary parts go through the scratch `red` ((kGramWaves - 1) x 256
// doubles) one after the other, so that it fits in the rotation accumulator's LDS, which
// is free until the inner solve starts.  Called by every thread of the workgroup.
__device__ __forceinline__ void reduce_block(zc (*A)[JC + 1], double (*red)[256], d4 gre, d4 gim, int wave, int lane,
                                             int ro, int co, bool diag) {
  double xr[4];
  if (wave > 0 && wave < kGramWaves) {
#pragma unroll
    for (int g = 0; g < 4; ++g) red[wave - 1][g * 64 + lane] = gre[g];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      xr[g] = gre[g];
      for (int w = 0; w < kGramWaves - 1; ++w) xr[g] += red[w][g * 64 + lane];
    }
  }
  __syncthreads();
  if (wave > 0 && wave < kGramWaves) {
#pragma unroll
    for (int g = 0; g < 4; ++g) red[wave - 1][g * 64 + lane] = gim[g];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double xi = gim[g];
      for (int w = 0; w < kGramWaves - 1; ++w) xi += red[w][g * 64 + lane];
      const int p = (lane >> 4) + 4 * g, q = lane & 15;
      if (!diag) {
        A[ro + p][co + q] = zmk(xr[g], xi);
        A[co + q][ro + p] = zmk(xr[g], -xi);
      } else if (p < q) {
        A[ro + p][co + q] = zmk(xr[g], xi);
        A[co + q][ro + p] = zmk(xr[g], -xi);
      } else if (p == q) {
        A[ro + p][co + p] = zmk(xr[g], 0.0);
      }
    }
  }
  __syncthreads();
}

constexpr int kInnerSweeps = 30;

// Index pair t (t < JC/2) of inner round rd (rd < JC - 1), lower index first: rr_pair
// on JC indices, with its reduction mod JC - 1 done by one compare (the argument is
// below 2 (JC - 1)) instead of a general modulo, which costs several multiplies per
// use inside the inner rounds.
__device__ __forceinline__ void inner_pair(int rd, int t, int& p, int& q) {
  auto pos = [rd](int k) {
    int x = k - 1 + rd;
    x = x >= JC - 1 ? x - (JC - 1) : x;
    return k == 0 ? 0 : 1 + x;
  };
  const int a = pos(t), b = pos(JC - 1 - t);
  p = a < b ? a : b;
  q = a < b ? b : a;
}

// Cyclic Jacobi on A (JC x JC Hermitian, LDS); Q (LDS) accumulates the rotations.
// A round rotates 16 disjoint index pairs (round-robin order) in two phases:
//   phase 1  lanes 0..15 of wave 0 compute the round's rotations into one of two
//            buffers and write each rotated pair's own 2x2 block of A, which the
//            rotation makes diagonal: a' = a - t|g|, b' = b + t|g| (t = s / c), the
//            annihilated entries exactly zero.  Meanwhile waves 1..7 apply the previous
//            round's rotations to Q (Q_:p <- c Q_:p - s conj(w) Q_:q,
//            Q_:q <- s w Q_:p + c Q_:q; 512 (row, pair) groups), which nothing in this
//            round reads;
//   phase 2  the 240 threads of waves 0..3 update the off-diagonal blocks
//            [p_r, q_r] x [p_k, q_k], r < k, of A (J^H A J acts on each such block on
//            its own, so the update is in place): two threads per block, one per row,
//            each also writing the mirrored entries of block (k, r) (A is Hermitian).
// Blocks whose two pairs both stay put are skipped, and so is phase 2 of a round that
// rotates nothing.  The last round's Q update runs after the loop.  A Gram entry is
// accurate relative to sqrt(A_pp A_qq), so the solve uses a purely relative test (no
// rotation when |A_pq| <= 1e-15 sqrt(A_pp A_qq)) and ends when a sweep would rotate
// nothing.  That is decided by testing every pair once after a sweep: a sweep that
// rotates nothing leaves A unchanged, so it rotates nothing iff no pair passes the test
// now.
__device__ __forceinline__ bool needs_rotation(zc g, double a, double b) {
  const double g2 = zabs2(g);
  return !(g2 <= 1e-30 * fabs(a * b) || g2 == 0.0);
}

static_assert(kRoundThreads == JC * JC / 2, "one thread per group of Q after the last round");
constexpr int kOffBlocks = (JC / 2) * (JC / 2 - 1) / 2;  // 2x2 blocks above the diagonal
static_assert(2 * kOffBlocks <= 256, "phase 2 runs on waves 0..3");

// Rotations of one round, pair t: c, s w, and p | q << 8 | (rotates) << 16.
struct RotBuf {
  double c[JC / 2];
  zc sw[JC / 2];
  int pq[JC / 2];
};

// Applies the rotations of buffer rb to group grp = 16 k + t of Q: row k, pair t.
__device__ __forceinline__ void q_rotate(zc (*Q)[JC + 1], const RotBuf& rb, int grp) {
  const int k = grp >> 4, tp = grp & 15;
  const int pq = rb.pq[tp];
  const double c = rb.c[tp];
  const zc sw = rb.sw[tp];
  if (!(pq >> 16)) return;
  const int p = pq & 255, q = (pq >> 8) & 255;
  const zc xp = Q[k][p], xq = Q[k][q];
  Q[k][p] = zsub(zscale(xp, c), zmul(zconj(sw), xq));
  Q[k][q] = zadd(zmul(sw, xp), zscale(xq, c));
}

__device__ void inner_jacobi(zc (*A)[JC + 1], zc (*Q)[JC + 1], RotBuf* rb, int* s_round, int* s_sweep,
                             int* s_more, int tid, int max_sweeps) {
  for (int e = tid; e < JC * JC; e += kRoundThreads) {
    const int r = e % JC, c = e / JC;
    Q[r][c] = zmk(r == c ? 1.0 : 0.0, 0.0);
  }
  // phase 2 owner: block (tr, tk), tr < tk, of the upper triangle, row `half` of it
  const int ob = tid >> 1, half = tid & 1;
  int tr = 0, tk = 0;
  if (ob < kOffBlocks) {
    int rem = ob;
    while (rem >= JC / 2 - 1 - tr) {
      rem -= JC / 2 - 1 - tr;
      ++tr;
    }
    tk = tr + 1 + rem;
  }
  int g = 0;  // rounds started, over all sweeps: round g uses buffer g & 1
  for (int isw = 0; isw < max_sweeps; ++isw) {
    if (tid == 0) s_sweep[isw & 1] = 0;
    for (int rd = 0; rd < JC - 1; ++rd, ++g) {
      const int cur = g & 1;
      if (tid < 64) {
        if (tid == 0) s_round[cur] = 0;
        if (tid < JC / 2) {
          int pp, qq;
          inner_pair(rd, tid, pp, qq);
          const zc gv = A[pp][qq];
          const double a = A[pp][pp].x, b = A[qq][qq].x;
          double c = 1.0;
          zc sw = zmk(0.0, 0.0);
          int rot = 0;
          if (needs_rotation(gv, a, b)) {
            double tg;
            jrot(a, b, gv, c, sw, tg);
            A[pp][pp] = zmk(a - tg, 0.0);
            A[qq][qq] = zmk(b + tg, 0.0);
            A[pp][qq] = zmk(0.0, 0.0);
            A[qq][pp] = zmk(0.0, 0.0);
            rot = 1;
            s_round[cur] = 1;
            s_sweep[isw & 1] = 1;
          }
          rb[cur].c[tid] = c;
          rb[cur].sw[tid] = sw;
          rb[cur].pq[tid] = pp | qq << 8 | rot << 16;
        }
      } else if (g > 0 && s_round[cur ^ 1]) {
        const int u = tid - 64;  // 448 threads for the 512 groups
        q_rotate(Q, rb[cur ^ 1], u);
        if (u < JC * JC / 2 - (kRoundThreads - 64)) q_rotate(Q, rb[cur ^ 1], u + (kRoundThreads - 64));
      }
      __syncthreads();
      if (!s_round[cur]) continue;
      if (ob < kOffBlocks) {
        const RotBuf& R = rb[cur];
        const int pqr = R.pq[tr], pqk = R.pq[tk];
        const double cr = R.c[tr], ck = R.c[tk];
        const zc swr = R.sw[tr], swk = R.sw[tk];
        if ((pqr | pqk) >> 16) {
          const int pr = pqr & 255, qr = (pqr >> 8) & 255, pk = pqk & 255, qk = (pqk >> 8) & 255;
          const int own = half ? qr : pr, oth = half ? pr : qr;
          const zc own0 = A[own][pk], own1 = A[own][qk], oth0 = A[oth][pk], oth1 = A[oth][qk];
          // this thread's row after the row rotation (a pair that stays put has c = 1,
          // s w = 0: the identity): A_p <- c A_p - s w A_q ; A_q <- s conj(w) A_p + c A_q
          const zc f = half ? zconj(swr) : zmk(-swr.x, -swr.y);
          const zc x0 = zadd(zscale(own0, cr), zmul(f, oth0));
          const zc x1 = zadd(zscale(own1, cr), zmul(f, oth1));
          // then the column rotation: A_:p <- c A_:p - s conj(w) A_:q ; A_:q <- s w A_:p + c A_:q
          const zc n0 = zsub(zscale(x0, ck), zmul(zconj(swk), x1));
          const zc n1 = zadd(zmul(swk, x0), zscale(x1, ck));
          A[own][pk] = n0;
          A[own][qk] = n1;
          A[pk][own] = zconj(n0);
          A[qk][own] = zconj(n1);
        }
      }
      __syncthreads();
    }
    if (!s_sweep[isw & 1]) break;
    if (isw + 1 < max_sweeps) {  // would the next sweep rotate anything?
      if (tid == 0) *s_more = 0;
      __syncthreads();
      for (int e = tid; e < JC * JC; e += kRoundThreads) {
        const int pp = e % JC, qq = e / JC;
        if (pp < qq && needs_rotation(A[pp][qq], A[pp][pp].x, A[qq][qq].x)) *s_more = 1;
      }
      __syncthreads();
      if (!*s_more) break;
    }
  }
  __syncthreads();
  if (g > 0 && s_round[(g - 1) & 1]) q_rotate(Q, rb[(g - 1) & 1], tid);
  __syncthreads();
}

template <bool FULL>
__global__ __launch_bounds__(kRoundThreads) void k_jacobi_round(const BondTask* __restrict__ tasks, int sweep, int round,
                                                      double tol_rel, double tol_abs, int* __restrict__ wl_count,
                                                      int2* __restrict__ wl_items, int wl_cap, int inner_sweeps) {
  const BondTask t = tasks[blockIdx.y];
  const int p = blockIdx.x;
  if (p >= t.P || round >= 2 * t.P - 1) return;
  int* cnt = t.counter;
  if (sweep > 0 && cnt[(sweep - 1) & 1] == 0) return;  // converged in the previous sweep
  const int nblk = 2 * t.P;
  int bi, bj;
  rr_pair(nblk, round, p, bi, bj);
  const int lo = bi < bj ? bi : bj, hi = bi < bj ? bj : bi;
  const unsigned vlo = t.blkver[lo], vhi = t.blkver[hi];
  const unsigned long long stamp = ((unsigned long long)vlo << 32) | vhi;
  unsigned long long* seen = t.pairseen + (size_t)lo * nblk + hi;
  if (!FULL && *seen == stamp) return;  // unchanged since it was verified orthogonal

  __shared__ zc A[JC][JC + 1];
  // The rotation accumulator Q; until the inner solve starts, its storage holds the
  // Gram's reduction scratch (red) and the column-norm partials (nred), which keeps
  // the workgroup at ~35 KB of LDS: registers, not LDS, bound the workgroups per CU.
  __shared__ zc Q[JC][JC + 1];
  double (*red)[256] = reinterpret_cast<double (*)[256]>(&Q[0][0]);
  double (*nred)[2][16] = reinterpret_cast<double (*)[2][16]>(reinterpret_cast<double*>(&Q[0][0]) + (kGramWaves - 1) * 256);
  static_assert(((kGramWaves - 1) * 256 + kGramWaves * 2 * 16) * sizeof(double) <= sizeof(zc) * JC * (JC + 1),
                "the Gram scratch must fit in Q's storage");
  __shared__ RotBuf rb[2];
  __shared__ int s_flag, s_round[2], s_sweep[2], s_more;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const zc* Wi = t.W + (size_t)t.R * (bi * JB);
  const zc* Wj = t.W + (size_t)t.R * (bj * JB);

  // ---- Gram: A = [Wi Wj]^H [Wi Wj]
  {
    d4 gre, gim;
    double nx, ny;
    auto gram = [&](const zc* X, const zc* Y) { gram_block(X, Y, t.R, wave, lane, gre, gim, nx, ny); };
    gram(Wi, Wj);
    reduce_block(A, red, gre, gim, wave, lane, 0, JB, false);
    if (FULL) {
      gram(Wi, Wi);
      reduce_block(A, red, gre, gim, wave, lane, 0, 0, true);
      gram(Wj, Wj);
      reduce_block(A, red, gre, gim, wave, lane, JB, JB, true);
    } else {
      nx += __shfl_xor(nx, 16, 64);
      nx += __shfl_xor(nx, 32, 64);
      ny += __shfl_xor(ny, 16, 64);
      ny += __shfl_xor(ny, 32, 64);
      if (lane < 16 && wave < kGramWaves) {
        nred[wave][0][lane] = nx;
        nred[wave][1][lane] = ny;
      }
      // within-block entries: zero (the blocks' columns are orthogonal)
      for (int e = tid; e < 2 * JB * JB; e += kRoundThreads) {
        const int blk = e / (JB * JB), r = (e / JB) % JB, c = e % JB;
        if (r != c) A[blk * JB + r][blk * JB + c] = zmk(0.0, 0.0);
      }
      __syncthreads();
      if (tid < JC) {
        const int side = tid / JB, col = tid % JB;
        double v = 0.0;
        for (int w = 0; w < kGramWaves; ++w) v += nred[w][side][col];
        A[tid][tid] = zmk(v, 0.0);
      }
      __syncthreads();
    }
  }

  // ---- does this pair need a rotation?  |A_pq| > tol_rel sqrt(A_pp A_qq) and
  // |A_pq| > tol_abs max_c|w_c| (|w_p| + |w_q|).  The second bound is the rounding
  // floor: rotations add errors of order eps * sigma_max to every column they touch.
  {
    const double fnorm = sqrt(t.result[2]);  // largest column norm
    int need = 0;
    for (int e = tid; e < JC * JC; e += kRoundThreads) {
      const int pp = e % JC, qq = e / JC;
      if (pp >= qq) continue;
      const double g2 = zabs2(A[pp][qq]);
      const double app = fmax(A[pp][pp].x, 0.0), aqq = fmax(A[qq][qq].x, 0.0);
      const double fl = tol_abs * fnorm * (sqrt(app) + sqrt(aqq));
      if (g2 > tol_rel * tol_rel * app * aqq && g2 > fl * fl) need = 1;
    }
    if (tid == 0) s_flag = 0;
    __syncthreads();
    if (need) s_flag = 1;
    __syncthreads();
    if (!s_flag) {
      if (tid == 0) {
        *seen = stamp;
        atomicAdd(&cnt[2], 1);
      }
      return;
    }
  }

  // ---- eigenvectors of A -> Q, to the pair's slot; queue the pair for rotation
  inner_jacobi(A, Q, rb, s_round, s_sweep, &s_more, tid, inner_sweeps);
  zc* qg = t.Qg + (size_t)p * JC * JC;
  for (int e = tid; e < JC * JC; e += kRoundThreads) qg[e] = Q[e / JC][e % JC];  // qg[k JC + c] = Q[k][c]
  if (tid == 0) {
    atomicAdd(&cnt[sweep & 1], 1);
    atomicAdd(&cnt[2], 1);
    atomicAdd(&cnt[3], 1);
    t.blkver[lo] = vlo + 1;
    t.blkver[hi] = vhi + 1;
    const int slot = atomicAdd(&wl_count[round & 1], 1);
    if (slot < wl_cap) wl_items[(size_t)(round & 1) * wl_cap + slot] = make_int2(blockIdx.y, p);
  }
}

constexpr int kRotRows = 64;  // rows of W or V per work item of k_jacobi_rotate

// Work item w = (queued pair w / nchunk, 64-row slab w % nchunk): slabs of W first,
// then of V.  The pair's Q comes from its slot; each wave rotates 16 rows as
// (X Q)^T = Q^T X^T: A operand Q^T (16 output columns x 4 k), B operand X^T (4 k x
// 16 rows), so every load and store touches 4 columns x 16 consecutive rows.
__global__ __launch_bounds__(256) void k_jacobi_rotate(const BondTask* __restrict__ tasks, int round,
                                                       int* __restrict__ wl_count, const int2* __restrict__ wl_items,
                                                       int wl_cap, int nchunk_w, int nchunk) {
  if (blockIdx.x == 0 && threadIdx.x == 0) wl_count[(round + 1) & 1] = 0;  // the next round's queue
  const int n = min(wl_count[round & 1], wl_cap);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  for (int w = blockIdx.x; w < n * nchunk; w += gridDim.x) {
    const int2 item = wl_items[(size_t)(round & 1) * wl_cap + w / nchunk];
    const int chunk = w % nchunk;
    const BondTask& t = tasks[item.x];
    const int p = item.y;
    int bi, bj;
    rr_pair(2 * t.P, round, p, bi, bj);
    zc* M;
    int rows, ld, r0;
    if (chunk < nchunk_w) {
      M = t.W; rows = t.R; ld = t.R; r0 = chunk * kRotRows;
    } else {
      M = t.V; rows = t.V ? t.NV : 0; ld = t.NV; r0 = (chunk - nchunk_w) * kRotRows;
    }
    if (r0 >= rows) continue;
    const zc* qg = t.Qg + (size_t)p * JC * JC;
    double qre[2][8], qim[2][8];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const zc q = qg[(4 * ks + lk) * JC + 16 * ct + lr];
        qre[ct][ks] = q.x;
        qim[ct][ks] = q.y;
      }
    auto col_of = [&](int c) { return c < JB ? bi * JB + c : bj * JB + (c - JB); };
    const int r = r0 + 16 * wave + lr;
    const bool in = r < rows;
    zc x[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) x[ks] = in ? M[(size_t)ld * col_of(4 * ks + lk) + r] : zmk(0.0, 0.0);
    d4 ore[2], oim[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      ore[ct] = d4{0.0, 0.0, 0.0, 0.0};
      oim[ct] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        ore[ct] = mfma4(qre[ct][ks], x[ks].x, ore[ct]);
        ore[ct] = mfma4(-qim[ct][ks], x[ks].y, ore[ct]);
        oim[ct] = mfma4(qre[ct][ks], x[ks].y, oim[ct]);
        oim[ct] = mfma4(qim[ct][ks], x[ks].x, oim[ct]);
      }
    }
    if (in) {
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int g = 0; g < 4; ++g) M[(size_t)ld * col_of(16 * ct + lk + 4 * g) + r] = zmk(ore[ct][g], oim[ct][g]);
    }
  }
}

// Start of a Jacobi solve: zero the counters and block versions, invalidate the
// verified-pair stamps.  One workgroup per task.
__global__ void k_jacobi_begin(const BondTask* __restrict__ tasks) {
  const BondTask t = tasks[blockIdx.x];
  const int nblk = 2 * t.P;
  if (threadIdx.x < 4) t.counter[threadIdx.x] = 0;
  for (int e = threadIdx.x; e < nblk; e += blockDim.x) t.blkver[e] = 0u;
  for (int e = threadIdx.x; e < nblk * nblk; e += blockDim.x) t.pairseen[e] = ~0ull;
}

// Start of sweep `sweep`: clear that sweep's rotation counter (its parity slot last
// held sweep - 2, which is no longer read) and both rotation queues.
__global__ void k_jacobi_sweep_begin(const BondTask* __restrict__ tasks, int ntasks, int sweep, int* wl_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntasks) tasks[i].counter[sweep & 1] = 0;
  if (i < 2) wl_count[i] = 0;
}

}  // namespace

hipError_t launch_jacobi_begin(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_jacobi_begin, dim3(ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

hipError_t launch_jacobi_sweep_begin(const BondTask* d_tasks, int ntasks, int sweep, int* wl_count, hipStream_t st) {
  hipLaunchKernelGGL(k_jacobi_sweep_begin, dim3((ntasks + 255) / 256), dim3(256), 0, st, d_tasks, ntasks, sweep,
                     wl_count);
  return hipGetLastError();
}

hipError_t launch_jacobi_round(const BondTask* d_tasks, int ntasks, int max_pairs, int sweep, int round,
                               double tol_rel, double tol_abs, int max_rows_w, int max_rows_v, int* wl_count,
                               int2* wl_items, int wl_cap, int inner_sweeps, hipStream_t st) {
  if (inner_sweeps < 1 || inner_sweeps > kInnerSweeps) inner_sweeps = kInnerSweeps;
  if (round == 0)
    hipLaunchKernelGGL(k_jacobi_round<true>, dim3(max_pairs, ntasks), dim3(kRoundThreads), 0, st, d_tasks, sweep, round,
                       tol_rel, tol_abs, wl_count, wl_items, wl_cap, inner_sweeps);
  else
    hipLaunchKernelGGL(k_jacobi_round<false>, dim3(max_pairs, ntasks), dim3(kRoundThreads), 0, st, d_tasks, sweep, round,
                       tol_rel, tol_abs, wl_count, wl_items, wl_cap, inner_sweeps);
  return hipGetLastError();
}

hipError_t launch_jacobi_rotate(const BondTask* d_tasks, int round, int max_rows_w, int max_rows_v, int* wl_count,
                                int2* wl_items, int wl_cap, hipStream_t st) {
  const int nw = (max_rows_w + kRotRows - 1) / kRotRows, nv = (max_rows_v + kRotRows - 1) / kRotRows;
  hipLaunchKernelGGL(k_jacobi_rotate, dim3(1024), dim3(256), 0, st, d_tasks, round, wl_count, wl_items, wl_cap, nw,
                     nw + nv);
  return hipGetLastError();
}
