// One-sided block-Jacobi SVD of the bond matrix (replaces ZGESVD in SVDTruncation,
// local_operations_module.f90:322-345).
//
// W (R x Cpad, column-major) is orthogonalised in place and V accumulates the right
// rotations.  The columns form 2P blocks of JB; a round pairs every block with one
// other (round-robin "circle" order, 2P-1 rounds per sweep) and one workgroup per
// pair
//   1. forms the pair's Gram matrix with FP64 MFMA (v_mfma_f64_16x16x4f64),
//   2. decides whether the pair needs a rotation,
//   3. diagonalises the 32x32 Gram by cyclic two-sided Jacobi in LDS -> Q,
//   4. rotates the 32 columns of W and V by Q, again on MFMA, with Q held in
//      registers and W / V streamed through with coalesced 256-byte segments.
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

// Round-robin ("circle") schedule over nb = 2P blocks: block 0 stays, the others
// rotate; over nb-1 rounds every pair of blocks meets exactly once.
__device__ __forceinline__ void rr_pair(int nb, int round, int p, int& bi, int& bj) {
  auto pos = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + round) % (nb - 1); };
  bi = pos(p);
  bj = pos(nb - 1 - p);
}

// Hermitian 2x2 Jacobi rotation J = [[c, s w], [-s conj(w), c]], w = g/|g|, that
// zeroes g in [[a, g], [conj(g), b]] via J^dagger H J.
__device__ __forceinline__ void jrot(double a, double b, zc g, double& c, double& s, zc& w) {
  const double ag = sqrt(zabs2(g));
  if (ag == 0.0) { c = 1.0; s = 0.0; w = zmk(1.0, 0.0); return; }
  w = zmk(g.x / ag, g.y / ag);
  const double zeta = (b - a) / (2.0 * ag);
  const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  c = 1.0 / sqrt(1.0 + tt * tt);
  s = tt * c;
}

// Gram block D = X^H Y (16 x 16) of two 16-column blocks over rows [0, R), as
// per-wave partial sums: wave w takes rows 8*(w + 4 it) .. +7 in two MFMA k-steps
// of 4 rows.  Also accumulates the lane's partial squared norms of X and Y.
__device__ __forceinline__ void gram_block(const zc* __restrict__ X, const zc* __restrict__ Y, int R, int wave,
                                           int lane, d4& gre, d4& gim, double& nx, double& ny) {
  const int lr = lane & 15, lk = lane >> 4;
  gre = d4{0.0, 0.0, 0.0, 0.0};
  gim = d4{0.0, 0.0, 0.0, 0.0};
  nx = 0.0;
  ny = 0.0;
  const zc* xc = X + (size_t)R * lr;
  const zc* yc = Y + (size_t)R * lr;
  for (int r0 = 8 * wave; r0 < R; r0 += 32) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 4 * h + lk;
      zc a = zmk(0.0, 0.0), b = zmk(0.0, 0.0);
      if (r < R) { a = xc[r]; b = yc[r]; }
      // A[p][k] = conj(X[k][p]), B[k][q] = Y[k][q]
      gre = mfma4(a.x, b.x, gre);
      gre = mfma4(a.y, b.y, gre);
      gim = mfma4(a.x, b.y, gim);
      gim = mfma4(-a.y, b.x, gim);
      nx += zabs2(a);
      ny += zabs2(b);
    }
  }
}

constexpr int kInnerSweeps = 30;

// Sums the four waves' partial 16x16 blocks in a fixed order (bitwise reproducible)
// and stores rows [ro, ro+16) x cols [co, co+16) of A.  With `diag` the block is a
// diagonal block of the Hermitian A: only the upper triangle is taken and mirrored.
__device__ __forceinline__ void reduce_block(zc (*A)[JC + 1], double (*red)[2][256], d4 gre, d4 gim, int wave,
                                             int lane, int ro, int co, bool diag) {
  if (wave > 0) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      red[wave - 1][0][g * 64 + lane] = gre[g];
      red[wave - 1][1][g * 64 + lane] = gim[g];
    }
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      double xr = gre[g], xi = gim[g];
      for (int w = 0; w < 3; ++w) {
        xr += red[w][0][g * 64 + lane];
        xi += red[w][1][g * 64 + lane];
      }
      const int p = (lane >> 4) + 4 * g, q = lane & 15;
      if (!diag) {
        A[ro + p][co + q] = zmk(xr, xi);
        A[co + q][ro + p] = zmk(xr, -xi);
      } else if (p < q) {
        A[ro + p][co + q] = zmk(xr, xi);
        A[co + q][ro + p] = zmk(xr, -xi);
      } else if (p == q) {
        A[ro + p][co + p] = zmk(xr, 0.0);
      }
    }
  }
  __syncthreads();
}

template <bool FULL>
__global__ __launch_bounds__(256) void k_jacobi_round(const BondTask* __restrict__ tasks, int sweep, int round,
                                                      double tol_rel, double tol_abs) {
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
  __shared__ zc Q[JC][JC + 1];
  __shared__ double red[3][2][256];
  __shared__ double nred[4][2][16];
  __shared__ double rc_c[JC / 2], rc_s[JC / 2];
  __shared__ zc rc_w[JC / 2];
  __shared__ int rc_p[JC / 2], rc_q[JC / 2];
  __shared__ int s_flag, s_rot;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const zc* Wi = t.W + (size_t)t.R * (bi * JB);
  const zc* Wj = t.W + (size_t)t.R * (bj * JB);

  // ---- Gram: A = [Wi Wj]^H [Wi Wj]
  {
    d4 gre, gim;
    double nx, ny;
    gram_block(Wi, Wj, t.R, wave, lane, gre, gim, nx, ny);
    reduce_block(A, red, gre, gim, wave, lane, 0, JB, false);
    if (FULL) {
      gram_block(Wi, Wi, t.R, wave, lane, gre, gim, nx, ny);
      reduce_block(A, red, gre, gim, wave, lane, 0, 0, true);
      gram_block(Wj, Wj, t.R, wave, lane, gre, gim, nx, ny);
      reduce_block(A, red, gre, gim, wave, lane, JB, JB, true);
    } else {
      nx += __shfl_xor(nx, 16, 64);
      nx += __shfl_xor(nx, 32, 64);
      ny += __shfl_xor(ny, 16, 64);
      ny += __shfl_xor(ny, 32, 64);
      if (lane < 16) {
        nred[wave][0][lane] = nx;
        nred[wave][1][lane] = ny;
      }
      // within-block entries: zero (the blocks' columns are orthogonal)
      for (int e = tid; e < 2 * JB * JB; e += 256) {
        const int blk = e / (JB * JB), r = (e / JB) % JB, c = e % JB;
        if (r != c) A[blk * JB + r][blk * JB + c] = zmk(0.0, 0.0);
      }
      __syncthreads();
      if (tid < JC) {
        const int side = tid / JB, col = tid % JB;
        double v = 0.0;
        for (int w = 0; w < 4; ++w) v += nred[w][side][col];
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
    for (int e = tid; e < JC * JC; e += 256) {
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
  if (tid == 0) {
    atomicAdd(&cnt[sweep & 1], 1);
    atomicAdd(&cnt[2], 1);
    atomicAdd(&cnt[3], 1);
  }

  // ---- cyclic Jacobi on A (JC x JC Hermitian) in LDS, Q accumulates the rotations.
  // A Gram entry is accurate relative to sqrt(A_pp A_qq), so the inner solve uses a
  // purely relative test (skip when |A_pq| <= 1e-15 sqrt(A_pp A_qq)), sets each
  // annihilated entry to exactly zero and ends after a sweep without rotations.
  // The pair's columns keep their block (re-sorting columns between blocks stalls
  // the convergence of the cyclic block sweep).
  for (int e = tid; e < JC * JC; e += 256) {
    const int r = e % JC, c = e / JC;
    Q[r][c] = zmk(r == c ? 1.0 : 0.0, 0.0);
  }
  for (int isw = 0; isw < kInnerSweeps; ++isw) {
    if (tid == 0) s_rot = 0;
    __syncthreads();
    for (int rd = 0; rd < JC - 1; ++rd) {
      if (tid < JC / 2) {
        int pp, qq;
        rr_pair(JC, rd, tid, pp, qq);
        if (pp > qq) { const int tmp = pp; pp = qq; qq = tmp; }
        double c, s;
        zc w;
        const zc g = A[pp][qq];
        const double a = A[pp][pp].x, b = A[qq][qq].x;
        const double g2 = zabs2(g);
        if (g2 <= 1e-30 * fabs(a * b) || g2 == 0.0) { c = 1.0; s = 0.0; w = zmk(1.0, 0.0); }
        else { jrot(a, b, g, c, s, w); s_rot = 1; }
        rc_c[tid] = c; rc_s[tid] = s; rc_w[tid] = w; rc_p[tid] = pp; rc_q[tid] = qq;
      }
      __syncthreads();
      // rows: A_p <- c A_p - s w A_q ; A_q <- s conj(w) A_p + c A_q
      for (int e = tid; e < (JC / 2) * JC; e += 256) {
        const int pr = e / JC, k = e % JC;
        const double c = rc_c[pr], s = rc_s[pr];
        if (s == 0.0) continue;
        const zc w = rc_w[pr];
        const int pp = rc_p[pr], qq = rc_q[pr];
        const zc xp = A[pp][k], xq = A[qq][k];
        const zc sw = zscale(w, s), swc = zscale(zconj(w), s);
        A[pp][k] = zsub(zscale(xp, c), zmul(sw, xq));
        A[qq][k] = zadd(zmul(swc, xp), zscale(xq, c));
      }
      __syncthreads();
      // cols: A_:p <- c A_:p - s conj(w) A_:q ; A_:q <- s w A_:p + c A_:q  (same for Q)
      for (int e = tid; e < (JC / 2) * JC; e += 256) {
        const int pr = e / JC, k = e % JC;
        const double c = rc_c[pr], s = rc_s[pr];
        if (s == 0.0) continue;
        const zc w = rc_w[pr];
        const int pp = rc_p[pr], qq = rc_q[pr];
        const zc sw = zscale(w, s), swc = zscale(zconj(w), s);
        zc xp = A[k][pp], xq = A[k][qq];
        A[k][pp] = zsub(zscale(xp, c), zmul(swc, xq));
        A[k][qq] = zadd(zmul(sw, xp), zscale(xq, c));
        xp = Q[k][pp]; xq = Q[k][qq];
        Q[k][pp] = zsub(zscale(xp, c), zmul(swc, xq));
        Q[k][qq] = zadd(zmul(sw, xp), zscale(xq, c));
      }
      __syncthreads();
      if (tid < JC / 2 && rc_s[tid] != 0.0) {
        A[rc_p[tid]][rc_q[tid]] = zmk(0.0, 0.0);
        A[rc_q[tid]][rc_p[tid]] = zmk(0.0, 0.0);
      }
      __syncthreads();
    }
    if (!s_rot) break;
    __syncthreads();
  }

  // ---- [W; V](:, pair) <- [W; V](:, pair) Q on MFMA, computed as (X Q)^T = Q^T X^T:
  // A operand Q^T (16 output columns x 4 k), B operand X^T (4 k x 16 rows), so each
  // load and store instruction touches 4 columns x 16 consecutive rows.
  const int lr = lane & 15, lk = lane >> 4;
  double qre[2][8], qim[2][8];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const zc q = Q[4 * ks + lk][16 * ct + lr];
      qre[ct][ks] = q.x;
      qim[ct][ks] = q.y;
    }
  auto col_of = [&](int c) { return c < JB ? bi * JB + c : bj * JB + (c - JB); };
  auto rotate = [&](zc* M, int rows, int ld) {
    const zc* src[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) src[ks] = M + (size_t)ld * col_of(4 * ks + lk);
    for (int r0 = 16 * wave; r0 < rows; r0 += 64) {
      const int r = r0 + lr;
      const bool in = r < rows;
      zc x[8];
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) x[ks] = in ? src[ks][r] : zmk(0.0, 0.0);
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
  };
  rotate(t.W, t.R, t.R);
  if (t.V) rotate(t.V, t.NV, t.NV);
  if (tid == 0) {
    t.blkver[lo] = vlo + 1;
    t.blkver[hi] = vhi + 1;
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

// Start of sweep `sweep`: clear that sweep's rotation counter (its parity slot
// last held sweep - 2, which is no longer read).
__global__ void k_jacobi_sweep_begin(const BondTask* __restrict__ tasks, int ntasks, int sweep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntasks) tasks[i].counter[sweep & 1] = 0;
}

}  // namespace

hipError_t launch_jacobi_begin(const BondTask* d_tasks, int ntasks, hipStream_t st) {
  hipLaunchKernelGGL(k_jacobi_begin, dim3(ntasks), dim3(256), 0, st, d_tasks);
  return hipGetLastError();
}

hipError_t launch_jacobi_sweep_begin(const BondTask* d_tasks, int ntasks, int sweep, hipStream_t st) {
  hipLaunchKernelGGL(k_jacobi_sweep_begin, dim3((ntasks + 255) / 256), dim3(256), 0, st, d_tasks, ntasks, sweep);
  return hipGetLastError();
}

hipError_t launch_jacobi_round(const BondTask* d_tasks, int ntasks, int max_pairs, int sweep, int round,
                               double tol_rel, double tol_abs, hipStream_t st) {
  if (round == 0)
    hipLaunchKernelGGL(k_jacobi_round<true>, dim3(max_pairs, ntasks), dim3(256), 0, st, d_tasks, sweep, round,
                       tol_rel, tol_abs);
  else
    hipLaunchKernelGGL(k_jacobi_round<false>, dim3(max_pairs, ntasks), dim3(256), 0, st, d_tasks, sweep, round,
                       tol_rel, tol_abs);
  return hipGetLastError();
}
