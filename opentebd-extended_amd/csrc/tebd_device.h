// Device-side types shared by the gfx950 kernels of the two-site update.
//
// Layout in HBM (one context = one chain):
//   Gamma of site s : packed complex(8) tensor Gamma(a,i,b), a < chi_s, i < d,
//                     b < chi_{s+1}, element at a + chi_s*(i + d*b) -- the
//                     reference's column-major order (TEBDtools_module.f90:496-513)
//                     with the *current* bond dimensions as leading dimensions,
//                     inside a slab of capacity chi_max*d*chi_max.
//   lambda of bond l: chi_max doubles, entries >= chi_l are zero.
//   theta/W         : R x Cpad complex column-major work matrix of one bond
//                     (R = max(m,n), C = min(m,n), Cpad = C rounded up to a
//                     multiple of the Jacobi pair width).
//   V               : Cpad x Cpad complex, accumulated right rotations.
#pragma once
#include <hip/hip_runtime.h>

struct __align__(16) zc {
  double x, y;
};

// Batched complex GEMM: C (M x N) = op(A) (M x K) * B (K x N); op(A) = A (stored
// M x K, lda) or A^dagger (A stored K x M, lda).  All column-major.
struct GemmTask {
  const zc* A;
  const zc* B;
  zc* C;
  int M, N, K, lda, ldb, ldc;
  double alpha;    // C = alpha op(A) B (+ C when accumulate)
  int accumulate;
};

constexpr int RESULT_STRIDE = 8;

// One block of a number-conserving bond update: theta is block diagonal in the number
// on the left, and block k gathers the rows (a, i) and columns (g, j) of number k.
struct NcBlock {
  const zc* theta;   // the bond's theta, m x n (ld m)
  int m;
  const int* rows;   // lgt theta rows a + chi_l i
  const int* cols;   // trv theta columns g + chi_r j
  int lgt, trv;
  int transposed;    // the work matrix is B^dagger (lgt < trv)
  zc* W;             // SVD work matrix (R x C, ld R); null: only the squared sum is formed
  int R;
  zc* U;             // lgt x min(lgt, trv): left singular vectors (the bond's vector store)
  zc* V;             // trv x min(lgt, trv): right singular vectors
};

// Workspace of the split-operand theta (one slot per task of a batch): the int8 slices of
// A = Gamma_l lambda_c (rows (a, k) -> a d + k) and of B = Gamma_r (columns (l, b) ->
// b d + l), plane (2 s + part) of `rows` rows x `pitch` bytes each (pitch: the middle
// bond padded to 64), each row's and column's power-of-two scale, and the power-of-two
// balance of each middle index g (A's column g times 2^eg, B's row g times 2^-eg).
constexpr int kSplitSlices = 7;
struct SplitWs {
  signed char* a;
  signed char* b;
  int* ea;          // rows per task
  int* eb;
  int* eg;          // pitch per task
  size_t rows;      // row capacity: d round_up(chi_max, 16)
  size_t stride;    // bytes per task: 2 kSplitSlices rows pitch
  int pitch;        // round_up(chi_max, 64)
};

// One kept Schmidt state b of a number-conserving bond update: its block's vectors and
// position, and where its column of Gamma_l and row of Gamma_r go.
struct NcKept {
  const int* rows;
  const int* cols;
  int lgt, trv;
  const zc* U;
  const zc* V;
  int pos;
  zc* gl;            // Gamma_l (chi_l x d x knew)
  zc* gr;            // Gamma_r (knew x d x chi_r)
  const double* ll;
  const double* lr;
  int chil, chir, d, knew, b;
};

// A-posteriori check of one sketched bond (learned-rank path), by two sets of p
// complex Gaussian probes.  theta (m x n), E = theta - Q Q^dagger theta the part of theta
// outside the sketch's range basis Q (m x r), kept triplets (s_j, u_j, v_j), j < k:
//   set 1: Z1 = theta V G1 - U S G1 = E V G1        (G1: k x p), since theta v_j = s_j u_j + E v_j;
//   set 2: Z2 = (I - Q Q^dagger) theta G2 = E G2   (G2: n x p);
// ||Z1||_F^2 / p estimates sum_j ||E v_j||^2 (how far the kept vectors are from singular
// vectors of theta), ||Z2||_F^2 / p estimates ||E||_F^2 (the weight the sketch missed).
constexpr int kProbes = 16;
struct ProbeTask {
  const double* s;   // k kept singular values of theta, descending
  int k, m, n;
  int rng_key;       // the bond's position in its chain
  zc* C1;            // k x p: G1
  zc* C2;            // k x p: S G1
  zc* Om;            // n x 2p: [V G1 | G2] (the coefficient kernel writes G2)
  zc* Z;             // m x 2p: [Z1 | Z2]
  double* out;       // [0] ||Z1||_F^2, [1] ||Z2||_F^2
};

__device__ __forceinline__ zc zmk(double x, double y) { zc r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ zc zadd(zc a, zc b) { return zmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ zc zsub(zc a, zc b) { return zmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ zc zmul(zc a, zc b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ zc zscale(zc a, double s) { return zmk(a.x * s, a.y * s); }
__device__ __forceinline__ zc zconj(zc a) { return zmk(a.x, -a.y); }
// acc += a * b
__device__ __forceinline__ void zfma(zc& acc, zc a, zc b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
__device__ __forceinline__ void zfmac(zc& acc, zc a, zc b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
__device__ __forceinline__ double zabs2(zc a) { return a.x * a.x + a.y * a.y; }

// Jacobi block geometry: pairs of JB-column blocks, i.e. JC = 2*JB columns
// are orthogonalised together by one workgroup.
constexpr int JB = 16;
constexpr int JC = 2 * JB;

// One bond update ("task") as the kernels see it.
struct BondTask {
  // inputs of the theta contraction
  const zc* gl;      // Gamma_l  (chi_l x d x chi_c)
  const zc* gr;      // Gamma_r  (chi_c x d x chi_r)
  const double* ll;  // lambda_l (chi_l)
  const double* lc;  // lambda_c (chi_c)
  const double* lr;  // lambda_r (chi_r)
  const zc* gate;    // d^2 x d^2 column-major, row (i*d+j)
  int chil, chic, chir, d;
  int link;          // 0-based bond (sites link, link+1)
  // SVD work
  int transposed;    // 0: W = theta (m x n); 1: W = theta^dagger (n x m)
  int R, C, Cpad, P; // rows, true columns, padded columns, Jacobi pairs (Cpad / JC)
  zc* W;             // R x Cpad, ld = R
  zc* V;             // NV x Cpad, ld = NV: accumulated right rotations (identity start) or a
                     // right basis to rotate along with W; nullptr = not needed
  int NV;
  double* s;         // Cpad: sorted singular values (descending)
  int* perm;         // Cpad: original column of each sorted value
  int* counter;      // Jacobi counters: [0..1] rotated pairs per sweep parity, [2] Gram visits, [3] rotations
  unsigned* blkver;  // 2P: rotations of each column block in the current solve
  unsigned long long* pairseen;  // (2P)^2: block versions when a pair was last verified orthogonal
  zc* Qg;            // P x 32 x 32: each pair's rotation of the current round, row-major
  double* result;    // [0]=kept k, [1]=truncerr, [2]=max column |w|^2 (Jacobi floor), [3]=||theta||^2 when
                     // W is a sketch (else <= 0), [4]=weight captured by W, [5]=smallest kept s^2
  // outputs
  zc* gl_out;        // new Gamma_l (chi_l x d x k)
  zc* gr_out;        // new Gamma_r (k x d x chi_r)
  double* lc_out;    // new lambda_c (chi_max entries, zero-filled past k)
  double* lc_commit; // when set, k_write_gammas copies lc_out here (sketched path)
  int chi_max;
  // column-pivoted QR preconditioning (qr = 1): Wq (Rb x Cb) holds the bond matrix and then its
  // Householder QR; the Jacobi (W, R, C above) runs on X = R(0:r, :)^dagger (Cb x r)
  int qr;
  int Rb, Cb;
  zc* Wq;
  double* qnrm;      // Cb remaining column norms^2
  int* qpiv;         // Cb column permutation
  zc* tau;           // Cb reflector scalars
  double* qstate;    // [0] numerical rank (-1 while running), [1] stop threshold on |col|^2
  zc* Tq;            // QB x QB compact-WY factors T of the panels (panel p at Tq + p QB^2, QB = 16)
  zc* Yr;            // Rb x chi_max: Rb-side singular vectors of the kept values, sorted
  zc* Zc;            // Cb x chi_max: Cb-side singular vectors of the kept values, sorted
  zc* Qs;            // Rb x r: explicit orthonormal factor (sketched SVD's range basis)
  const double* wtot;  // when set, the reference total weight (||theta||^2) for the truncation
  int rng_key;       // sketch stream of the bond: its position in its chain (not its batch slot)
};
