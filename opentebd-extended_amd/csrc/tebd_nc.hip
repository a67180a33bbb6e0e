// Kernels of the number-conserving (U(1) block-sparse) two-site update, which replaces
// TwoSiteOpNC_c (local_operations_module.f90:1845-1964).  With the numbers of every
// Schmidt state known, theta (the dense product lambda Gamma lambda Gamma lambda and the
// gate, from the ordinary theta kernel) is block diagonal in the number on the left; the
// host lists each block's rows (a, i) and columns (g, j) as the reference's IndexLeft /
// IndexRight do (:1339-1443), and these kernels
//   k_nc_gather     copy each block out of theta into its own SVD work matrix (FormBlockTheta,
//                   :1445-1468; transposed when it has more columns than rows) and sum its
//                   squared entries (blocks without weight skip the SVD);
//   k_nc_vectors    copy each block's left and right singular vectors out of the SVD's
//                   work arrays into the bond's vector store, so that the next blocks can
//                   reuse the work slots;
//   k_nc_gammas     write Gamma_l(a, i, b) = U(f, p) / lambda_l(a) and
//                   Gamma_r(b, j, g) = conj(V(h, p)) / lambda_r(g) for every kept state b, its
//                   block and position p (FormGamma1NC / FormGamma2NC, :1659-1721).
#include "tebd_device.h"
#include "tebd_kernels.h"

namespace {

constexpr int kGatherSplit = 16;  // workgroups per block in k_nc_gather

__global__ __launch_bounds__(256) void k_nc_gather(const NcBlock* __restrict__ blocks, double* __restrict__ partial) {
  const NcBlock b = blocks[blockIdx.y];
  const size_t n = (size_t)b.lgt * b.trv;
  double acc = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int f = (int)(e % b.lgt), h = (int)(e / b.lgt);
    const zc v = b.theta[b.rows[f] + (size_t)b.m * b.cols[h]];
    acc += v.x * v.x + v.y * v.y;
    if (!b.W) continue;
    if (!b.transposed) b.W[f + (size_t)b.R * h] = v;
    else b.W[h + (size_t)b.R * f] = zconj(v);
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Column c of the left (U, lgt rows) or right (V, trv rows) singular vectors of block B,
// read as k_write_gammas reads a bond's: sorted and normalised in Yr / Zc after the QR
// preconditioning, else the Jacobi's W / s and V at the sorted value's column perm[c].
// The work matrix is B (R = lgt, C = trv) or, transposed, B^dagger (R = trv, C = lgt).
__device__ __forceinline__ zc nc_left(const BondTask& t, int tr, int row, int c) {
  if (t.qr) return !tr ? t.Yr[row + (size_t)t.Rb * c] : t.Zc[row + (size_t)t.Cb * c];
  const int col = t.perm[c];
  if (!tr) return zscale(t.W[row + (size_t)t.R * col], t.s[c] > 0.0 ? 1.0 / t.s[c] : 0.0);
  return t.V[row + (size_t)t.NV * col];
}
__device__ __forceinline__ zc nc_right(const BondTask& t, int tr, int row, int c) {
  if (t.qr) return !tr ? t.Zc[row + (size_t)t.Cb * c] : t.Yr[row + (size_t)t.Rb * c];
  const int col = t.perm[c];
  if (!tr) return t.V[row + (size_t)t.NV * col];
  return zscale(t.W[row + (size_t)t.R * col], t.s[c] > 0.0 ? 1.0 / t.s[c] : 0.0);
}

// One workgroup per (block, vector column): U(:, c) (lgt) and V(:, c) (trv), c < the
// block's computed count.
__global__ __launch_bounds__(256) void k_nc_vectors(const NcBlock* __restrict__ blocks,
                                                    const BondTask* __restrict__ tasks) {
  const NcBlock b = blocks[blockIdx.y];
  const BondTask t = tasks[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= (int)t.result[0]) return;
  for (int f = threadIdx.x; f < b.lgt; f += blockDim.x) b.U[f + (size_t)b.lgt * c] = nc_left(t, b.transposed, f, c);
  for (int h = threadIdx.x; h < b.trv; h += blockDim.x) b.V[h + (size_t)b.trv * c] = nc_right(t, b.transposed, h, c);
}

// One workgroup per kept state b of a bond: its column of Gamma_l and its row of Gamma_r,
// zero outside its block.  The reference divides only where lambda > 10.0**(-10) (a single
// precision literal).
__global__ __launch_bounds__(256) void k_nc_gammas(const NcKept* __restrict__ kept, double cut) {
  const NcKept q = kept[blockIdx.x];
  const size_t nl = (size_t)q.chil * q.d, nr = (size_t)q.d * q.chir;
  for (size_t e = threadIdx.x; e < nl; e += blockDim.x) q.gl[e + nl * q.b] = zmk(0.0, 0.0);
  for (size_t e = threadIdx.x; e < nr; e += blockDim.x) q.gr[q.b + (size_t)q.knew * e] = zmk(0.0, 0.0);
  __syncthreads();
  for (int f = threadIdx.x; f < q.lgt; f += blockDim.x) {
    const int r = q.rows[f];                 // theta row a + chi_l i
    const int a = r % q.chil;
    const double la = q.ll[a];
    const zc u = q.U[f + (size_t)q.lgt * q.pos];
    q.gl[r + nl * q.b] = la > cut ? zscale(u, 1.0 / la) : zmk(0.0, 0.0);
  }
  for (int h = threadIdx.x; h < q.trv; h += blockDim.x) {
    const int col = q.cols[h];               // theta column g + chi_r j
    const int g = col % q.chir, j = col / q.chir;
    const double lg = q.lr[g];
    const zc v = q.V[h + (size_t)q.trv * q.pos];
    q.gr[q.b + (size_t)q.knew * (j + (size_t)q.d * g)] = lg > cut ? zscale(zconj(v), 1.0 / lg) : zmk(0.0, 0.0);
  }
}

}  // namespace

hipError_t launch_nc_gather(const NcBlock* d_blocks, int nblocks, double* d_partial, hipStream_t st) {
  hipLaunchKernelGGL(k_nc_gather, dim3(kGatherSplit, nblocks), dim3(256), 0, st, d_blocks, d_partial);
  return hipGetLastError();
}

int nc_gather_split() { return kGatherSplit; }

hipError_t launch_nc_vectors(const NcBlock* d_blocks, const BondTask* d_tasks, int nblocks, int max_k,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_nc_vectors, dim3(max_k, nblocks), dim3(256), 0, st, d_blocks, d_tasks);
  return hipGetLastError();
}

hipError_t launch_nc_gammas(const NcKept* d_kept, int nkept, double cut, hipStream_t st) {
  hipLaunchKernelGGL(k_nc_gammas, dim3(nkept), dim3(256), 0, st, d_kept, cut);
  return hipGetLastError();
}
