// The rank model of the learned-rank path: a 66 -> 128 -> 128 -> 1 perceptron in FP32
// (ReLU hidden layers) with hand-written inference, forward/backward and Adam kernels.
//
// It predicts the rank a bond update will keep, as log2(k_new / chi_c), from features
// of the bond before the update (k_mlp_features); the learned-rank path sizes the
// bond's sketch from the prediction.  Loss: mean of w (out - y)^2 with w = w_under when
// the model is below the target (an under-predicted rank costs a rejected sketch, an
// over-predicted one only wider GEMMs).
//
// Parameter layout (floats): W1[128][66], b1[128], W2[128][128], b2[128], W3[128], b3.
#include "tebd_device.h"
#include "tebd_kernels.h"
#include "../../include/tebd_hip.h"

#include <cstring>
#include <vector>

namespace {

constexpr int NI = TEBD_MLP_INPUTS, NH = TEBD_MLP_HIDDEN;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH;
constexpr int NP = OB3 + 1;
static_assert(NP == TEBD_MLP_PARAMS, "parameter count");

// Forward pass of one sample by 128 threads (thread j = hidden unit j); x in LDS.
// Leaves h1, h2 in LDS and returns the output in every thread.
__device__ __forceinline__ float mlp_forward(const float* __restrict__ p, const float* xs, float* h1s, float* h2s,
                                             float* red) {
  const int j = threadIdx.x;
  float a = p[OB1 + j];
  for (int i = 0; i < NI; ++i) a = fmaf(p[OW1 + j * NI + i], xs[i], a);
  h1s[j] = fmaxf(a, 0.0f);
  __syncthreads();
  a = p[OB2 + j];
  for (int k = 0; k < NH; ++k) a = fmaf(p[OW2 + j * NH + k], h1s[k], a);
  const float h2 = fmaxf(a, 0.0f);
  h2s[j] = h2;
  float o = p[OW3 + j] * h2;
  for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off, 64);
  if ((j & 63) == 0) red[j >> 6] = o;
  __syncthreads();
  return red[0] + red[1] + p[OB3];
}

// Forward + backward over the samples s = blockIdx.x, blockIdx.x + gridDim.x, ...;
// this workgroup's gradient sum goes to part[blockIdx.x][NP], its loss sum to
// lpart[blockIdx.x] (both later reduced in a fixed order).
__global__ __launch_bounds__(NH) void k_mlp_fb(const float* __restrict__ p, const float* __restrict__ X,
                                               const float* __restrict__ y, int n, float w_under,
                                               float* __restrict__ part, float* __restrict__ lpart) {
  __shared__ float xs[NI], h1s[NH], h2s[NH], dh2s[NH], red[2];
  const int j = threadIdx.x;
  float gW1[NI], gW2[NH];
#pragma unroll
  for (int i = 0; i < NI; ++i) gW1[i] = 0.0f;
#pragma unroll
  for (int k = 0; k < NH; ++k) gW2[k] = 0.0f;
  float gb1 = 0.0f, gb2 = 0.0f, gW3 = 0.0f, gb3 = 0.0f, loss = 0.0f;
  const float inv_n = 1.0f / (float)n;
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    if (j < NI) xs[j] = X[(size_t)s * NI + j];
    __syncthreads();
    const float out = mlp_forward(p, xs, h1s, h2s, red);
    const float r = out - y[s];
    const float w = r < 0.0f ? w_under : 1.0f;
    loss += w * r * r * inv_n;
    const float dout = 2.0f * w * r * inv_n;
    const float h2 = h2s[j], h1 = h1s[j];
    gW3 = fmaf(dout, h2, gW3);
    gb3 += dout;
    const float dh2 = h2 > 0.0f ? dout * p[OW3 + j] : 0.0f;
    dh2s[j] = dh2;
    gb2 += dh2;
#pragma unroll
    for (int k = 0; k < NH; ++k) gW2[k] = fmaf(dh2, h1s[k], gW2[k]);
    __syncthreads();
    // dh1_k = relu'(h1_k) sum_j W2[j][k] dh2_j, thread j plays k
    float dh1 = 0.0f;
    if (h1 > 0.0f)
      for (int jj = 0; jj < NH; ++jj) dh1 = fmaf(p[OW2 + jj * NH + j], dh2s[jj], dh1);
    gb1 += dh1;
#pragma unroll
    for (int i = 0; i < NI; ++i) gW1[i] = fmaf(dh1, xs[i], gW1[i]);
    __syncthreads();
  }
  float* g = part + (size_t)blockIdx.x * NP;
#pragma unroll
  for (int i = 0; i < NI; ++i) g[OW1 + j * NI + i] = gW1[i];
  g[OB1 + j] = gb1;
#pragma unroll
  for (int k = 0; k < NH; ++k) g[OW2 + j * NH + k] = gW2[k];
  g[OB2 + j] = gb2;
  g[OW3 + j] = gW3;
  if (j == 0) {  // every thread holds the same loss sum
    g[OB3] = gb3;
    lpart[blockIdx.x] = loss;
  }
}

// grad[i] = sum_b part[b][i] in block order; loss likewise.
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ lpart, int nb,
                             float* __restrict__ grad, float* __restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NP) {
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += part[(size_t)b * NP + i];
    grad[i] = s;
  }
  if (i == 0) {
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += lpart[b];
    *loss = s;
  }
}

// Adam (Kingma & Ba) with bias correction, step t >= 1.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int t, float lr, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NP) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)t)), vh = vi / (1.0f - powf(b2, (float)t));
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// out[s] = model(X[s]); one workgroup per sample.
__global__ __launch_bounds__(NH) void k_mlp_predict(const float* __restrict__ p, const float* __restrict__ X, int n,
                                                    float* __restrict__ out) {
  __shared__ float xs[NI], h1s[NH], h2s[NH], red[2];
  const int s = blockIdx.x;
  if (s >= n) return;
  if (threadIdx.x < NI) xs[threadIdx.x] = X[(size_t)s * NI + threadIdx.x];
  __syncthreads();
  const float o = mlp_forward(p, xs, h1s, h2s, red);
  if (threadIdx.x == 0) out[s] = o;
}

// Features of every task's bond (the state before the update): one workgroup of 64
// threads per task.  x[0..62]: the last (smallest) 63 of its chi_c Schmidt values as
// log10(lambda)/10, in order (-3 in front when fewer); x[63]: log2(kmax / chi_c) / 4, the
// room to grow, with kmax = min(chi_max, d chi_l, d chi_r) the largest rank the update can
// keep; x[64]: log2(chi_c) / 12; x[65]: log10(||G - 1||_F) / 4 of the bond's gate G, how
// far the update moves theta (dt ||H_bond|| for a real-time gate).
constexpr int kSchmidtFeatures = NI - 3;
__global__ void k_mlp_features(const BondTask* __restrict__ tasks, float* __restrict__ X) {
  const BondTask t = tasks[blockIdx.x];
  float* x = X + (size_t)blockIdx.x * NI;
  for (int q = threadIdx.x; q < kSchmidtFeatures; q += blockDim.x) {
    const int idx = t.chic - kSchmidtFeatures + q;  // last values, in order
    const double l = idx >= 0 ? t.lc[idx] : 0.0;
    x[q] = l > 1e-30 ? (float)(log10(l) / 10.0) : -3.0f;
  }
  const int d2 = t.d * t.d;
  double g2 = 0.0;
  for (int e = threadIdx.x; e < d2 * d2; e += blockDim.x) {
    zc g = t.gate[e];
    if (e % d2 == e / d2) g.x -= 1.0;
    g2 += g.x * g.x + g.y * g.y;
  }
  for (int off = 32; off > 0; off >>= 1) g2 += __shfl_xor(g2, off, 64);
  if (threadIdx.x == 0) {
    int kmax = t.chi_max < t.d * t.chil ? t.chi_max : t.d * t.chil;
    kmax = kmax < t.d * t.chir ? kmax : t.d * t.chir;
    x[NI - 3] = (float)(log2((double)kmax / (double)t.chic) / 4.0);
    x[NI - 2] = (float)(log2((double)t.chic) / 12.0);
    x[NI - 1] = (float)(log10(fmax(sqrt(g2), 1e-16)) / 4.0);
  }
}

int bad(hipError_t e) { return e != hipSuccess; }

// Device buffers of one training / gradient call.
struct MlpBuffers {
  float *p = nullptr, *X = nullptr, *y = nullptr, *part = nullptr, *lpart = nullptr, *g = nullptr, *m = nullptr,
        *v = nullptr, *loss = nullptr;
  hipStream_t st = nullptr;
  int nb = 0;
  ~MlpBuffers() {
    for (float* q : {p, X, y, part, lpart, g, m, v, loss}) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
  int init(const float* params, const float* Xh, const float* yh, int n) {
    nb = n < 256 ? n : 256;
    if (bad(hipStreamCreateWithFlags(&st, hipStreamNonBlocking))) return TEBD_ERR_HIP;
    if (bad(hipMalloc(&p, sizeof(float) * NP)) || bad(hipMalloc(&X, sizeof(float) * (size_t)n * NI)) ||
        bad(hipMalloc(&y, sizeof(float) * n)) || bad(hipMalloc(&part, sizeof(float) * (size_t)nb * NP)) ||
        bad(hipMalloc(&lpart, sizeof(float) * nb)) || bad(hipMalloc(&g, sizeof(float) * NP)) ||
        bad(hipMalloc(&m, sizeof(float) * NP)) || bad(hipMalloc(&v, sizeof(float) * NP)) ||
        bad(hipMalloc(&loss, sizeof(float))))
      return TEBD_ERR_HIP;
    if (bad(hipMemcpy(p, params, sizeof(float) * NP, hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(X, Xh, sizeof(float) * (size_t)n * NI, hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(y, yh, sizeof(float) * n, hipMemcpyHostToDevice)) || bad(hipMemset(m, 0, sizeof(float) * NP)) ||
        bad(hipMemset(v, 0, sizeof(float) * NP)))
      return TEBD_ERR_HIP;
    return TEBD_OK;
  }
  hipError_t gradient(int n, float w_under) {
    hipLaunchKernelGGL(k_mlp_fb, dim3(nb), dim3(NH), 0, st, p, X, y, n, w_under, part, lpart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NP + 255) / 256), dim3(256), 0, st, part, lpart, nb, g, loss);
    return hipGetLastError();
  }
};

}  // namespace

hipError_t launch_mlp_features(const BondTask* d_tasks, int ntasks, float* d_X, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_features, dim3(ntasks), dim3(64), 0, st, d_tasks, d_X);
  return hipGetLastError();
}

hipError_t launch_mlp_predict(const float* d_params, const float* d_X, int n, float* d_out, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_predict, dim3(n), dim3(NH), 0, st, d_params, d_X, n, d_out);
  return hipGetLastError();
}

extern "C" {

int tebd_mlp_gradient(const float* params, const float* X, const float* y, int n, float w_under, float* grad,
                      float* loss) {
  if (!params || !X || !y || !grad || !loss || n < 1) return TEBD_ERR_ARG;
  MlpBuffers b;
  int e = b.init(params, X, y, n);
  if (e) return e;
  if (bad(b.gradient(n, w_under))) return TEBD_ERR_HIP;
  if (bad(hipMemcpyAsync(grad, b.g, sizeof(float) * NP, hipMemcpyDeviceToHost, b.st)) ||
      bad(hipMemcpyAsync(loss, b.loss, sizeof(float), hipMemcpyDeviceToHost, b.st)) || bad(hipStreamSynchronize(b.st)))
    return TEBD_ERR_HIP;
  return TEBD_OK;
}

int tebd_mlp_train(float* params, const float* X, const float* y, int n, int epochs, float lr, float w_under,
                   float* loss_history) {
  if (!params || !X || !y || n < 1 || epochs < 1 || !(lr > 0.0f)) return TEBD_ERR_ARG;
  MlpBuffers b;
  int e = b.init(params, X, y, n);
  if (e) return e;
  for (int it = 1; it <= epochs; ++it) {
    if (bad(b.gradient(n, w_under))) return TEBD_ERR_HIP;
    hipLaunchKernelGGL(k_mlp_adam, dim3((NP + 255) / 256), dim3(256), 0, b.st, b.p, b.g, b.m, b.v, it, lr, 0.9f,
                       0.999f, 1e-8f);
    if (bad(hipGetLastError())) return TEBD_ERR_HIP;
    if (loss_history && bad(hipMemcpyAsync(loss_history + it - 1, b.loss, sizeof(float), hipMemcpyDeviceToHost, b.st)))
      return TEBD_ERR_HIP;
  }
  if (bad(hipMemcpyAsync(params, b.p, sizeof(float) * NP, hipMemcpyDeviceToHost, b.st)) ||
      bad(hipStreamSynchronize(b.st)))
    return TEBD_ERR_HIP;
  return TEBD_OK;
}

int tebd_mlp_predict(const float* params, const float* X, int n, float* out) {
  if (!params || !X || !out || n < 1) return TEBD_ERR_ARG;
  MlpBuffers b;
  std::vector<float> yz(n, 0.0f);
  int e = b.init(params, X, yz.data(), n);
  if (e) return e;
  float* d_out = b.y;
  if (bad(launch_mlp_predict(b.p, b.X, n, d_out, b.st))) return TEBD_ERR_HIP;
  if (bad(hipMemcpyAsync(out, d_out, sizeof(float) * n, hipMemcpyDeviceToHost, b.st)) || bad(hipStreamSynchronize(b.st)))
    return TEBD_ERR_HIP;
  return TEBD_OK;
}

}  // extern "C"
