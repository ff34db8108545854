// Learned schedule selector: sparsity features of a CSR matrix and a small
// MLP (32 -> 128 -> 128 -> 24 schedule logits + 1 log-time) with forward,
// backward and Adam kernels, all FP32 and free of atomics in the network
// path, so training is bitwise reproducible.
//
// No reference counterpart: taco picks schedules by hand (the scheduling
// commands of test/tests-scheduling-eval.cpp); the selector only chooses
// which of the implemented, correct schedules runs.
//
// Features.  Two kernels fill integer statistics (exact, so the host
// restatement reproduces them bit for bit):
//   row pass over all rows: sum of lengths, sum of squared lengths, max,
//     empty rows, a 16-bin log2 histogram of lengths, entries in rows of
//     length >= 64 and >= 1024;
//   sample pass over S = min(m, 4096) evenly spaced rows (row i*m/S): their
//     entries, the entries within 64 columns of the row's diagonal column
//     (row*n/m), the entries in the row's 8x8 diagonal block, and the sum of
//     floor(4096 * |col*m - row*n| / (n*m)) (distance from the diagonal).
// The host turns them into 32 floats (taco_hip_features_from_stats).
//
// Network.  Parameters in one flat f32 array: W1[128][32], b1[128],
// W2[128][128], b2[128], W3[25][128], b3[25] (23,961 values).  Inputs are
// standardised features.  Output j < 24 is the logit of schedule class j;
// output 24 predicts log(time).  Classes [0, 13) are SpMV schedules and
// [13, 24) SpMM schedules; the softmax runs over the classes of the
// sample's operation only.  Loss per sample: cross-entropy of the measured
// best class plus kTimeWeight * (out24 - log time)^2, averaged over the
// batch.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "device.h"

namespace taco_hip {
namespace {

constexpr int kStats = 30;
constexpr int kSampleRows = 4096;
constexpr int kIn = 32, kHid = 128, kOut = 25, kClasses = 24, kSpmvClasses = 13;
constexpr int64_t kParams = (int64_t)kHid * kIn + kHid + (int64_t)kHid * kHid + kHid +
                            (int64_t)kOut * kHid + kOut;
constexpr float kTimeWeight = 0.1f;

__device__ __forceinline__ int log2_bin(int32_t len) {
  if (len <= 0) return 0;
  const int b = 32 - __clz(len);  // 1 for len 1, 2 for 2..3, ...
  return b < 15 ? b : 15;
}

// stats: [0] sum len, [1] sum len^2, [2] max len, [3] empty rows,
// [4..19] histogram, [20] entries in rows >= 64, [21] entries in rows >= 1024.
__global__ __launch_bounds__(256) void k_row_stats(int32_t m, const int32_t* __restrict__ pos,
                                                   unsigned long long* __restrict__ stats) {
  __shared__ unsigned long long s[22];
  const int tid = threadIdx.x;
  if (tid < 22) s[tid] = 0;
  __syncthreads();
  unsigned long long sum = 0, sq = 0, mx = 0, empty = 0, h64 = 0, h1k = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + tid; r < m;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t len = pos[r + 1] - pos[r];
    sum += (unsigned long long)len;
    sq += (unsigned long long)len * (unsigned long long)len;
    mx = mx > (unsigned long long)len ? mx : (unsigned long long)len;
    empty += len == 0;
    if (len >= 64) h64 += (unsigned long long)len;
    if (len >= 1024) h1k += (unsigned long long)len;
    atomicAdd(&s[4 + log2_bin(len)], 1ull);
  }
  atomicAdd(&s[0], sum);
  atomicAdd(&s[1], sq);
  atomicMax(&s[2], mx);
  atomicAdd(&s[3], empty);
  atomicAdd(&s[20], h64);
  atomicAdd(&s[21], h1k);
  __syncthreads();
  if (tid < 22) {
    if (tid == 2)
      atomicMax(&stats[2], s[2]);
    else if (s[tid])
      atomicAdd(&stats[tid], s[tid]);
  }
}

// stats: [22] sampled rows, [23] sampled entries, [24] within 64 of the
// diagonal, [25] in the 8x8 diagonal block, [26] distance sum.
__global__ __launch_bounds__(256) void k_sample_stats(int32_t m, int32_t n, int32_t S,
                                                      const int32_t* __restrict__ pos,
                                                      const int32_t* __restrict__ crd,
                                                      unsigned long long* __restrict__ stats) {
  const int lane = threadIdx.x % kWave;
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (i >= S) return;
  const int64_t row = i * (int64_t)m / S;
  const int64_t diag = row * (int64_t)n / m;
  const int64_t nm = (int64_t)n * m;
  unsigned long long cnt = 0, band = 0, block = 0, dist = 0;
  for (int32_t q = pos[row] + lane; q < pos[row + 1]; q += kWave) {
    const int64_t c = crd[q];
    ++cnt;
    band += (c - diag <= 64 && diag - c <= 64);
    block += (c / 8 == row / 8);
    const int64_t d = c * m - row * (int64_t)n;
    dist += (unsigned long long)((d < 0 ? -d : d) * 4096 / nm);
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off);
    band += __shfl_xor(band, off);
    block += __shfl_xor(block, off);
    dist += __shfl_xor(dist, off);
  }
  if (lane == 0) {
    atomicAdd(&stats[22], 1ull);
    atomicAdd(&stats[23], cnt);
    atomicAdd(&stats[24], band);
    atomicAdd(&stats[25], block);
    atomicAdd(&stats[26], dist);
  }
}

// ------------------------------------------------------------------ MLP --

// Y[b][o] = act(b[o] + sum_i X[b][i] * W[o][i]); act = ReLU when relu.
__global__ __launch_bounds__(256) void k_linear(int32_t B, int32_t I, int32_t O,
                                                const float* __restrict__ X,
                                                const float* __restrict__ W,
                                                const float* __restrict__ bias,
                                                float* __restrict__ Y, int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * O) return;
  const int32_t b = (int32_t)(t / O), o = (int32_t)(t % O);
  const float* x = X + (int64_t)b * I;
  const float* w = W + (int64_t)o * I;
  float acc = bias[o];
  for (int32_t i = 0; i < I; ++i) acc = fmaf(x[i], w[i], acc);
  Y[t] = relu && acc < 0.f ? 0.f : acc;
}

// Masked softmax cross-entropy over the classes of the sample's operation,
// plus the weighted squared error of the log-time output.  Writes dZ
// (already divided by B) and each sample's loss.
__global__ __launch_bounds__(256) void k_loss(int32_t B, const float* __restrict__ Z,
                                              const int32_t* __restrict__ cls,
                                              const float* __restrict__ logt,
                                              float* __restrict__ dZ,
                                              float* __restrict__ loss) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float* z = Z + (int64_t)b * kOut;
  float* dz = dZ + (int64_t)b * kOut;
  const int32_t c = cls[b];
  if (c < 0 || c >= kClasses) {  // not a class: no gradient, NaN loss
    for (int j = 0; j < kOut; ++j) dz[j] = 0.f;
    loss[b] = NAN;
    return;
  }
  const int lo = c < kSpmvClasses ? 0 : kSpmvClasses;
  const int hi = c < kSpmvClasses ? kSpmvClasses : kClasses;
  float mx = -INFINITY;
  for (int j = lo; j < hi; ++j) mx = fmaxf(mx, z[j]);
  float se = 0.f;
  for (int j = lo; j < hi; ++j) se += expf(z[j] - mx);
  const float lse = mx + logf(se);
  const float invB = 1.f / (float)B;
  for (int j = 0; j < kClasses; ++j) {
    float g = 0.f;
    if (j >= lo && j < hi) g = expf(z[j] - lse) - (j == c ? 1.f : 0.f);
    dz[j] = g * invB;
  }
  const float d = z[kClasses] - logt[b];
  dz[kClasses] = 2.f * kTimeWeight * d * invB;
  loss[b] = (lse - z[c]) + kTimeWeight * d * d;
}

// dX[b][i] = sum_o dY[b][o] * W[o][i], times [X[b][i] > 0] when the input
// came out of a ReLU.
__global__ __launch_bounds__(256) void k_linear_dx(int32_t B, int32_t I, int32_t O,
                                                   const float* __restrict__ dY,
                                                   const float* __restrict__ W,
                                                   const float* __restrict__ X,
                                                   float* __restrict__ dX, int relu_in) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * I) return;
  const int32_t b = (int32_t)(t / I), i = (int32_t)(t % I);
  const float* dy = dY + (int64_t)b * O;
  float acc = 0.f;
  for (int32_t o = 0; o < O; ++o) acc = fmaf(dy[o], W[(int64_t)o * I + i], acc);
  dX[t] = relu_in && !(X[t] > 0.f) ? 0.f : acc;
}

// dW[o][i] = sum_b dY[b][o] * X[b][i]; thread (o, I) computes db[o].
__global__ __launch_bounds__(256) void k_linear_dw(int32_t B, int32_t I, int32_t O,
                                                   const float* __restrict__ dY,
                                                   const float* __restrict__ X,
                                                   float* __restrict__ dW,
                                                   float* __restrict__ db) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)O * (I + 1)) return;
  const int32_t o = (int32_t)(t / (I + 1)), i = (int32_t)(t % (I + 1));
  float acc = 0.f;
  if (i < I) {
    for (int32_t b = 0; b < B; ++b)
      acc = fmaf(dY[(int64_t)b * O + o], X[(int64_t)b * I + i], acc);
    dW[(int64_t)o * I + i] = acc;
  } else {
    for (int32_t b = 0; b < B; ++b) acc += dY[(int64_t)b * O + o];
    db[o] = acc;
  }
}

__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ p,
                                              const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2,
                                              float lr, float beta1, float beta2, float eps,
                                              float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float a = beta1 * m1[i] + (1.f - beta1) * gi;
  const float b = beta2 * m2[i] + (1.f - beta2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace taco_hip

using namespace taco_hip;

// Offsets of the parameter blocks in the flat array.
namespace {
constexpr int64_t kW1 = 0, kB1 = kW1 + (int64_t)kHid * kIn, kW2 = kB1 + kHid,
                  kB2 = kW2 + (int64_t)kHid * kHid, kW3 = kB2 + kHid,
                  kB3 = kW3 + (int64_t)kOut * kHid;
static_assert(kB3 + kOut == kParams, "parameter layout");
}  // namespace

struct taco_hip_mlp {
  float* params = nullptr;  // kParams
  float* grads = nullptr;   // kParams
  float* m1 = nullptr;      // Adam moments
  float* m2 = nullptr;
  int64_t steps = 0;
  int32_t cap = 0;  // batch capacity of the activation buffers
  float *h1 = nullptr, *h2 = nullptr, *z = nullptr, *dz = nullptr, *d2 = nullptr,
        *d1 = nullptr, *loss = nullptr;
};

namespace {
void mlp_release_acts(taco_hip_mlp* n) {
  for (float* q : {n->h1, n->h2, n->z, n->dz, n->d2, n->d1, n->loss})
    if (q) (void)hipFree(q);
  n->h1 = n->h2 = n->z = n->dz = n->d2 = n->d1 = n->loss = nullptr;
  n->cap = 0;
}

int mlp_reserve(taco_hip_mlp* n, int32_t B) {
  if (B <= n->cap) return 0;
  mlp_release_acts(n);
  const size_t b = (size_t)B;
  if (hipMalloc((void**)&n->h1, sizeof(float) * b * kHid) ||
      hipMalloc((void**)&n->h2, sizeof(float) * b * kHid) ||
      hipMalloc((void**)&n->z, sizeof(float) * b * kOut) ||
      hipMalloc((void**)&n->dz, sizeof(float) * b * kOut) ||
      hipMalloc((void**)&n->d2, sizeof(float) * b * kHid) ||
      hipMalloc((void**)&n->d1, sizeof(float) * b * kHid) ||
      hipMalloc((void**)&n->loss, sizeof(float) * b)) {
    mlp_release_acts(n);
    return fail("mlp: device allocation failed");
  }
  n->cap = B;
  return 0;
}

int mlp_forward(taco_hip_mlp* n, const float* X, int32_t B, hipStream_t st) {
  const float* P = n->params;
  hipLaunchKernelGGL(k_linear, dim3(grid_for((int64_t)B * kHid)), dim3(256), 0, st, B, kIn,
                     kHid, X, P + kW1, P + kB1, n->h1, 1);
  hipLaunchKernelGGL(k_linear, dim3(grid_for((int64_t)B * kHid)), dim3(256), 0, st, B, kHid,
                     kHid, n->h1, P + kW2, P + kB2, n->h2, 1);
  hipLaunchKernelGGL(k_linear, dim3(grid_for((int64_t)B * kOut)), dim3(256), 0, st, B, kHid,
                     kOut, n->h2, P + kW3, P + kB3, n->z, 0);
  TH_LAUNCH_CHECK("mlp forward");
  return 0;
}

// Forward, loss and gradients into n->grads; the batch loss (mean) to *loss.
int mlp_gradients(taco_hip_mlp* n, const float* X, const int32_t* cls, const float* logt,
                  int32_t B, float* loss, hipStream_t st) {
  if (int rc = mlp_forward(n, X, B, st)) return rc;
  const float* P = n->params;
  float* G = n->grads;
  hipLaunchKernelGGL(k_loss, dim3(grid_for(B)), dim3(256), 0, st, B, n->z, cls, logt, n->dz,
                     n->loss);
  // layer 3
  hipLaunchKernelGGL(k_linear_dw, dim3(grid_for((int64_t)kOut * (kHid + 1))), dim3(256), 0,
                     st, B, kHid, kOut, n->dz, n->h2, G + kW3, G + kB3);
  hipLaunchKernelGGL(k_linear_dx, dim3(grid_for((int64_t)B * kHid)), dim3(256), 0, st, B,
                     kHid, kOut, n->dz, P + kW3, n->h2, n->d2, 1);
  // layer 2
  hipLaunchKernelGGL(k_linear_dw, dim3(grid_for((int64_t)kHid * (kHid + 1))), dim3(256), 0,
                     st, B, kHid, kHid, n->d2, n->h1, G + kW2, G + kB2);
  hipLaunchKernelGGL(k_linear_dx, dim3(grid_for((int64_t)B * kHid)), dim3(256), 0, st, B,
                     kHid, kHid, n->d2, P + kW2, n->h1, n->d1, 1);
  // layer 1
  hipLaunchKernelGGL(k_linear_dw, dim3(grid_for((int64_t)kHid * (kIn + 1))), dim3(256), 0, st,
                     B, kIn, kHid, n->d1, X, G + kW1, G + kB1);
  TH_LAUNCH_CHECK("mlp backward");
  if (loss) {
    std::vector<float> l((size_t)B);
    TH_CHECK(hipMemcpyAsync(l.data(), n->loss, sizeof(float) * B, hipMemcpyDeviceToHost, st));
    TH_CHECK(hipStreamSynchronize(st));
    double s = 0;
    for (float v : l) s += v;
    *loss = (float)(s / B);
  }
  return 0;
}
}  // namespace

extern "C" {

int taco_hip_csr_stats(taco_tensor_t* A, int64_t* stats) {
  TH_REQUIRE(stats, "features: null output");
  CsrView a;
  if (int rc = view_csr(A, "A", &a)) return rc;
  hipStream_t st = current_stream();
  auto* d = (unsigned long long*)workspace(sizeof(unsigned long long) * 32, 2);
  if (!d) return 1;
  TH_CHECK(hipMemsetAsync(d, 0, sizeof(unsigned long long) * 32, st));
  if (a.m > 0) {
    const unsigned rg = (unsigned)std::min<int64_t>((a.m + 255) / 256, 2048);
    hipLaunchKernelGGL(k_row_stats, dim3(rg), dim3(256), 0, st, a.m, a.pos, d);
    const int32_t S = a.m < kSampleRows ? a.m : kSampleRows;
    if (a.n > 0)
      hipLaunchKernelGGL(k_sample_stats, dim3(grid_for((int64_t)S * kWave)), dim3(256), 0, st,
                         a.m, a.n, S, a.pos, a.crd, d);
    TH_LAUNCH_CHECK("feature kernels");
  }
  unsigned long long h[32];
  TH_CHECK(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  stats[0] = a.m;
  stats[1] = a.n;
  stats[2] = a.nnz;
  for (int i = 0; i < 27; ++i) stats[3 + i] = (int64_t)h[i];
  return 0;
}

int taco_hip_features_from_stats(const int64_t* s, int32_t op, int32_t K, float* f) {
  TH_REQUIRE(s && f && (op == 0 || op == 1), "features: bad arguments");
  // s: [0] m, [1] n, [2] nnz, [3 + k] kernel statistic k.
  const double m = (double)s[0], n = (double)s[1], nnz = (double)s[2];
  const double mean = m > 0 ? (double)s[3] / m : 0.0;
  const double var = m > 0 ? std::max(0.0, (double)s[4] / m - mean * mean) : 0.0;
  const double mx = (double)s[5];
  const double sampled = (double)s[26];
  std::memset(f, 0, sizeof(float) * kIn);
  f[0] = (float)op;
  f[1] = op == 1 ? (float)std::log2((double)(K > 0 ? K : 1)) : 0.f;
  f[2] = (float)std::log2(m + 1);
  f[3] = (float)std::log2(n + 1);
  f[4] = (float)std::log2(nnz + 1);
  f[5] = (float)std::log2(1 + mean);
  f[6] = (float)(mean > 0 ? std::sqrt(var) / mean : 0.0);
  f[7] = (float)std::log2(1 + mx);
  f[8] = (float)std::log2(1 + (mean > 0 ? mx / mean : 0.0));
  f[9] = (float)(m > 0 ? (double)s[6] / m : 0.0);
  for (int k = 0; k < 16; ++k) f[10 + k] = (float)(m > 0 ? (double)s[7 + k] / m : 0.0);
  f[26] = (float)(nnz > 0 ? (double)s[23] / nnz : 0.0);
  f[27] = (float)(nnz > 0 ? (double)s[24] / nnz : 0.0);
  f[28] = (float)(sampled > 0 ? (double)s[29] / (4096.0 * sampled) : 0.0);
  f[29] = (float)(sampled > 0 ? (double)s[27] / sampled : 0.0);
  f[30] = (float)(sampled > 0 ? (double)s[28] / sampled : 0.0);
  const double row_bytes = op == 1 ? 4.0 * (K > 0 ? K : 1) : 8.0;
  f[31] = (float)std::log2(1 + n * row_bytes);
  return 0;
}

int taco_hip_features(taco_tensor_t* A, int32_t op, int32_t K, float* out) {
  int64_t s[kStats];
  if (int rc = taco_hip_csr_stats(A, s)) return rc;
  return taco_hip_features_from_stats(s, op, K, out);
}

int64_t taco_hip_mlp_param_count(void) { return kParams; }

int taco_hip_mlp_create(taco_hip_mlp** out) {
  TH_REQUIRE(out, "mlp: null output");
  *out = nullptr;
  auto* n = new taco_hip_mlp();
  const size_t b = sizeof(float) * (size_t)kParams;
  if (hipMalloc((void**)&n->params, b) || hipMalloc((void**)&n->grads, b) ||
      hipMalloc((void**)&n->m1, b) || hipMalloc((void**)&n->m2, b)) {
    taco_hip_mlp_destroy(n);
    return fail("mlp: device allocation failed");
  }
  hipStream_t st = current_stream();
  for (float* q : {n->params, n->grads, n->m1, n->m2})
    TH_CHECK(hipMemsetAsync(q, 0, b, st));
  TH_CHECK(hipStreamSynchronize(st));
  *out = n;
  return 0;
}

void taco_hip_mlp_destroy(taco_hip_mlp* n) {
  if (!n) return;
  mlp_release_acts(n);
  for (float* q : {n->params, n->grads, n->m1, n->m2})
    if (q) (void)hipFree(q);
  delete n;
}

int taco_hip_mlp_set_params(taco_hip_mlp* n, const float* host) {
  TH_REQUIRE(n && host, "mlp: null argument");
  hipStream_t st = current_stream();
  const size_t b = sizeof(float) * (size_t)kParams;
  TH_CHECK(hipMemcpyAsync(n->params, host, b, hipMemcpyHostToDevice, st));
  TH_CHECK(hipMemsetAsync(n->m1, 0, b, st));
  TH_CHECK(hipMemsetAsync(n->m2, 0, b, st));
  TH_CHECK(hipStreamSynchronize(st));
  n->steps = 0;
  return 0;
}

int taco_hip_mlp_get_params(taco_hip_mlp* n, float* host) {
  TH_REQUIRE(n && host, "mlp: null argument");
  hipStream_t st = current_stream();
  TH_CHECK(hipMemcpyAsync(host, n->params, sizeof(float) * (size_t)kParams,
                          hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  return 0;
}

int taco_hip_mlp_forward(taco_hip_mlp* n, const float* X, int32_t B, float* Z) {
  TH_REQUIRE(n && X && Z && B >= 0, "mlp: bad forward arguments");
  if (B == 0) return 0;
  if (int rc = mlp_reserve(n, B)) return rc;
  hipStream_t st = current_stream();
  if (int rc = mlp_forward(n, X, B, st)) return rc;
  TH_CHECK(hipMemcpyAsync(Z, n->z, sizeof(float) * (size_t)B * kOut, hipMemcpyDeviceToDevice,
                          st));
  return finish("taco_hip_mlp_forward");
}

int taco_hip_mlp_gradients(taco_hip_mlp* n, const float* X, const int32_t* cls,
                           const float* logt, int32_t B, float* grads_host, float* loss) {
  TH_REQUIRE(n && X && cls && logt && B > 0, "mlp: bad gradient arguments");
  if (int rc = mlp_reserve(n, B)) return rc;
  hipStream_t st = current_stream();
  if (int rc = mlp_gradients(n, X, cls, logt, B, loss, st)) return rc;
  if (grads_host) {
    TH_CHECK(hipMemcpyAsync(grads_host, n->grads, sizeof(float) * (size_t)kParams,
                            hipMemcpyDeviceToHost, st));
    TH_CHECK(hipStreamSynchronize(st));
  }
  return 0;
}

int taco_hip_mlp_train_step(taco_hip_mlp* n, const float* X, const int32_t* cls,
                            const float* logt, int32_t B, float lr, float weight_decay,
                            float* loss) {
  TH_REQUIRE(n && X && cls && logt && B > 0, "mlp: bad training arguments");
  if (int rc = mlp_reserve(n, B)) return rc;
  hipStream_t st = current_stream();
  if (int rc = mlp_gradients(n, X, cls, logt, B, loss, st)) return rc;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  n->steps += 1;
  const float c1 = 1.f - std::pow(beta1, (float)n->steps);
  const float c2 = 1.f - std::pow(beta2, (float)n->steps);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(kParams)), dim3(256), 0, st, kParams, n->params,
                     n->grads, n->m1, n->m2, lr, beta1, beta2, eps, c1, c2, weight_decay);
  TH_LAUNCH_CHECK("mlp adam");
  return finish("taco_hip_mlp_train_step");
}

}  // extern "C"
