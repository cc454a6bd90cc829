// Per-column nnz predictor: a 6 -> 64 -> 64 -> 1 ReLU MLP in FP32 with its
// own forward, backward and Adam kernels (no libraries).
//
// Parameter layout (CBX_MLP_PARAMS floats):
//   W1[64][6] b1[64] W2[64][64] b2[64] W3[64] b3[1] mu[6] sd[6]
// The network reads standardised features (x - mu) / sd and predicts
// log2(1 + nnz_j); the loss is the mean squared error of that value.
//
// Training runs mini-batches: k_mlp_fwd_bwd (one thread per sample) keeps the
// per-sample activations and back-propagated deltas in global scratch, then
// k_mlp_outer reduces the weight gradients as sums of per-sample outer
// products over the batch (deterministic order), and k_mlp_adam updates.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "cbx_device.h"
#include "cbx_spgemm.h"

namespace cbx {

constexpr int kF = CBX_MLP_FEATURES;
constexpr int kH = CBX_MLP_HIDDEN;
constexpr int oW1 = 0, oB1 = oW1 + kH * kF, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH,
              oMU = oB3 + 1, oSD = oMU + kF, kP = oSD + kF;
static_assert(kP == CBX_MLP_PARAMS, "parameter layout");
constexpr int kTrainable = oMU;  // W1..b3

namespace {

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

// Inference: weights staged in LDS, one thread per sample.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ P, const float* __restrict__ X, int64_t n,
                                                   float* __restrict__ out) {
  __shared__ float w[kP];
  for (int i = threadIdx.x; i < kP; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kF];
#pragma unroll
  for (int f = 0; f < kF; ++f) x[f] = (X[s * kF + f] - w[oMU + f]) / w[oSD + f];
  float h1[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    float a = w[oB1 + j];
#pragma unroll
    for (int f = 0; f < kF; ++f) a += w[oW1 + j * kF + f] * x[f];
    h1[j] = relu(a);
  }
  float o = w[oB3];
  for (int j = 0; j < kH; ++j) {
    float a = w[oB2 + j];
#pragma unroll
    for (int i = 0; i < kH; ++i) a += w[oW2 + j * kH + i] * h1[i];
    o += w[oW3 + j] * relu(a);
  }
  out[s] = o;
}

// Forward + backward of one mini-batch.  Sample t of the batch is row
// idx[t] of X/y.  Stores the standardised input, h1, h2 and the deltas
// d1 = dL/d(pre-activation 1), d2 = dL/d(pre-activation 2), d3 = dL/d(out).
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ X,
                                                     const float* __restrict__ Y, const int64_t* __restrict__ idx,
                                                     int B, float* __restrict__ xs, float* __restrict__ h1s,
                                                     float* __restrict__ h2s, float* __restrict__ d1s,
                                                     float* __restrict__ d2s, float* __restrict__ d3s,
                                                     float* __restrict__ sqerr) {
  __shared__ float w[kP];
  for (int i = threadIdx.x; i < kP; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B) return;
  const int64_t s = idx ? idx[t] : t;
  float x[kF];
#pragma unroll
  for (int f = 0; f < kF; ++f) {
    x[f] = (X[s * kF + f] - w[oMU + f]) / w[oSD + f];
    xs[t * kF + f] = x[f];
  }
  float h1[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    float a = w[oB1 + j];
#pragma unroll
    for (int f = 0; f < kF; ++f) a += w[oW1 + j * kF + f] * x[f];
    h1[j] = relu(a);
    h1s[t * kH + j] = h1[j];
  }
  float o = w[oB3];
  for (int j = 0; j < kH; ++j) {
    float a = w[oB2 + j];
#pragma unroll
    for (int i = 0; i < kH; ++i) a += w[oW2 + j * kH + i] * h1[i];
    const float h = relu(a);
    h2s[t * kH + j] = h;
    o += w[oW3 + j] * h;
  }
  const float e = o - Y[s];
  sqerr[t] = e * e;
  const float d3 = 2.f * e / (float)B;  // d(mean squared error)/d(out)
  d3s[t] = d3;
  float g1[kH];
#pragma unroll
  for (int i = 0; i < kH; ++i) g1[i] = 0.f;
  for (int j = 0; j < kH; ++j) {
    const float d2 = h2s[t * kH + j] > 0.f ? d3 * w[oW3 + j] : 0.f;
    d2s[t * kH + j] = d2;
#pragma unroll
    for (int i = 0; i < kH; ++i) g1[i] += d2 * w[oW2 + j * kH + i];
  }
#pragma unroll
  for (int i = 0; i < kH; ++i) d1s[t * kH + i] = h1[i] > 0.f ? g1[i] : 0.f;
}

// G[a][b] = sum_t L[t][a] * R[t][b]  (R = nullptr means a column of ones),
// one thread per (a, b), samples summed in order.
__global__ void k_mlp_outer(const float* __restrict__ L, int na, const float* __restrict__ R, int nb, int B,
                            float* __restrict__ G) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= na * nb) return;
  const int a = q / nb, b = q % nb;
  float acc = 0.f;
  for (int t = 0; t < B; ++t) acc += L[t * na + a] * (R ? R[t * nb + b] : 1.f);
  G[q] = acc;
}

__global__ void k_mlp_sum(const float* __restrict__ v, int n, float* __restrict__ out) {
  if (blockIdx.x != 0) return;
  float a = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a += v[i];
  __shared__ float red[1024];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                           float* __restrict__ V, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Mini-batch sample indices from a counter-based hash (seed, step).
__global__ void k_mlp_batch(int64_t n, int B, uint64_t seed, uint64_t step, int64_t* __restrict__ idx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B) return;
  uint64_t z = seed ^ (step * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)t * 0xD1B54A32D192ED03ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[t] = (int64_t)(z % (uint64_t)n);
}

struct Scratch {
  float *xs = nullptr, *h1 = nullptr, *h2 = nullptr, *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *se = nullptr,
        *g = nullptr, *loss = nullptr;
  int64_t* idx = nullptr;
  hipError_t alloc(int B) {
    hipError_t e = hipSuccess;
    if ((e = hipMalloc(&xs, sizeof(float) * B * kF))) return e;
    if ((e = hipMalloc(&h1, sizeof(float) * B * kH))) return e;
    if ((e = hipMalloc(&h2, sizeof(float) * B * kH))) return e;
    if ((e = hipMalloc(&d1, sizeof(float) * B * kH))) return e;
    if ((e = hipMalloc(&d2, sizeof(float) * B * kH))) return e;
    if ((e = hipMalloc(&d3, sizeof(float) * B))) return e;
    if ((e = hipMalloc(&se, sizeof(float) * B))) return e;
    if ((e = hipMalloc(&g, sizeof(float) * kP))) return e;
    if ((e = hipMalloc(&loss, sizeof(float)))) return e;
    if ((e = hipMalloc(&idx, sizeof(int64_t) * B))) return e;
    return hipMemset(g, 0, sizeof(float) * kP);
  }
  ~Scratch() {
    for (void* p : {(void*)xs, (void*)h1, (void*)h2, (void*)d1, (void*)d2, (void*)d3, (void*)se, (void*)g,
                    (void*)loss, (void*)idx})
      if (p) (void)hipFree(p);
  }
};

inline unsigned gb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// One forward/backward pass over `B` samples (indices idx or the first B rows)
// leaving the trainable-parameter gradient in sc.g and the loss in *loss.
void grad_step(const float* P, const float* X, const float* Y, const int64_t* idx, int B, Scratch& sc,
               hipStream_t s) {
  k_mlp_fwd_bwd<<<gb(B, 256), 256, 0, s>>>(P, X, Y, idx, B, sc.xs, sc.h1, sc.h2, sc.d1, sc.d2, sc.d3, sc.se);
  k_mlp_outer<<<gb(kH * kF, 256), 256, 0, s>>>(sc.d1, kH, sc.xs, kF, B, sc.g + oW1);
  k_mlp_outer<<<gb(kH, 256), 256, 0, s>>>(sc.d1, kH, nullptr, 1, B, sc.g + oB1);
  k_mlp_outer<<<gb(kH * kH, 256), 256, 0, s>>>(sc.d2, kH, sc.h1, kH, B, sc.g + oW2);
  k_mlp_outer<<<gb(kH, 256), 256, 0, s>>>(sc.d2, kH, nullptr, 1, B, sc.g + oB2);
  k_mlp_outer<<<gb(kH, 256), 256, 0, s>>>(sc.d3, 1, sc.h2, kH, B, sc.g + oW3);
  k_mlp_outer<<<1, 64, 0, s>>>(sc.d3, 1, nullptr, 1, B, sc.g + oB3);
  k_mlp_sum<<<1, 1024, 0, s>>>(sc.se, B, sc.loss);
}

}  // namespace

void mlp_infer_launch(const float* params, const float* X, int64_t n, float* out, hipStream_t s) {
  if (n > 0) k_mlp_infer<<<gb(n, 256), 256, 0, s>>>(params, X, n, out);
}

}  // namespace cbx

using namespace cbx;

extern "C" cbx_status cbx_mlp_forward(const float* params, const float* X, int64_t n, float* out, void* stream) {
  CBX_RETURN_IF(!params || (n > 0 && (!X || !out)), CBX_ERR_INVALID, "cbx_mlp_forward: null argument");
  if (n == 0) return CBX_OK;
  k_mlp_infer<<<gb(n, 256), 256, 0, (hipStream_t)stream>>>(params, X, n, out);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CBX_OK;
}

extern "C" cbx_status cbx_mlp_gradient(const float* params, const float* X, const float* Y, int32_t n, float* grad,
                                       float* loss) {
  CBX_RETURN_IF(!params || !X || !Y || !grad || !loss || n <= 0, CBX_ERR_INVALID, "cbx_mlp_gradient: bad argument");
  Scratch sc;
  CBX_HIP(sc.alloc(n));
  grad_step(params, X, Y, nullptr, n, sc, nullptr);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipMemcpy(grad, sc.g, sizeof(float) * kP, hipMemcpyDeviceToDevice));
  float l = 0;
  CBX_HIP(hipMemcpy(&l, sc.loss, sizeof(float), hipMemcpyDeviceToHost));
  *loss = l / (float)n;
  return CBX_OK;
}

extern "C" cbx_status cbx_mlp_train(float* params, const float* X, const float* Y, int64_t n,
                                    const cbx_mlp_train_opts* o, float* loss_curve) {
  CBX_RETURN_IF(!params || !X || !Y || !o || n <= 0, CBX_ERR_INVALID, "cbx_mlp_train: bad argument");
  CBX_RETURN_IF(o->batch <= 0 || o->steps <= 0, CBX_ERR_INVALID, "cbx_mlp_train: steps and batch must be positive");
  const int B = (int)std::min<int64_t>(o->batch, n);
  Scratch sc;
  CBX_HIP(sc.alloc(B));
  float *M = nullptr, *V = nullptr;
  CBX_HIP(hipMalloc(&M, sizeof(float) * kP));
  CBX_HIP(hipMalloc(&V, sizeof(float) * kP));
  CBX_HIP(hipMemset(M, 0, sizeof(float) * kP));
  CBX_HIP(hipMemset(V, 0, sizeof(float) * kP));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.f, c2 = 1.f;
  hipStream_t s = nullptr;
  for (int step = 0; step < o->steps; ++step) {
    k_mlp_batch<<<gb(B, 256), 256, 0, s>>>(n, B, o->seed, (uint64_t)step, sc.idx);
    grad_step(params, X, Y, sc.idx, B, sc, s);
    c1 *= b1;
    c2 *= b2;
    k_mlp_adam<<<gb(kTrainable, 256), 256, 0, s>>>(params, sc.g, M, V, kTrainable, o->lr, b1, b2, eps, 1.f - c1,
                                                   1.f - c2);
    if (loss_curve) {
      float l = 0;
      (void)hipMemcpyAsync(&l, sc.loss, sizeof(float), hipMemcpyDeviceToHost, s);
      (void)hipStreamSynchronize(s);
      loss_curve[step] = l / (float)B;
    }
  }
  hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(M);
  (void)hipFree(V);
  CBX_HIP(e);
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}
