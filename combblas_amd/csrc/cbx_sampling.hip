// Sampling estimate of the product's column sizes on the device: the
// reference's estimateNNZ_sampling (include/CombBLAS/mtSpGEMM.h:936-1054), a
// min-hash (Cohen) estimator, kept per column instead of summed.
//
//   samples[i], i < m * r    exponential(1) draws, one stream for all rows:
//                            the i-th draw of the reference's generator
//                            (std::default_random_engine, i.e. minstd_rand0
//                            x <- 16807 x mod 2^31-1 from x = 1, turned into
//                            u = float(x - 1) / 2^31 and -log(1 - u)).  Every
//                            thread jumps to its first index by modular
//                            exponentiation and steps the recurrence, so the
//                            stream does not depend on the launch shape.
//   mid[p][k]                min over the rows of A's p-th stored column
//   final_j[k]               min over B(:,j)'s rows of mid[.][k]
//   est_j                    (r - 1) / (final_j[0] + ... + final_j[r-1])
//
// The result is the reference's per-column estimate up to the last bit of
// -log (the device logf against the host's), which moves an estimate by a
// few units in the last place.  This is the non-learned predictor the
// learned single pass is compared with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "cbx_device.h"
#include "cbx_spgemm.h"

namespace cbx {

namespace {

constexpr uint32_t kMinstdM = 2147483647u;  // 2^31 - 1
constexpr uint32_t kMinstdA = 16807u;

__device__ __forceinline__ uint32_t mulmod_m31(uint64_t a, uint64_t b) {
  uint64_t t = a * b;  // < 2^62
  t = (t & kMinstdM) + (t >> 31);
  t = (t & kMinstdM) + (t >> 31);
  return (uint32_t)(t >= kMinstdM ? t - kMinstdM : t);
}

// 16807^e mod 2^31-1
__device__ __forceinline__ uint32_t pow_a(uint64_t e) {
  uint64_t base = kMinstdA, r = 1;
  while (e) {
    if (e & 1) r = mulmod_m31(r, base);
    base = mulmod_m31(base, base);
    e >>= 1;
  }
  return (uint32_t)r;
}

constexpr int kSampPerThread = 64;

// samples[i] for i < count: the (i+1)-th output x of the generator seeded
// with 1 (x_{i+1} = 16807^{i+1} mod 2^31-1), as an exponential(1) draw.
__global__ void k_exp_samples(float* __restrict__ samples, int64_t count) {
  const int64_t i0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * kSampPerThread;
  if (i0 >= count) return;
  uint32_t x = pow_a((uint64_t)i0 + 1);
  const int64_t n = min<int64_t>(kSampPerThread, count - i0);
  for (int64_t j = 0; j < n; ++j) {
    float u = (float)(x - 1u) * 4.656612873077392578125e-10f;  // / 2^31, exact
    if (u >= 1.0f) u = 0x1.fffffep-1f;                             // largest float below 1
    samples[i0 + j] = -logf(1.0f - u);
    x = mulmod_m31(x, kMinstdA);
  }
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// mid[p*r + k] = min over the rows of A's p-th stored column (one wave each).
template <int R>
__global__ void k_minhash_mid(const int64_t* __restrict__ cpA, int64_t nzcA, const int32_t* __restrict__ rowsA,
                              const float* __restrict__ samples, float* __restrict__ mid) {
  const int64_t p = global_wave();
  if (p >= nzcA) return;
  float mn[R];
#pragma unroll
  for (int k = 0; k < R; ++k) mn[k] = FLT_MAX;
  for (int64_t e = cpA[p] + lane_id(); e < cpA[p + 1]; e += kWave) {
    const float* s = samples + (int64_t)rowsA[e] * R;
#pragma unroll
    for (int k = 0; k < R; ++k) mn[k] = fminf(mn[k], s[k]);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, kWave));
  }
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < R; ++k) mid[p * R + k] = mn[k];
  }
}

// est[j] for B's stored column j: the two-level minimum over B(:,j)'s rows
// (bpos: their stored-column positions in A, -1 for an empty A column).
template <int R>
__global__ void k_minhash_est(const int64_t* __restrict__ cpB, int64_t nzcB, const int32_t* __restrict__ bpos,
                              const float* __restrict__ mid, float* __restrict__ est) {
  const int64_t j = global_wave();
  if (j >= nzcB) return;
  float mn[R];
#pragma unroll
  for (int k = 0; k < R; ++k) mn[k] = FLT_MAX;
  for (int64_t e = cpB[j] + lane_id(); e < cpB[j + 1]; e += kWave) {
    const int32_t p = bpos[e];
    if (p < 0) continue;
#pragma unroll
    for (int k = 0; k < R; ++k) mn[k] = fminf(mn[k], mid[(int64_t)p * R + k]);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, kWave));
  }
  if (lane_id() == 0) {
    float c = 0.0f;
#pragma unroll
    for (int k = 0; k < R; ++k) c += mn[k];
    est[j] = (float)(R - 1) / c;
  }
}

template <int R>
void launch_estimate(const cbx_dcsc* A, const cbx_dcsc* B, const int32_t* rowsA, const int32_t* bpos,
                     float* samples, float* mid, float* est, hipStream_t s) {
  const int64_t count = A->m * R;
  const int64_t threads = (count + kSampPerThread - 1) / kSampPerThread;
  k_exp_samples<<<(unsigned)std::max<int64_t>((threads + 255) / 256, 1), 256, 0, s>>>(samples, count);
  k_minhash_mid<R><<<wave_grid(A->nzc), 256, 0, s>>>(A->cp, A->nzc, rowsA, samples, mid);
  k_minhash_est<R><<<wave_grid(B->nzc), 256, 0, s>>>(B->cp, B->nzc, bpos, mid, est);
}

}  // namespace

size_t sampling_scratch_floats(const cbx_dcsc* A, int rounds) {
  return (size_t)(A->m * rounds) + (size_t)(A->nzc * rounds);
}

cbx_status sampling_estimate_launch(const cbx_dcsc* A, const cbx_dcsc* B, int rounds, const int32_t* rowsA,
                                    const int32_t* bpos, float* scratch, float* est, hipStream_t s) {
  float* samples = scratch;
  float* mid = scratch + A->m * rounds;
  switch (rounds) {
#define CBX_R(N)                                                         \
  case N:                                                                \
    launch_estimate<N>(A, B, rowsA, bpos, samples, mid, est, s);         \
    break;
    CBX_R(2) CBX_R(3) CBX_R(4) CBX_R(5) CBX_R(6) CBX_R(7) CBX_R(8)
#undef CBX_R
    default:
      return set_error(CBX_ERR_LIMIT, "sampling estimate: %d rounds (2 to 8 are provided)", rounds);
  }
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}

}  // namespace cbx
