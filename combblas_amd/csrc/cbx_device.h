// Device helpers shared by the cbx kernels (wave64 scans, block scans,
// error checks).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "cbx_internal.h"

#define CBX_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::cbx::set_error(CBX_ERR_HIP, "%s failed: %s (%s:%d)", #call,            \
                              hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

namespace cbx {

constexpr int kWave = 64;

// Index of this wave in a 1-D or 2-D grid of 1-D blocks (a 2-D grid keeps
// the per-dimension work-item count of wave-per-column launches below 2^32).
__device__ __forceinline__ int64_t global_wave() {
  return (((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
}

// Grid with one wave per item for blocks of `bs` threads.
inline dim3 wave_grid(int64_t nwaves, int bs = 256) {
  int64_t blocks = (nwaves * kWave + bs - 1) / bs;
  if (blocks < 1) blocks = 1;
  const int64_t X = (int64_t)1 << 22;
  return blocks <= X ? dim3((unsigned)blocks) : dim3((unsigned)X, (unsigned)((blocks + X - 1) / X));
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x >> 6); }

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// Exclusive prefix sum over a block of NT threads; `wsum` is LDS scratch of
// NT/64 ints.  Returns this thread's exclusive prefix and the block total.
// Contains barriers: every thread of the block must call it.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
  constexpr int NW = NT / kWave;
  const int lane = lane_id(), w = wave_id();
  const int inc = wave_incl_scan(v);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  if (w == 0) {
    int x = lane < NW ? wsum[lane] : 0;
    x = wave_incl_scan(x);
    if (lane < NW) wsum[lane] = x;
  }
  __syncthreads();
  const int base = w ? wsum[w - 1] : 0;
  total = wsum[NW - 1];
  __syncthreads();
  return base + inc - v;
}

__device__ __forceinline__ long long wave_incl_scan_ll(long long v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    long long t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// 64-bit exclusive prefix sum over a block of NT threads (wsum: NT/64 slots).
template <int NT>
__device__ __forceinline__ long long block_excl_scan_ll(long long v, long long* wsum, long long& total) {
  constexpr int NW = NT / kWave;
  const int lane = lane_id(), w = wave_id();
  const long long inc = wave_incl_scan_ll(v);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  if (w == 0) {
    long long x = lane < NW ? wsum[lane] : 0;
    x = wave_incl_scan_ll(x);
    if (lane < NW) wsum[lane] = x;
  }
  __syncthreads();
  const long long base = w ? wsum[w - 1] : 0;
  total = wsum[NW - 1];
  __syncthreads();
  return base + inc - v;
}

// Sum over a block of NT threads (result valid in every thread).
template <int NT>
__device__ __forceinline__ long long block_sum_ll(long long v, long long* wsum) {
  constexpr int NW = NT / kWave;
  v = wave_sum_ll(v);
  if (lane_id() == 0) wsum[wave_id()] = v;
  __syncthreads();
  long long t = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += wsum[i];
  __syncthreads();
  return t;
}

// Adds two per-thread counts into *pa and *pb with one device atomic each per
// workgroup (same-address device atomics serialise, so reduce first).
// Contains a barrier: every thread of the workgroup must call it.
template <int NT>
__device__ __forceinline__ void block_add_counts(long long a, long long b, unsigned long long* pa,
                                                 unsigned long long* pb) {
  __shared__ long long s[2][NT / kWave];
  a = wave_sum_ll(a);
  b = wave_sum_ll(b);
  if (lane_id() == 0) {
    s[0][wave_id()] = a;
    s[1][wave_id()] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long x = 0, y = 0;
    for (int w = 0; w < NT / kWave; ++w) {
      x += s[0][w];
      y += s[1][w];
    }
    if (x) atomicAdd(pa, (unsigned long long)x);
    if (y) atomicAdd(pb, (unsigned long long)y);
  }
}

}  // namespace cbx
