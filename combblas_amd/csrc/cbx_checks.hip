// Device-side structural validation and per-column checksums of a DCSC.
//
// The checksum of stored column j is sum_i u(i) * C(i,j) over its entries,
// in wrapping uint64 arithmetic, with u(i) = mix64(i ^ salt).  It is linear
// in C, so for C = A*B (plus-times) it equals sum_k (u^T A)_k * B(k,j),
// which a caller can evaluate on the host in O(nnz(A) + nnz(B)): a
// full-size, bit-exact check of every column without materialising C on
// the host.  Bit patterns of floating values are summed as integers, so the
// check is exact only for integer-valued semirings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "cbx_device.h"
#include "cbx_spgemm.h"

namespace cbx {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <class T>
__device__ __forceinline__ uint64_t val_bits(T v) {
  if constexpr (sizeof(T) == 8) return __builtin_bit_cast(uint64_t, v);
  else if constexpr (sizeof(T) == 4) return (uint64_t)__builtin_bit_cast(uint32_t, v);
  else return (uint64_t)v;
}

template <class T>
__global__ void k_col_checksum(const int64_t* __restrict__ cp, const int64_t* __restrict__ ir,
                               const T* __restrict__ val, int64_t nzc, uint64_t salt, uint64_t* __restrict__ out) {
  const int64_t col = global_wave();
  if (col >= nzc) return;
  uint64_t acc = 0;
  for (int64_t e = cp[col] + lane_id(); e < cp[col + 1]; e += kWave)
    acc += mix64((uint64_t)ir[e] ^ salt) * val_bits(val[e]);
  acc = (uint64_t)wave_sum_ll((long long)acc);
  if (lane_id() == 0) out[col] = acc;
}

// Same rules as the SpGEMM input validation (rows strictly increasing within
// a column and below m, jc strictly increasing and below n, cp consistent),
// in two fully parallel passes.  Per entry: rows in range, and every adjacent
// pair with ir[e+1] <= ir[e] (a descent) is counted.  Per column: cp and jc,
// and the descents that fall on a column start (ir[cp[q]] <= ir[cp[q]-1]).
// With cp valid, the column starts are exactly the positions where a descent
// is allowed, so the rows ascend within every column iff both counts agree.
// cnt[0] = violations, cnt[1] = descents, cnt[2] = descents at column starts.
// Counts are summed per workgroup before one atomic each: same-address device
// atomics serialise, so there are few of them (at most kCheckBlocks).
constexpr int kCheckNT = 256, kCheckBlocks = 1024;

__global__ __launch_bounds__(kCheckNT) void k_check_entries(const int64_t* __restrict__ ir, int64_t nnz, int64_t nrows,
                                unsigned long long* __restrict__ cnt) {
  long long bad = 0, desc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += stride) {
    const int64_t r = ir[e];
    bad += r < 0 || r >= nrows;
    if (e + 1 < nnz) desc += ir[e + 1] <= r;
  }
  block_add_counts<kCheckNT>(bad, desc, &cnt[0], &cnt[1]);
}

__global__ __launch_bounds__(kCheckNT) void k_check_columns(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc,
                                const int64_t* __restrict__ ir, int64_t nzc, int64_t nnz, int64_t ncols,
                                unsigned long long* __restrict__ cnt) {
  long long bad = 0, desc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nzc; q += stride) {
    const int64_t b = cp[q], e = cp[q + 1], j = jc[q];
    bool err = b < 0 || e <= b || e > nnz || j < 0 || j >= ncols;
    if (q + 1 < nzc) err |= jc[q + 1] <= j;
    if (q == 0) err |= b != 0;
    if (q == nzc - 1) err |= e != nnz;
    bad += err;
    // a column start inside the entries (a malformed cp elsewhere may still
    // put b at 0: then there is nothing before it to compare with)
    if (!err && q > 0 && b > 0) desc += ir[b] <= ir[b - 1];
  }
  block_add_counts<kCheckNT>(bad, desc, &cnt[0], &cnt[2]);
}

__global__ void k_check_zero(unsigned long long* __restrict__ cnt) {
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0ull;
}

static unsigned check_blocks(int64_t n) {
  return (unsigned)std::min<int64_t>(std::max<int64_t>((n + kCheckNT - 1) / kCheckNT, 1), kCheckBlocks);
}

void dcsc_check_launch(const cbx_dcsc* X, unsigned long long* cnt, hipStream_t s) {
  dcsc_check_columns_launch(X, cnt, s);
  if (X->nnz > 0) k_check_entries<<<check_blocks(X->nnz), kCheckNT, 0, s>>>(X->ir, X->nnz, X->m, cnt);
}

void dcsc_check_columns_launch(const cbx_dcsc* X, unsigned long long* cnt, hipStream_t s) {
  k_check_zero<<<1, 64, 0, s>>>(cnt);
  k_check_columns<<<check_blocks(X->nzc), kCheckNT, 0, s>>>(X->cp, X->jc, X->ir, X->nzc, X->nnz, X->n, cnt);
}

cbx_status dcsc_check_result(const unsigned long long* cnt, hipStream_t s) {
  unsigned long long h[3] = {0, 0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  CBX_HIP(e);
  return h[0] || h[1] != h[2] ? set_error(CBX_ERR_INVALID, "cbx_dcsc_check: malformed DCSC") : CBX_OK;
}

}  // namespace cbx

using namespace cbx;

extern "C" cbx_status cbx_dcsc_column_checksums(const cbx_dcsc* X, uint64_t salt, uint64_t* host_out) {
  CBX_RETURN_IF(!X || !host_out, CBX_ERR_INVALID, "cbx_dcsc_column_checksums: null argument");
  CBX_RETURN_IF(X->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_dcsc_column_checksums: expected device memory");
  if (X->nzc == 0) return CBX_OK;
  uint64_t* d = nullptr;
  CBX_HIP(hipMalloc(&d, sizeof(uint64_t) * X->nzc));
  const dim3 g = wave_grid(X->nzc);
  switch (X->dtype) {
    case CBX_I64: k_col_checksum<int64_t><<<g, 256>>>(X->cp, X->ir, (const int64_t*)X->val, X->nzc, salt, d); break;
    case CBX_F64: k_col_checksum<double><<<g, 256>>>(X->cp, X->ir, (const double*)X->val, X->nzc, salt, d); break;
    case CBX_F32: k_col_checksum<float><<<g, 256>>>(X->cp, X->ir, (const float*)X->val, X->nzc, salt, d); break;
    case CBX_BOOL: k_col_checksum<uint8_t><<<g, 256>>>(X->cp, X->ir, (const uint8_t*)X->val, X->nzc, salt, d); break;
    default: (void)hipFree(d); return set_error(CBX_ERR_INVALID, "cbx_dcsc_column_checksums: unknown dtype");
  }
  hipError_t e = hipMemcpy(host_out, d, sizeof(uint64_t) * X->nzc, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  CBX_HIP(e);
  return CBX_OK;
}

extern "C" cbx_status cbx_dcsc_check(const cbx_dcsc* X) {
  CBX_RETURN_IF(!X, CBX_ERR_INVALID, "cbx_dcsc_check: null argument");
  CBX_RETURN_IF(X->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_dcsc_check: expected device memory");
  if (X->nzc == 0) return X->nnz == 0 ? CBX_OK : set_error(CBX_ERR_INVALID, "cbx_dcsc_check: nnz without columns");
  const cbx_status init = init_device_once();
  if (init != CBX_OK) return init;
  // stream-ordered scratch from the pool (a plain hipMalloc/hipFree pair costs
  // more than the check itself on a matrix of a few million entries)
  const hipStream_t s0 = nullptr;
  unsigned long long* d = nullptr;
  CBX_HIP(hipMallocAsync((void**)&d, 3 * sizeof(unsigned long long), s0));
  dcsc_check_launch(X, d, s0);
  const cbx_status st = dcsc_check_result(d, s0);
  (void)hipFreeAsync(d, s0);
  return st;
}

extern "C" cbx_status cbx_dcsc_structure_to_host(const cbx_dcsc* X, int64_t* cp_host, int64_t* jc_host) {
  CBX_RETURN_IF(!X, CBX_ERR_INVALID, "cbx_dcsc_structure_to_host: null argument");
  CBX_RETURN_IF(X->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_dcsc_structure_to_host: expected device memory");
  if (cp_host && X->cp) CBX_HIP(hipMemcpy(cp_host, X->cp, sizeof(int64_t) * (X->nzc + 1), hipMemcpyDeviceToHost));
  if (jc_host && X->nzc) CBX_HIP(hipMemcpy(jc_host, X->jc, sizeof(int64_t) * X->nzc, hipMemcpyDeviceToHost));
  return CBX_OK;
}
