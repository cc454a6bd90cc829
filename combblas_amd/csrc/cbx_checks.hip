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
// a column and below m, jc strictly increasing and below n, cp consistent).
__global__ void k_check_dcsc(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc,
                             const int64_t* __restrict__ ir, int64_t nzc, int64_t nnz, int64_t nrows, int64_t ncols,
                             int* __restrict__ bad) {
  const int64_t col = global_wave();
  if (col >= nzc) return;
  const int lane = lane_id();
  const int64_t b = cp[col], e = cp[col + 1];
  bool err = false;
  if (lane == 0) {
    err |= b < 0 || e <= b || e > nnz || jc[col] < 0 || jc[col] >= ncols;
    if (col + 1 < nzc) err |= jc[col + 1] <= jc[col];
    if (col == 0) err |= b != 0;
    if (col == nzc - 1) err |= e != nnz;
  }
  if (!(b < 0 || e < b || e > nnz))
    for (int64_t i = b + lane; i < e; i += kWave) {
      const int64_t r = ir[i];
      err |= r < 0 || r >= nrows;
      if (i + 1 < e) err |= ir[i + 1] <= r;
    }
  if (err) atomicOr(bad, 1);
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
  int* d = nullptr;
  CBX_HIP(hipMalloc(&d, sizeof(int)));
  CBX_HIP(hipMemset(d, 0, sizeof(int)));
  k_check_dcsc<<<wave_grid(X->nzc), 256>>>(X->cp, X->jc, X->ir, X->nzc, X->nnz, X->m, X->n, d);
  int bad = 0;
  hipError_t e = hipMemcpy(&bad, d, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  CBX_HIP(e);
  return bad ? set_error(CBX_ERR_INVALID, "cbx_dcsc_check: malformed DCSC") : CBX_OK;
}

extern "C" cbx_status cbx_dcsc_structure_to_host(const cbx_dcsc* X, int64_t* cp_host, int64_t* jc_host) {
  CBX_RETURN_IF(!X, CBX_ERR_INVALID, "cbx_dcsc_structure_to_host: null argument");
  CBX_RETURN_IF(X->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_dcsc_structure_to_host: expected device memory");
  if (cp_host && X->cp) CBX_HIP(hipMemcpy(cp_host, X->cp, sizeof(int64_t) * (X->nzc + 1), hipMemcpyDeviceToHost));
  if (jc_host && X->nzc) CBX_HIP(hipMemcpy(jc_host, X->jc, sizeof(int64_t) * X->nzc, hipMemcpyDeviceToHost));
  return CBX_OK;
}
