// Stage merge: the sum, under a semiring's add, of several sparse matrices of
// one shape, returned as DCSC.
//
// cbx_dcsc_merge replaces MultiwayMerge<SR> (MultiwayMerge.h:428-541; the
// per-split SerialMerge at :193-250) and MultiwayMergeHash with sorted=true
// (:553), which SUMMA uses to combine the local products of its stages, and
// the SpDCCols(SpTuples) conversion after it (SpDCCols.cpp:109-184): entries
// with equal (row, col) are combined with SR::add, nothing is dropped
// (explicit zeros stay), the result is column-major with ascending rows.  The
// reference merges k column-sorted lists with a heap; here every entry gets a
// 64-bit (col, row) key (each part's keys are already in order), a tree of
// pairwise merges orders all of them in ceil(log2 count) passes, a
// reduce-by-key applies the semiring's add to equal keys, and a run-length
// encoding of the key's column half gives the DCSC columns.  Equal keys are
// combined in part order (the merges are stable), the reference in heap
// order: this changes only floating-point rounding.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_merge.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cbx_device.h"
#include "cbx_semiring.h"
#include "cbx_spgemm.h"

namespace cbx {
namespace {

constexpr int kMergeNT = 256;

// keys[out + e] = (col << 32) | row for every entry of one part, values copied;
// one wave per stored column
template <class T>
__global__ void k_merge_keys(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc,
                             const int64_t* __restrict__ ir, const T* __restrict__ val, int64_t nzc, int64_t out,
                             uint64_t* __restrict__ keys, T* __restrict__ vals) {
  const int64_t q = global_wave();
  if (q >= nzc) return;
  const uint64_t hi = (uint64_t)jc[q] << 32;
  for (int64_t e = cp[q] + lane_id(); e < cp[q + 1]; e += kWave) {
    keys[out + e] = hi | (uint64_t)ir[e];
    vals[out + e] = val[e];
  }
}

// The semiring's add on stored values (SR::add of the reference's semirings).
template <int K, class T>
struct SemAdd {
  __device__ __forceinline__ T operator()(const T& a, const T& b) const {
    using SR = Sem<K, T>;
    using Acc = typename SR::Acc;
    return (T)SR::out(SR::add((Acc)a, (Acc)b));
  }
};

struct KeyCol {
  __host__ __device__ __forceinline__ int64_t operator()(const uint64_t& k) const { return (int64_t)(k >> 32); }
};

__global__ void k_merge_rows(const uint64_t* __restrict__ keys, int64_t n, int64_t* __restrict__ ir) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) ir[i] = (int64_t)(keys[i] & 0xffffffffull);
}

__global__ void k_merge_close(int64_t* __restrict__ cp, int64_t nzc, int64_t nnz) {
  if (threadIdx.x == 0 && blockIdx.x == 0) cp[nzc] = nnz;  // after the exclusive scan of the column counts
}

using ColIt = hipcub::TransformInputIterator<int64_t, KeyCol, const uint64_t*>;

// Temporary bytes of reduce_assemble for up to N sorted entries.
template <int K, class T>
hipError_t assemble_tmp_bytes(int N, hipStream_t s, size_t& bytes) {
  size_t t2 = 0, t3 = 0, t4 = 0;
  uint64_t* k = nullptr;
  T* v = nullptr;
  int64_t* c = nullptr;
  hipError_t e = hipcub::DeviceReduce::ReduceByKey(nullptr, t2, k, k, v, v, c, SemAdd<K, T>(), N, s);
  if (e == hipSuccess) e = hipcub::DeviceRunLengthEncode::Encode(nullptr, t3, ColIt(k, KeyCol()), c, c, c, N, s);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, t4, c, c, N, s);
  bytes = std::max(t2, std::max(t3, t4));
  return e;
}

// From N (col, row)-sorted keys and values to a new device DCSC in *out:
// equal keys combined with the semiring's add (into ukeys/uvals, N entries
// each), columns from the run lengths of the keys' column halves (counts: N
// entries; nruns: 2).  out's shape and dtype are set by the caller.
template <int K, class T>
cbx_status reduce_assemble(const uint64_t* keys, const T* vals, int N, uint64_t* ukeys, T* uvals, int64_t* nruns,
                           int64_t* counts, void* tmp, size_t tb, cbx_dcsc* out, hipStream_t s, const char* who) {
  hipError_t e = hipSuccess;
  int64_t nnz = 0;  // no entries (every one a dropped loop): an empty matrix
  if (N > 0) {
    e = hipcub::DeviceReduce::ReduceByKey(tmp, tb, keys, ukeys, vals, uvals, nruns, SemAdd<K, T>(), N, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&nnz, nruns, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) return set_error(CBX_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
  cbx_status st = CBX_OK;
  if ((st = dev_alloc((void**)&out->jc, sizeof(int64_t) * std::max<int64_t>(nnz, 1))) != CBX_OK ||
      (st = dev_alloc((void**)&out->ir, sizeof(int64_t) * std::max<int64_t>(nnz, 1))) != CBX_OK ||
      (st = dev_alloc(&out->val, sizeof(T) * std::max<int64_t>(nnz, 1))) != CBX_OK) {
    cbx_dcsc_free(out);
    return st;
  }
  int64_t nzc = 0;
  if (nnz > 0) {
    e = hipcub::DeviceRunLengthEncode::Encode(tmp, tb, ColIt(ukeys, KeyCol()), out->jc, counts, nruns + 1, (int)nnz,
                                              s);
    if (e == hipSuccess) e = hipMemcpyAsync(&nzc, nruns + 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e == hipSuccess && (st = dev_alloc((void**)&out->cp, sizeof(int64_t) * (nzc + 1))) != CBX_OK) {
    cbx_dcsc_free(out);
    return st;
  }
  if (e == hipSuccess && nzc > 0) e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, counts, out->cp, (int)nzc, s);
  if (e == hipSuccess) {
    k_merge_close<<<1, 64, 0, s>>>(out->cp, nzc, nnz);
    if (nnz > 0) {
      k_merge_rows<<<(unsigned)((nnz + kMergeNT - 1) / kMergeNT), kMergeNT, 0, s>>>(ukeys, nnz, out->ir);
      e = hipMemcpyAsync(out->val, uvals, sizeof(T) * nnz, hipMemcpyDeviceToDevice, s);
    }
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    cbx_dcsc_free(out);
    return set_error(CBX_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
  }
  out->nnz = nnz;
  out->nzc = nzc;
  return CBX_OK;
}

template <int K, class T>
cbx_status merge_typed(int32_t count, const cbx_dcsc* parts, cbx_dcsc* out, hipStream_t s) {
  const int64_t m = parts[0].m, n = parts[0].n;
  int64_t total = 0;
  std::vector<int64_t> base(count);
  for (int32_t i = 0; i < count; ++i) {
    base[i] = total;
    total += parts[i].nnz;
  }
  std::memset(out, 0, sizeof(*out));
  out->m = m;
  out->n = n;
  out->dtype = parts[0].dtype;
  out->location = CBX_DEVICE;
  CBX_RETURN_IF(total >= (1ll << 31), CBX_ERR_LIMIT, "cbx_dcsc_merge: more than 2^31 entries in one merge");
  if (total == 0) {  // an empty m x n matrix
    cbx_status st = CBX_OK;
    if ((st = dev_alloc((void**)&out->cp, sizeof(int64_t))) != CBX_OK ||
        (st = dev_alloc((void**)&out->jc, sizeof(int64_t))) != CBX_OK ||
        (st = dev_alloc((void**)&out->ir, sizeof(int64_t))) != CBX_OK ||
        (st = dev_alloc(&out->val, sizeof(int64_t))) != CBX_OK) {
      cbx_dcsc_free(out);
      return st;
    }
    CBX_HIP(hipMemsetAsync(out->cp, 0, sizeof(int64_t), s));
    CBX_HIP(hipStreamSynchronize(s));
    return CBX_OK;
  }
  const int N = (int)total;
  uint64_t *k0 = nullptr, *k1 = nullptr, *ukeys = nullptr;
  T *v0 = nullptr, *v1 = nullptr;
  int64_t *nruns = nullptr, *counts = nullptr;
  void* tmp = nullptr;
  size_t tb = 0, t1 = 0, t2 = 0;
  CBX_HIP(rocprim::merge(nullptr, t1, k0, k0, k1, v0, v0, v1, (size_t)N, (size_t)N, rocprim::less<uint64_t>(), s));
  CBX_HIP((assemble_tmp_bytes<K, T>(N, s, t2)));
  tb = std::max(t1, t2);
  ScratchSet ws(s);
  auto cleanup = [&]() { ws.release(); };
  if (!ws.take(&k0, N) || !ws.take(&k1, N) || !ws.take(&ukeys, N) || !ws.take(&v0, N) || !ws.take(&v1, N) ||
      !ws.take(&nruns, 2) || !ws.take(&counts, N) || !ws.take_bytes(&tmp, std::max<size_t>(tb, 16))) {
    (void)hipGetLastError();
    cleanup();
    return set_error(CBX_ERR_NOMEM, "cbx_dcsc_merge: workspace allocation failed");
  }
  for (int32_t i = 0; i < count; ++i)
    if (parts[i].nzc)
      k_merge_keys<T><<<wave_grid(parts[i].nzc), 256, 0, s>>>(parts[i].cp, parts[i].jc, parts[i].ir,
                                                                (const T*)parts[i].val, parts[i].nzc, base[i], k0,
                                                                v0);
  hipError_t e = hipGetLastError();
  // pairwise merges of adjacent sorted runs until one remains (in k1/v1)
  std::vector<int64_t> runs(base);
  runs.push_back(total);  // run r is [runs[r], runs[r+1])
  uint64_t *ksrc = k0, *kdst = k1;
  T *vsrc = v0, *vdst = v1;
  while (e == hipSuccess && runs.size() > 2) {
    std::vector<int64_t> next;
    for (size_t r = 0; r + 1 < runs.size() && e == hipSuccess; r += 2) {
      next.push_back(runs[r]);
      const int64_t a = runs[r], b = runs[r + 1], c = r + 2 < runs.size() ? runs[r + 2] : b;
      if (c > b)
        e = rocprim::merge(tmp, tb, ksrc + a, ksrc + b, kdst + a, vsrc + a, vsrc + b, vdst + a, (size_t)(b - a),
                           (size_t)(c - b), rocprim::less<uint64_t>(), s);
      else if (b > a)  // an odd run out: copied through
        e = hipMemcpyAsync(kdst + a, ksrc + a, sizeof(uint64_t) * (b - a), hipMemcpyDeviceToDevice, s) ||
                    hipMemcpyAsync(vdst + a, vsrc + a, sizeof(T) * (b - a), hipMemcpyDeviceToDevice, s)
                ? hipErrorUnknown
                : hipSuccess;
    }
    next.push_back(total);
    runs.swap(next);
    std::swap(ksrc, kdst);
    std::swap(vsrc, vdst);
  }
  if (e == hipSuccess && ksrc != k1) {  // one run: bring it to k1/v1
    e = hipMemcpyAsync(k1, ksrc, sizeof(uint64_t) * N, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(v1, vsrc, sizeof(T) * N, hipMemcpyDeviceToDevice, s);
  }
  if (e != hipSuccess) {
    cleanup();
    return set_error(CBX_ERR_HIP, "cbx_dcsc_merge: %s", hipGetErrorString(e));
  }
  // distinct (col, row) keys and their combined values (values reuse v0)
  const cbx_status st = reduce_assemble<K, T>(k1, v1, N, ukeys, v0, nruns, counts, tmp, tb, out, s, "cbx_dcsc_merge");
  cleanup();
  return st;
}

// Edge-list input: keys of COO entries.  bad[0] counts entries out of range,
// bad[1] the self loops dropped (their key is the sentinel (n << 32), past
// every valid key, so they sort to the end).  vals may be null (value 1).
template <class T>
__global__ __launch_bounds__(kMergeNT) void k_coo_keys(const int64_t* __restrict__ rows,
                                                       const int64_t* __restrict__ cols,
                                                       const T* __restrict__ vals, int64_t nnz, int64_t m, int64_t n,
                                                       int drop_loops, uint64_t* __restrict__ keys,
                                                       T* __restrict__ out, unsigned long long* __restrict__ bad) {
  long long oob = 0, loops = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += stride) {
    const int64_t r = rows[i], c = cols[i];
    const bool ok = r >= 0 && r < m && c >= 0 && c < n;
    const bool loop = drop_loops && r == c;
    oob += !ok;
    loops += ok && loop;
    keys[i] = ok && !loop ? ((uint64_t)c << 32) | (uint64_t)r : ((uint64_t)n << 32);
    out[i] = vals ? vals[i] : (T)1;
  }
  block_add_counts<kMergeNT>(oob, loops, &bad[0], &bad[1]);
}

int bits_for(int64_t v) {  // bits to hold 0..v
  int b = 1;
  while (b < 62 && (int64_t(1) << b) <= v) ++b;
  return b;
}

template <int K, class T>
cbx_status coo_typed(int64_t m, int64_t n, const int64_t* rows, const int64_t* cols, const T* vals, int64_t nnz,
                     int drop_loops, cbx_dcsc* out, hipStream_t s) {
  const int N = (int)nnz;
  uint64_t *k0 = nullptr, *k1 = nullptr, *ukeys = nullptr;
  T *v0 = nullptr, *v1 = nullptr;
  int64_t *nruns = nullptr, *counts = nullptr;
  unsigned long long* bad = nullptr;
  void* tmp = nullptr;
  size_t tb = 0, t1 = 0, t2 = 0;
  const int end_bit = 32 + bits_for(n);  // the loop sentinel (col = n) included
  CBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, k0, k1, v0, v1, N, 0, end_bit, s));
  CBX_HIP((assemble_tmp_bytes<K, T>(N, s, t2)));
  tb = std::max(t1, t2);
  ScratchSet ws(s);
  auto cleanup = [&]() { ws.release(); };
  if (!ws.take(&k0, N) || !ws.take(&k1, N) || !ws.take(&ukeys, N) || !ws.take(&v0, N) || !ws.take(&v1, N) ||
      !ws.take(&nruns, 2) || !ws.take(&counts, N) || !ws.take(&bad, 2) ||
      !ws.take_bytes(&tmp, std::max<size_t>(tb, 16))) {
    (void)hipGetLastError();
    cleanup();
    return set_error(CBX_ERR_NOMEM, "cbx_dcsc_from_coo: workspace allocation failed");
  }
  unsigned long long hbad[2] = {0, 0};
  hipError_t e = hipMemsetAsync(bad, 0, 2 * sizeof(unsigned long long), s);
  if (e == hipSuccess) {
    const unsigned g = (unsigned)std::min<int64_t>((nnz + kMergeNT - 1) / kMergeNT, 4096);
    k_coo_keys<T><<<g, kMergeNT, 0, s>>>(rows, cols, vals, nnz, m, n, drop_loops, k0, v0, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess || hbad[0]) {
    cleanup();
    return e != hipSuccess ? set_error(CBX_ERR_HIP, "cbx_dcsc_from_coo: %s", hipGetErrorString(e))
                           : set_error(CBX_ERR_INVALID, "cbx_dcsc_from_coo: %llu entries outside the %lld x %lld shape",
                                       hbad[0], (long long)m, (long long)n);
  }
  e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, N, 0, end_bit, s);
  if (e != hipSuccess) {
    cleanup();
    return set_error(CBX_ERR_HIP, "cbx_dcsc_from_coo: %s", hipGetErrorString(e));
  }
  // the dropped loops sit at the end of the sorted keys
  const cbx_status st = reduce_assemble<K, T>(k1, v1, N - (int)hbad[1], ukeys, v0, nruns, counts, tmp, tb, out, s,
                                              "cbx_dcsc_from_coo");
  cleanup();
  return st;
}

}  // namespace
}  // namespace cbx

using namespace cbx;

extern "C" cbx_status cbx_dcsc_merge(int32_t count, const cbx_dcsc* parts, cbx_semiring semiring, cbx_dcsc* out,
                                     void* stream) {
  CBX_RETURN_IF(!parts || !out || count < 1, CBX_ERR_INVALID, "cbx_dcsc_merge: need at least one part and an output");
  cbx_status st = init_device_once();
  if (st != CBX_OK) return st;
  for (int32_t i = 0; i < count; ++i) {
    if ((st = check_dcsc(&parts[i], "part", CBX_DEVICE)) != CBX_OK) return st;
    CBX_RETURN_IF(parts[i].m != parts[0].m || parts[i].n != parts[0].n || parts[i].dtype != parts[0].dtype,
                  CBX_ERR_INVALID, "cbx_dcsc_merge: part %d differs in shape or dtype from part 0", i);
    if ((st = cbx_dcsc_check(&parts[i])) != CBX_OK) return st;  // keys index by the parts' rows and columns
  }
  hipStream_t s = (hipStream_t)stream;
  const cbx_dtype dt = parts[0].dtype;
  switch (semiring * 8 + dt) {
#define CBX_MG(SRK, SRE, T, DT) \
  case SRE * 8 + DT:            \
    return merge_typed<SRK, T>(count, parts, out, s);
    CBX_MG(kPlusTimes, CBX_SR_PLUS_TIMES, int64_t, CBX_I64)
    CBX_MG(kPlusTimes, CBX_SR_PLUS_TIMES, double, CBX_F64)
    CBX_MG(kPlusTimes, CBX_SR_PLUS_TIMES, float, CBX_F32)
    CBX_MG(kPlusTimes, CBX_SR_PLUS_TIMES, uint8_t, CBX_BOOL)
    CBX_MG(kMinPlus, CBX_SR_MIN_PLUS, int64_t, CBX_I64)
    CBX_MG(kMinPlus, CBX_SR_MIN_PLUS, double, CBX_F64)
    CBX_MG(kMinPlus, CBX_SR_MIN_PLUS, float, CBX_F32)
    CBX_MG(kSelectMax, CBX_SR_SELECT_MAX, int64_t, CBX_I64)
    CBX_MG(kSelectMax, CBX_SR_SELECT_MAX, double, CBX_F64)
    CBX_MG(kSelectMax, CBX_SR_SELECT_MAX, float, CBX_F32)
#undef CBX_MG
  }
  return set_error(CBX_ERR_UNSUPPORTED, "cbx_dcsc_merge: semiring %d with dtype %d is not provided", (int)semiring,
                   (int)dt);
}

extern "C" cbx_status cbx_dcsc_from_coo(int64_t m, int64_t n, const int64_t* rows, const int64_t* cols,
                                        const void* vals, int64_t nnz, cbx_dtype dtype, cbx_semiring semiring,
                                        int32_t remove_loops, cbx_dcsc* out, void* stream) {
  CBX_RETURN_IF(!out || (nnz > 0 && (!rows || !cols)), CBX_ERR_INVALID, "cbx_dcsc_from_coo: null argument");
  CBX_RETURN_IF(m < 0 || n < 0 || nnz < 0, CBX_ERR_INVALID, "cbx_dcsc_from_coo: negative size");
  CBX_RETURN_IF(m >= (1ll << 31) || n >= (1ll << 31), CBX_ERR_LIMIT, "cbx_dcsc_from_coo: dimensions must be below 2^31");
  CBX_RETURN_IF(nnz >= (1ll << 31), CBX_ERR_LIMIT, "cbx_dcsc_from_coo: at most 2^31 - 1 entries");
  cbx_status st = init_device_once();
  if (st != CBX_OK) return st;
  std::memset(out, 0, sizeof(*out));
  out->m = m;
  out->n = n;
  out->dtype = dtype;
  out->location = CBX_DEVICE;
  hipStream_t s = (hipStream_t)stream;
  if (nnz == 0) {  // an empty m x n matrix
    if ((st = dev_alloc((void**)&out->cp, sizeof(int64_t))) != CBX_OK ||
        (st = dev_alloc((void**)&out->jc, sizeof(int64_t))) != CBX_OK ||
        (st = dev_alloc((void**)&out->ir, sizeof(int64_t))) != CBX_OK ||
        (st = dev_alloc(&out->val, sizeof(int64_t))) != CBX_OK) {
      cbx_dcsc_free(out);
      return st;
    }
    CBX_HIP(hipMemsetAsync(out->cp, 0, sizeof(int64_t), s));
    CBX_HIP(hipStreamSynchronize(s));
    return CBX_OK;
  }
  const int loops = remove_loops ? 1 : 0;
  switch (semiring * 8 + dtype) {
#define CBX_CO(SRK, SRE, T, DT) \
  case SRE * 8 + DT:            \
    return coo_typed<SRK, T>(m, n, rows, cols, (const T*)vals, nnz, loops, out, s);
    CBX_CO(kPlusTimes, CBX_SR_PLUS_TIMES, int64_t, CBX_I64)
    CBX_CO(kPlusTimes, CBX_SR_PLUS_TIMES, double, CBX_F64)
    CBX_CO(kPlusTimes, CBX_SR_PLUS_TIMES, float, CBX_F32)
    CBX_CO(kPlusTimes, CBX_SR_PLUS_TIMES, uint8_t, CBX_BOOL)
    CBX_CO(kMinPlus, CBX_SR_MIN_PLUS, int64_t, CBX_I64)
    CBX_CO(kMinPlus, CBX_SR_MIN_PLUS, double, CBX_F64)
    CBX_CO(kMinPlus, CBX_SR_MIN_PLUS, float, CBX_F32)
    CBX_CO(kSelectMax, CBX_SR_SELECT_MAX, int64_t, CBX_I64)
    CBX_CO(kSelectMax, CBX_SR_SELECT_MAX, double, CBX_F64)
    CBX_CO(kSelectMax, CBX_SR_SELECT_MAX, float, CBX_F32)
#undef CBX_CO
  }
  return set_error(CBX_ERR_UNSUPPORTED, "cbx_dcsc_from_coo: semiring %d with dtype %d is not provided", (int)semiring,
                   (int)dtype);
}
