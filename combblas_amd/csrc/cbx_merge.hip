// Stage merge: the sum, under a semiring's add, of several sparse matrices of
// one shape, returned as DCSC.
//
// cbx_dcsc_merge replaces MultiwayMerge<SR> (MultiwayMerge.h:428-541; the
// per-split SerialMerge at :193-250) and MultiwayMergeHash with sorted=true
// (:553), which SUMMA uses to combine the local products of its stages, and
// the SpDCCols(SpTuples) conversion after it (SpDCCols.cpp:109-184): entries
// with equal (row, col) are combined with SR::add, nothing is dropped
// (explicit zeros stay), the result is column-major with ascending rows.  The
// reference merges k column-sorted lists with a heap; here the merge of c
// parts P_s (m x n) is the product [P_0 ... P_{c-1}] * B of their column
// concatenation (m x cn) with B (cn x n) stacking c identities whose entries
// are the semiring's multiplicative one (1 for plus-times and select-max,
// -0.0 / 0 for min-plus: x + -0.0 = x keeps a -0.0 entry's sign), so C(:,j)
// combines the parts' columns j with the semiring's add: the SpGEMM's own
// tiers do the merge and write DCSC directly.  Equal entries are combined in
// the accumulation's order, the reference in heap order: this changes only
// floating-point rounding.
//
// cbx_dcsc_from_coo (the Graph500 edge-list constructor, further below) keys
// the entries, radix-sorts them and assembles with reduce-by-key.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cbx_device.h"
#include "cbx_semiring.h"
#include "cbx_spgemm.h"

namespace cbx {
namespace {

constexpr int kMergeNT = 256;

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

// The parts as column blocks of one matrix: cp shifted by the entries before
// the part, jc by the part's first column (its index times n).  Checks what
// the concatenation's own validation cannot see: column ids in [0, n) (one
// beyond would fall into the next part's block and merge into the wrong
// output column) and cp starting at 0 and closing at the part's nnz; *bad
// counts violations.
__global__ void k_cat_columns(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc, int64_t nzc,
                              int64_t nnz, int64_t n, int64_t entry_base, int64_t col_base,
                              int64_t* __restrict__ cpA, int64_t* __restrict__ jcA, unsigned long long* __restrict__ bad) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < nzc) {
    const int64_t b = cp[q], j = jc[q];
    if (j < 0 || j >= n || (q == 0 && b != 0) || (q + 1 == nzc && cp[q + 1] != nnz)) atomicAdd(bad, 1ull);
    cpA[q] = b + entry_base;
    jcA[q] = j + col_base;
  }
}

__global__ void k_cat_close(int64_t* __restrict__ cpA, int64_t nzc, int64_t nnz) {
  if (threadIdx.x == 0 && blockIdx.x == 0) cpA[nzc] = nnz;
}

// B (c n x n): column j holds c entries, at rows s n + j, each the one.
template <class T>
__global__ void k_stacked_identity(int64_t n, int c, T one, int64_t* __restrict__ cpB, int64_t* __restrict__ jcB,
                                   int64_t* __restrict__ irB, T* __restrict__ valB) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j > n) return;
  cpB[j] = j * c;
  if (j == n) return;
  jcB[j] = j;
  for (int p = 0; p < c; ++p) {
    irB[j * c + p] = p * n + j;
    valB[j * c + p] = one;
  }
}

template <int K, class T>
T semiring_one() {
  if constexpr (K == kMinPlus) {
    if constexpr (std::is_floating_point<T>::value) return (T)-0.0;
    else return (T)0;
  } else {
    return (T)1;
  }
}

template <int K, class T>
cbx_status merge_typed(int32_t count, const cbx_dcsc* parts, cbx_dcsc* out, hipStream_t s) {
  const int64_t m = parts[0].m, n = parts[0].n;
  int64_t nnz = 0, nzc = 0;
  for (int32_t i = 0; i < count; ++i) {
    nnz += parts[i].nnz;
    nzc += parts[i].nzc;
  }
  std::memset(out, 0, sizeof(*out));
  out->m = m;
  out->n = n;
  out->dtype = parts[0].dtype;
  out->location = CBX_DEVICE;
  if (nnz == 0) {  // an empty m x n matrix
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
  CBX_RETURN_IF((double)count * (double)n >= 9.0e18, CBX_ERR_LIMIT, "cbx_dcsc_merge: too many columns in the parts");
  cbx_dcsc A{}, B{};
  A.m = m;
  A.n = (int64_t)count * n;
  A.nnz = nnz;
  A.nzc = nzc;
  A.dtype = B.dtype = parts[0].dtype;
  A.location = B.location = CBX_DEVICE;
  B.m = A.n;
  B.n = n;
  B.nnz = (int64_t)count * n;
  B.nzc = n;
  ScratchSet ws(s);
  int64_t *cpA, *jcA, *irA, *cpB, *jcB, *irB;
  T *valA, *valB;
  unsigned long long* bad;
  if (!ws.take(&cpA, nzc + 1) || !ws.take(&jcA, nzc) || !ws.take(&irA, nnz) || !ws.take(&valA, nnz) ||
      !ws.take(&cpB, n + 1) || !ws.take(&jcB, n) || !ws.take(&irB, B.nnz) || !ws.take(&valB, B.nnz) ||
      !ws.take(&bad, 1))
    return set_error(CBX_ERR_NOMEM, "cbx_dcsc_merge: workspace allocation failed");
  CBX_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned long long), s));
  A.cp = cpA;
  A.jc = jcA;
  A.ir = irA;
  A.val = valA;
  B.cp = cpB;
  B.jc = jcB;
  B.ir = irB;
  B.val = valB;
  int64_t eb = 0, qb = 0;
  for (int32_t i = 0; i < count; ++i) {
    const cbx_dcsc& P = parts[i];
    if (P.nzc) {
      k_cat_columns<<<(unsigned)((P.nzc + kMergeNT - 1) / kMergeNT), kMergeNT, 0, s>>>(
          P.cp, P.jc, P.nzc, P.nnz, n, eb, (int64_t)i * n, cpA + qb, jcA + qb, bad);
      CBX_HIP(hipMemcpyAsync(irA + eb, P.ir, sizeof(int64_t) * P.nnz, hipMemcpyDeviceToDevice, s));
      CBX_HIP(hipMemcpyAsync(valA + eb, P.val, sizeof(T) * P.nnz, hipMemcpyDeviceToDevice, s));
    }
    eb += P.nnz;
    qb += P.nzc;
  }
  k_cat_close<<<1, 64, 0, s>>>(cpA, nzc, nnz);
  k_stacked_identity<T><<<(unsigned)((n + 1 + kMergeNT - 1) / kMergeNT), kMergeNT, 0, s>>>(
      n, count, semiring_one<K, T>(), cpB, jcB, irB, valB);
  CBX_HIP(hipGetLastError());
  unsigned long long hbad = 0;
  CBX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  if (hbad) return set_error(CBX_ERR_INVALID, "cbx_dcsc_check: a part is malformed (column outside [0, n) or cp)");
  // the rest of the parts' structure is validated as A's (rows in range and
  // ascending per column, columns ascending, cp consistent); B is built here
  SpgemmConfig cfg = current_config();
  cfg.trusted_b = true;
  const cbx_semiring sr = K == kPlusTimes ? CBX_SR_PLUS_TIMES : K == kMinPlus ? CBX_SR_MIN_PLUS : CBX_SR_SELECT_MAX;
  const cbx_status st = spgemm_dispatch(&A, &B, sr, CBX_SPGEMM_TWO_PHASE, cfg, out, nullptr, s);
  if (st != CBX_OK) cbx_dcsc_free(out);
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
