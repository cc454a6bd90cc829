// MCL expansion on gfx950: the prune / recover / select epilogue and the
// column-phased product that feeds it.
//
// Replaces, for one process's block, MCLPruneRecoverySelect
// (include/CombBLAS/ParFriends.h:185-345) and the phase loop of
// MemEfficientSpGEMM (ParFriends.h:452-780, B split with ColSplit, each phase
// multiplied, pruned, then ColConcatenate).  Per stored column j of C:
//   nP, sP   count and sum of the entries > hard threshold (the Reduce calls
//            over PrunedA, ParFriends.h:196-200)
//   recover  nP < recoverNum, nnz > nP and sP < recoverPct: the cut is the
//            recoverNum-th largest entry of the unpruned column (Kselect)
//   select   otherwise, if nP > selectNum: the cut is the selectNum-th
//            largest entry; if then fewer than recoverNum entries are >= the
//            cut and their sum is < recoverPct, the cut is the
//            recoverNum-th largest instead (ParFriends.h:270-325)
//   keep     entries with !(v < cut), the cut defaulting to the threshold
//            (PruneColumn with std::less, ParFriends.h:339)
// Kselect's k-th largest of a column with fewer than k entries is its
// smallest entry (SpParMat.cpp:1680-1687).
//
// Kernels:
//   k_mcl_stats   one wave per column: nP, sP (FP64), the count kept at the
//                 threshold, and the column's job (none / recover / select),
//                 appended to a job list
//   k_mcl_select  one workgroup per job column: MSB-first radix select on
//                 order-preserving integer keys (8-bit digits, LDS histogram),
//                 the recover-after-select test and the final kept count
//   k_mcl_write   one wave per column: order-preserving ballot compaction of
//                 the kept entries into the output DCSC
// The reference sums column statistics sequentially in the value type; here
// they are summed in FP64, so a recovery decision can differ only for a
// column whose sum lies within rounding of recoverPct.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cbx_device.h"
#include "cbx_spgemm.h"

namespace cbx {
namespace {

template <class T>
struct OrderKey;
template <>
struct OrderKey<float> {
  using U = uint32_t;
  static constexpr int kBits = 32;
  __device__ static U enc(float v) {
    const U b = __float_as_uint(v);
    return (b >> 31) ? ~b : (b | 0x80000000u);
  }
  __device__ static float dec(U k) { return __uint_as_float((k >> 31) ? (k & 0x7fffffffu) : ~k); }
};
template <>
struct OrderKey<double> {
  using U = uint64_t;
  static constexpr int kBits = 64;
  __device__ static U enc(double v) {
    const U b = (U)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
  }
  __device__ static double dec(U k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
  }
};

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

constexpr int kSelNT = 256;

// job codes
constexpr int kJobNone = 0, kJobRecover = 1, kJobSelect = 2;

template <class T>
__global__ __launch_bounds__(256) void k_mcl_stats(const int64_t* __restrict__ cp, const T* __restrict__ val,
                                                   int64_t nzc, T thr, int64_t select, int64_t recover, double pct,
                                                   T* __restrict__ cut, int64_t* __restrict__ cnt,
                                                   int32_t* __restrict__ job, int32_t* __restrict__ jobs,
                                                   int32_t* __restrict__ npos, unsigned long long* __restrict__ counters) {
  const int64_t q = global_wave();
  if (q >= nzc) return;  // whole wave
  const int lane = lane_id();
  const int64_t b = cp[q], e = cp[q + 1];
  long long nP = 0, nK = 0;
  double sP = 0;
  // kUnroll independent loads in flight per lane (columns are long and the
  // wave's loads would otherwise wait on each other)
  constexpr int kUnroll = 8;
  for (int64_t p0 = b; p0 < e; p0 += kWave * kUnroll) {
    T v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t p = p0 + u * kWave + lane;
      v[u] = p < e ? val[p] : thr;  // thr: counted as kept, not as above the threshold
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (p0 + u * kWave + lane >= e) continue;
      if (!(v[u] <= thr)) {
        ++nP;
        sP += (double)v[u];
      }
      nK += !(v[u] < thr);
    }
  }
  nP = wave_sum_ll(nP);
  nK = wave_sum_ll(nK);
  sP = wave_sum_f64(sP);
  if (lane == 0) {
    const long long N = e - b;
    const bool rec = nP < recover && N > nP && sP < pct;
    const bool sel = !rec && select > 0 && nP > select;
    const int j = rec ? kJobRecover : (sel ? kJobSelect : kJobNone);
    job[q] = j;
    npos[q] = (int32_t)nP;
    cut[q] = thr;
    cnt[q] = nK;
    if (j != kJobNone) {
      const unsigned long long slot = atomicAdd(&counters[0], 1ull);
      jobs[slot] = (int32_t)q;
      atomicAdd(&counters[j], 1ull);
    }
  }
}

// k-th largest (1-based, k <= N) of v[0..N) by MSB-first radix select over
// order-preserving keys.  All NT threads of the block must call it.
template <class T, int NT>
__device__ T select_kth(const T* __restrict__ v, int64_t N, int64_t k, unsigned* hist,
                        typename OrderKey<T>::U* sh_prefix, unsigned long long* sh_rank) {
  using K = OrderKey<T>;
  using U = typename K::U;
  U prefix = 0, mask = 0;
  unsigned long long rank = (unsigned long long)k;
  for (int shift = K::kBits - 8; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 256; i += NT) hist[i] = 0;
    __syncthreads();
    for (int64_t p = threadIdx.x; p < N; p += NT) {
      const U key = K::enc(v[p]);
      if ((key & mask) == prefix) atomicAdd(&hist[(unsigned)(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
      // lane l holds bins 255-4l .. 252-4l, so lanes run from the largest digit down
      const int lane = (int)threadIdx.x;
      const unsigned long long h[4] = {hist[255 - 4 * lane], hist[254 - 4 * lane], hist[253 - 4 * lane],
                                       hist[252 - 4 * lane]};
      const unsigned long long tot = h[0] + h[1] + h[2] + h[3];
      const unsigned long long inc = (unsigned long long)wave_incl_scan_ll((long long)tot);
      const unsigned long long exc = inc - tot;
      if (exc < rank && rank <= inc) {
        unsigned long long r = rank - exc;
        int i = 0;
        while (i < 3 && r > h[i]) r -= h[i++];
        *sh_prefix = prefix | ((U)(255 - 4 * lane - i) << shift);
        *sh_rank = r;
      }
    }
    __syncthreads();
    prefix = *sh_prefix;
    rank = *sh_rank;
    mask |= (U)255 << shift;
  }
  return K::dec(prefix);
}

template <int NT>
__device__ double block_sum_f64(double v, double* wsum) {
  constexpr int NW = NT / kWave;
  v = wave_sum_f64(v);
  if (lane_id() == 0) wsum[wave_id()] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) t += wsum[i];
  __syncthreads();
  return t;
}

// A selected column's cut is its k-th largest entry with k <= nP, so it lies
// among the nP entries above the hard threshold; when those fit in LDS they
// are gathered once and every later step (selection, the recovery test,
// the kept count) runs on them instead of re-reading the column.
constexpr int kCandCap = 8192;

template <class T, int NT>
__global__ __launch_bounds__(NT) void k_mcl_select(const int64_t* __restrict__ cp, const T* __restrict__ val,
                                                   const int32_t* __restrict__ jobs, const int32_t* __restrict__ job,
                                                   const int32_t* __restrict__ npos, T thr, int64_t select,
                                                   int64_t recover, double pct, T* __restrict__ cut,
                                                   int64_t* __restrict__ cnt,
                                                   unsigned long long* __restrict__ counters) {
  __shared__ unsigned hist[256];
  __shared__ typename OrderKey<T>::U sh_prefix;
  __shared__ unsigned long long sh_rank;
  __shared__ long long wsum[NT / kWave];
  __shared__ double wsumd[NT / kWave];
  __shared__ T cand[kCandCap];
  __shared__ int ncand;
  const int32_t q = jobs[blockIdx.x];
  const int64_t b = cp[q], N = cp[q + 1] - b;
  const T* v = val + b;
  const int j = job[q];
  const int64_t nP = npos[q];
  int64_t k0 = j == kJobRecover ? recover : select;
  k0 = k0 < N ? k0 : N;
  const bool use_cand = nP <= kCandCap && k0 <= nP;  // block-uniform
  if (use_cand) {
    if (threadIdx.x == 0) ncand = 0;
    __syncthreads();
    const int lane = lane_id();
    for (int64_t p0 = 0; p0 < N; p0 += NT) {
      const int64_t p = p0 + threadIdx.x;
      T x = T(0);
      bool take = false;
      if (p < N) {
        x = v[p];
        take = !(x <= thr);
      }
      const unsigned long long m = __ballot(take);
      int base = 0;
      if (lane == 0 && m) base = atomicAdd(&ncand, __popcll(m));
      base = __shfl(base, 0, kWave);
      if (take) cand[base + __popcll(m & ((1ull << lane) - 1ull))] = x;
    }
    __syncthreads();
  }
  const int M = use_cand ? ncand : 0;
  T c = use_cand ? select_kth<T, NT>(cand, M, k0, hist, &sh_prefix, &sh_rank)
                 : select_kth<T, NT>(v, N, k0, hist, &sh_prefix, &sh_rank);
  // entries >= c all lie among the candidates when c is above the threshold
  auto src_for = [&](T cc, const T*& src, int64_t& len) {
    if (use_cand && cc > thr) {
      src = cand;
      len = M;
    } else {
      src = v;
      len = N;
    }
  };
  if (j == kJobSelect && recover > 0) {
    const T* src;
    int64_t len;
    src_for(c, src, len);
    long long n1 = 0;
    double s1 = 0;
    for (int64_t p = threadIdx.x; p < len; p += NT) {
      const T x = src[p];
      if (!(x < c)) {
        ++n1;
        s1 += (double)x;
      }
    }
    n1 = block_sum_ll<NT>(n1, wsum);
    s1 = block_sum_f64<NT>(s1, wsumd);
    if (n1 < recover && s1 < pct) {
      const int64_t kr = recover < N ? recover : N;
      c = (use_cand && kr <= nP) ? select_kth<T, NT>(cand, M, kr, hist, &sh_prefix, &sh_rank)
                                 : select_kth<T, NT>(v, N, kr, hist, &sh_prefix, &sh_rank);
      if (threadIdx.x == 0) atomicAdd(&counters[3], 1ull);
    }
  }
  const T* src;
  int64_t len;
  src_for(c, src, len);
  long long kept = 0;
  for (int64_t p = threadIdx.x; p < len; p += NT) kept += !(src[p] < c);
  kept = block_sum_ll<NT>(kept, wsum);
  if (threadIdx.x == 0) {
    cut[q] = c;
    cnt[q] = kept;
  }
}

__global__ void k_nonempty(const int64_t* __restrict__ cnt, int64_t nzc, int32_t* __restrict__ flag) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < nzc) flag[q] = cnt[q] > 0;
}

template <class T>
__global__ __launch_bounds__(256) void k_mcl_write(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc,
                                                   const int64_t* __restrict__ ir, const T* __restrict__ val,
                                                   int64_t nzc, const T* __restrict__ cut,
                                                   const int64_t* __restrict__ cnt, const int64_t* __restrict__ off,
                                                   const int32_t* __restrict__ cpos, int64_t* __restrict__ ocp,
                                                   int64_t* __restrict__ ojc, int64_t* __restrict__ oir,
                                                   T* __restrict__ oval) {
  const int64_t q = global_wave();
  if (q >= nzc || cnt[q] == 0) return;  // whole wave
  const int lane = lane_id();
  const T c = cut[q];
  int64_t w = off[q];
  const int64_t b = cp[q], e = cp[q + 1];
  constexpr int kUnroll = 8;  // loads in flight per lane
  const int64_t end_w = w + cnt[q];
  for (int64_t base0 = b; base0 < e && w < end_w; base0 += kWave * kUnroll) {
    T v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t p = base0 + u * kWave + lane;
      v[u] = p < e ? val[p] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t p = base0 + u * kWave + lane;
      const bool keep = p < e && !(v[u] < c);
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int64_t at = w + __popcll(m & ((1ull << lane) - 1ull));
        oir[at] = ir[p];
        oval[at] = v[u];
      }
      w += __popcll(m);
    }
  }
  if (lane == 0) {
    ojc[cpos[q]] = jc[q];
    ocp[cpos[q]] = off[q];
  }
}

// Column slice of a DCSC: dst cp[i] = src cp[q0+i] - src cp[q0].
__global__ void k_rebase(const int64_t* __restrict__ src, int64_t q0, int64_t count, int64_t* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[q0 + i] - src[q0];
}

// Concatenation: dst[i] = src[i] + shift.
__global__ void k_shift(const int64_t* __restrict__ src, int64_t count, int64_t shift, int64_t* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i] + shift;
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Scratch from the library's caching allocator, released at scope exit.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) dev_free(p);
  }
  template <class X>
  cbx_status get(X** out, int64_t count) {
    void* p = nullptr;
    cbx_status st = dev_alloc(&p, sizeof(X) * (size_t)std::max<int64_t>(count, 1));
    if (st) return st;
    ptrs.push_back(p);
    *out = (X*)p;
    return CBX_OK;
  }
};

#define CBX_TRY(x)                 \
  do {                             \
    cbx_status st_ = (x);          \
    if (st_ != CBX_OK) return st_; \
  } while (0)

cbx_status alloc_dcsc(cbx_dcsc* X, int64_t m, int64_t n, int64_t nnz, int64_t nzc, cbx_dtype dt) {
  std::memset(X, 0, sizeof(*X));
  X->m = m;
  X->n = n;
  X->nnz = nnz;
  X->nzc = nzc;
  X->dtype = dt;
  X->location = CBX_DEVICE;
  cbx_status st;
  if ((st = dev_alloc((void**)&X->cp, sizeof(int64_t) * (size_t)(nzc + 1))) ||
      (st = dev_alloc((void**)&X->jc, sizeof(int64_t) * (size_t)std::max<int64_t>(nzc, 1))) ||
      (st = dev_alloc((void**)&X->ir, sizeof(int64_t) * (size_t)std::max<int64_t>(nnz, 1))) ||
      (st = dev_alloc(&X->val, dtype_size(dt) * (size_t)std::max<int64_t>(nnz, 1)))) {
    cbx_dcsc_free(X);
    return st;
  }
  return CBX_OK;
}

}  // namespace

template <class T>
static cbx_status mcl_prune_typed(const cbx_dcsc* C, double thr_d, int64_t select, int64_t recover, double pct_d,
                                  cbx_dcsc* out, cbx_mcl_stats* st, hipStream_t s) {
  const int64_t nzc = C->nzc;
  const T thr = (T)thr_d;
  const double pct = (double)(T)pct_d;  // the reference compares in the value type
  Scratch sc;
  T* cut;
  int64_t *cnt, *off;
  int32_t *job, *jobs, *flag, *cpos, *npos;
  unsigned long long* counters;
  CBX_TRY(sc.get(&cut, nzc));
  CBX_TRY(sc.get(&cnt, nzc + 1));
  CBX_TRY(sc.get(&off, nzc + 1));
  CBX_TRY(sc.get(&job, nzc));
  CBX_TRY(sc.get(&jobs, nzc));
  CBX_TRY(sc.get(&npos, nzc));
  CBX_TRY(sc.get(&flag, nzc + 1));
  CBX_TRY(sc.get(&cpos, nzc + 1));
  CBX_TRY(sc.get(&counters, 4));
  CBX_HIP(hipMemsetAsync(counters, 0, sizeof(unsigned long long) * 4, s));
  CBX_HIP(hipMemsetAsync(cnt + nzc, 0, sizeof(int64_t), s));
  CBX_HIP(hipMemsetAsync(flag + nzc, 0, sizeof(int32_t), s));
  if (nzc) {
    k_mcl_stats<T><<<wave_grid(nzc), 256, 0, s>>>(C->cp, (const T*)C->val, nzc, thr, select, recover,
                                                              pct, cut, cnt, job, jobs, npos, counters);
    CBX_HIP(hipGetLastError());
  }
  unsigned long long hc[4] = {0, 0, 0, 0};
  CBX_HIP(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  CBX_RETURN_IF(hc[0] * (unsigned long long)kSelNT >= (1ull << 32), CBX_ERR_LIMIT,
                "cbx_mcl_prune: too many selected columns for a single launch");
  if (hc[0]) {
    k_mcl_select<T, kSelNT><<<(unsigned)hc[0], kSelNT, 0, s>>>(C->cp, (const T*)C->val, jobs, job, npos, thr,
                                                               select, recover, pct, cut, cnt, counters);
    CBX_HIP(hipGetLastError());
  }
  size_t tmp_bytes = 0, tb2 = 0;
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, off, (int)(nzc + 1), s));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, flag, cpos, (int)(nzc + 1), s));
  tmp_bytes = std::max(tmp_bytes, tb2);
  uint8_t* tmp;
  CBX_TRY(sc.get(&tmp, (int64_t)tmp_bytes));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, (int)(nzc + 1), s));
  if (nzc) k_nonempty<<<nblocks(nzc, 256), 256, 0, s>>>(cnt, nzc, flag);
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flag, cpos, (int)(nzc + 1), s));
  int64_t onnz = 0;
  int32_t onzc = 0;
  CBX_HIP(hipMemcpyAsync(&onnz, off + nzc, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&onzc, cpos + nzc, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&hc[3], counters + 3, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  CBX_TRY(alloc_dcsc(out, C->m, C->n, onnz, onzc, C->dtype));
  if (nzc)
    k_mcl_write<T><<<wave_grid(nzc), 256, 0, s>>>(C->cp, C->jc, C->ir, (const T*)C->val, nzc, cut, cnt,
                                                              off, cpos, out->cp, out->jc, out->ir, (T*)out->val);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipMemcpyAsync(out->cp + onzc, &onnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
  CBX_HIP(hipStreamSynchronize(s));
  if (st) {
    st->nnz += onnz;
    st->recovered_cols += (int64_t)hc[kJobRecover];
    st->selected_cols += (int64_t)hc[kJobSelect];
    st->reselected_cols += (int64_t)hc[3];
  }
  return CBX_OK;
}

static cbx_status mcl_prune_dispatch(const cbx_dcsc* C, double thr, int64_t select, int64_t recover, double pct,
                                     cbx_dcsc* out, cbx_mcl_stats* st, hipStream_t s) {
  if (C->dtype == CBX_F32) return mcl_prune_typed<float>(C, thr, select, recover, pct, out, st, s);
  if (C->dtype == CBX_F64) return mcl_prune_typed<double>(C, thr, select, recover, pct, out, st, s);
  return set_error(CBX_ERR_UNSUPPORTED, "MCL epilogue: values must be float32 or float64");
}

// Stored columns [q0, q1) of B as a DCSC of its own (same m, n).
static cbx_status column_slice(const cbx_dcsc* B, const int64_t* cp_host, int64_t q0, int64_t q1, cbx_dcsc* out,
                               hipStream_t s) {
  const int64_t e0 = cp_host[q0], e1 = cp_host[q1], nz = e1 - e0, nc = q1 - q0;
  CBX_TRY(alloc_dcsc(out, B->m, B->n, nz, nc, B->dtype));
  const size_t vs = dtype_size(B->dtype);
  k_rebase<<<nblocks(nc + 1, 256), 256, 0, s>>>(B->cp, q0, nc + 1, out->cp);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipMemcpyAsync(out->jc, B->jc + q0, sizeof(int64_t) * nc, hipMemcpyDeviceToDevice, s));
  if (nz) {
    CBX_HIP(hipMemcpyAsync(out->ir, B->ir + e0, sizeof(int64_t) * nz, hipMemcpyDeviceToDevice, s));
    CBX_HIP(hipMemcpyAsync(out->val, (const char*)B->val + vs * e0, vs * nz, hipMemcpyDeviceToDevice, s));
  }
  return CBX_OK;
}

// ColConcatenate of pieces holding increasing, disjoint column ranges.
static cbx_status concatenate(std::vector<cbx_dcsc>& parts, int64_t m, int64_t n, cbx_dtype dt, cbx_dcsc* out,
                              hipStream_t s) {
  int64_t nnz = 0, nzc = 0;
  for (auto& p : parts) {
    nnz += p.nnz;
    nzc += p.nzc;
  }
  CBX_TRY(alloc_dcsc(out, m, n, nnz, nzc, dt));
  const size_t vs = dtype_size(dt);
  int64_t at_c = 0, at_e = 0;
  for (auto& p : parts) {
    if (p.nzc) {
      k_shift<<<nblocks(p.nzc, 256), 256, 0, s>>>(p.cp, p.nzc, at_e, out->cp + at_c);
      CBX_HIP(hipGetLastError());
      CBX_HIP(hipMemcpyAsync(out->jc + at_c, p.jc, sizeof(int64_t) * p.nzc, hipMemcpyDeviceToDevice, s));
    }
    if (p.nnz) {
      CBX_HIP(hipMemcpyAsync(out->ir + at_e, p.ir, sizeof(int64_t) * p.nnz, hipMemcpyDeviceToDevice, s));
      CBX_HIP(hipMemcpyAsync((char*)out->val + vs * at_e, p.val, vs * p.nnz, hipMemcpyDeviceToDevice, s));
    }
    at_c += p.nzc;
    at_e += p.nnz;
  }
  CBX_HIP(hipMemcpyAsync(out->cp + nzc, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

namespace {
struct Parts {
  std::vector<cbx_dcsc> v;
  ~Parts() {
    for (auto& p : v) cbx_dcsc_free(&p);
  }
};
struct Owned {
  cbx_dcsc x{};
  ~Owned() { cbx_dcsc_free(&x); }
};
struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  Events() {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~Events() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  float span() {
    float f = 0;
    (void)hipEventElapsedTime(&f, a, b);
    return f;
  }
};
}  // namespace

static cbx_status mcl_expand(const cbx_dcsc* A, const cbx_dcsc* B, double thr, int64_t select, int64_t recover,
                             double pct, int64_t phases, cbx_spgemm_mode mode, const SpgemmConfig& cfg,
                             cbx_dcsc* out, cbx_mcl_stats* st, hipStream_t s) {
  const int64_t nzcB = B->nzc;
  std::vector<int64_t> flop(std::max<int64_t>(nzcB, 1), 0), cpB(nzcB + 1, 0);
  CBX_TRY(spgemm_flops(A, B, flop.data(), s));
  CBX_HIP(hipMemcpy(cpB.data(), B->cp, sizeof(int64_t) * (nzcB + 1), hipMemcpyDeviceToHost));
  int64_t total = 0;
  for (int64_t q = 0; q < nzcB; ++q) total += flop[q];
  st->flops = total;
  if (phases <= 0) {  // size phases so a phase's product (bounded by its flops) fits in 40% of free memory
    size_t freeb = 0, totb = 0;
    CBX_HIP(hipMemGetInfo(&freeb, &totb));
    freeb += cache_idle_bytes();  // cached output blocks and the stream pool's idle reserve are reusable
    int dev = 0;
    hipMemPool_t pool;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t reserved = 0, used = 0;
      if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) == hipSuccess &&
          hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess && reserved > used)
        freeb += (size_t)(reserved - used);
    }
    const double per_flop = (double)(sizeof(int64_t) + dtype_size(A->dtype));
    phases = (int64_t)std::ceil((double)total * per_flop / (0.4 * (double)freeb));
  }
  phases = std::max<int64_t>(1, std::min<int64_t>(phases, std::max<int64_t>(nzcB, 1)));
  st->phases = phases;
  // With no recovery possible (recover_num 0), the epilogue's result on the
  // product thresholded at hard_threshold equals its result on the whole
  // product: its statistics count entries above the threshold only, and a
  // selection cut (the select_num-th largest of more than select_num such
  // entries) lies above it.  The two-phase product then emits only those
  // entries (fused threshold) instead of all of them.
  SpgemmConfig pcfg = cfg;
  if (recover <= 0 && mode == CBX_SPGEMM_TWO_PHASE && cfg.mcl_fuse) {
    pcfg.filter = true;
    pcfg.filter_thr = thr;
  }
  Parts parts;
  Events ev;
  int64_t q0 = 0, acc = 0;
  for (int64_t ph = 0; ph < phases && q0 < nzcB; ++ph) {
    int64_t q1 = q0;
    const double goal = (double)total * (double)(ph + 1) / (double)phases;
    while (q1 < nzcB && (q1 == q0 || (double)(acc + flop[q1]) <= goal)) acc += flop[q1++];
    if (ph == phases - 1) {
      while (q1 < nzcB) acc += flop[q1++];
    }
    Owned Bp, Cp;
    CBX_TRY(column_slice(B, cpB.data(), q0, q1, &Bp.x, s));
    cbx_spgemm_stats sst;
    CBX_HIP(hipEventRecord(ev.a, s));
    CBX_TRY(spgemm_dispatch(A, &Bp.x, CBX_SR_PLUS_TIMES, mode, pcfg, &Cp.x, &sst, s));
    CBX_HIP(hipEventRecord(ev.b, s));
    CBX_HIP(hipEventSynchronize(ev.b));
    st->ms_product += ev.span();
    st->nnz_product += sst.nnz_before_filter;
    st->overflow_cols += sst.overflow_cols;
    st->learned_fallback += sst.learned_fallback;
    cbx_dcsc P{};
    CBX_HIP(hipEventRecord(ev.a, s));
    cbx_status ps = mcl_prune_dispatch(&Cp.x, thr, select, recover, pct, &P, st, s);
    if (ps) return ps;
    parts.v.push_back(P);
    CBX_HIP(hipEventRecord(ev.b, s));
    CBX_HIP(hipEventSynchronize(ev.b));
    st->ms_prune += ev.span();
    q0 = q1;
  }
  if (parts.v.size() == 1) {
    *out = parts.v[0];
    parts.v.clear();
    return CBX_OK;
  }
  return concatenate(parts.v, A->m, B->n, A->dtype, out, s);
}

}  // namespace cbx

using namespace cbx;

extern "C" cbx_status cbx_mcl_prune(const cbx_dcsc* C, double hard_threshold, int64_t select_num,
                                    int64_t recover_num, double recover_pct, cbx_dcsc* out, cbx_mcl_stats* stats,
                                    void* stream) {
  cbx_status st;
  if ((st = check_dcsc(C, "C", CBX_DEVICE))) return st;
  CBX_RETURN_IF(!out, CBX_ERR_INVALID, "cbx_mcl_prune: out is null");
  CBX_RETURN_IF(C->nzc >= (1ll << 31), CBX_ERR_LIMIT, "cbx_mcl_prune: at most 2^31-1 stored columns");
  if ((st = init_device_once())) return st;
  cbx_mcl_stats local;
  cbx_mcl_stats* sp = stats ? stats : &local;
  std::memset(sp, 0, sizeof(*sp));
  hipEvent_t a, b;
  CBX_HIP(hipEventCreate(&a));
  CBX_HIP(hipEventCreate(&b));
  (void)hipEventRecord(a, (hipStream_t)stream);
  st = mcl_prune_dispatch(C, hard_threshold, select_num, recover_num, recover_pct, out, sp, (hipStream_t)stream);
  (void)hipEventRecord(b, (hipStream_t)stream);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  sp->nnz_product = C->nnz;
  sp->ms_prune = ms;
  sp->ms_total = ms;
  return st;
}

extern "C" cbx_status cbx_mcl_expand(const cbx_dcsc* A, const cbx_dcsc* B, double hard_threshold,
                                     int64_t select_num, int64_t recover_num, double recover_pct, int64_t phases,
                                     cbx_spgemm_mode mode, cbx_dcsc* out, cbx_mcl_stats* stats, void* stream) {
  cbx_status st;
  if ((st = check_dcsc(A, "A", CBX_DEVICE))) return st;
  if ((st = check_dcsc(B, "B", CBX_DEVICE))) return st;
  CBX_RETURN_IF(!out, CBX_ERR_INVALID, "cbx_mcl_expand: out is null");
  CBX_RETURN_IF(A->n != B->m, CBX_ERR_INVALID, "cbx_mcl_expand: inner dimensions differ");
  CBX_RETURN_IF(A->dtype != B->dtype || (A->dtype != CBX_F32 && A->dtype != CBX_F64), CBX_ERR_UNSUPPORTED,
                "cbx_mcl_expand: A and B must both be float32 or both float64");
  CBX_RETURN_IF(A->nnz >= (1ll << 31) || B->nnz >= (1ll << 31), CBX_ERR_LIMIT,
                "cbx_mcl_expand: operands must hold fewer than 2^31 nonzeros each");
  CBX_RETURN_IF(mode != CBX_SPGEMM_TWO_PHASE && mode != CBX_SPGEMM_LEARNED, CBX_ERR_INVALID,
                "cbx_mcl_expand: mode must be two-phase or learned");
  if ((st = init_device_once())) return st;
  SpgemmConfig cfg = current_config();
  CBX_RETURN_IF(mode == CBX_SPGEMM_LEARNED && !cfg.predictor && cfg.sampling_rounds == 0, CBX_ERR_INVALID,
                "cbx_mcl_expand: learned mode needs a predictor (cbx_spgemm_set_predictor or "
                "cbx_spgemm_set_sampling_predictor)");
  cbx_mcl_stats local;
  cbx_mcl_stats* sp = stats ? stats : &local;
  std::memset(sp, 0, sizeof(*sp));
  std::memset(out, 0, sizeof(*out));
  hipEvent_t a, b;
  CBX_HIP(hipEventCreate(&a));
  CBX_HIP(hipEventCreate(&b));
  (void)hipEventRecord(a, (hipStream_t)stream);
  st = mcl_expand(A, B, hard_threshold, select_num, recover_num, recover_pct, phases, mode, cfg, out, sp,
                  (hipStream_t)stream);
  (void)hipEventRecord(b, (hipStream_t)stream);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  sp->ms_total = ms;
  return st;
}
