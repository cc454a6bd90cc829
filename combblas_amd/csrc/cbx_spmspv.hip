// Sparse matrix times sparse vector (SpMSpV) and top-down BFS for gfx950.
//
// cbx_spmspv replaces the local y = A*x of CombBLAS's SpMV with a sparse
// vector: SpImpl<SR,IT,NUM,IVT,OVT>::SpMXSpV (SpImpl.cpp:55-158) and its
// boolean-matrix form SpImpl<SR,IT,bool,IVT,OVT>::SpMXSpV (SpImpl.cpp:167-213),
// reached through generic_gespmv (Friends.h:448-482).  Output indices are
// ascending, an entry exists wherever some A(i,k) x_k term exists, and a
// boolean A contributes x_k itself (the boolean specialisation never calls
// multiply, SpImpl.cpp:183-188).  The GPU form is a dense accumulator over the
// m rows with a presence bitmap; products are block-flattened across the
// frontier's columns, and the output is compacted in row order by a scan of
// the bitmap's word counts.
//
// The BFS plan runs the level loop of TopDownBFS.cpp:442-448 with
// SelectMaxSRing<bool,int64> (parent = largest frontier neighbour):
//   fringe.setNumToInd(); fringe = SpMV(A, fringe); fringe = EWiseMult(fringe,
//   parents, exclude=true, -1); parents.Set(fringe);
// fused per level into two kernels: expand (every frontier vertex k offers k to
// each unvisited neighbour with an atomic max; the first offer appends the
// neighbour to the next frontier) and finalize (parents = offers; mark
// visited).  Each level is one captured hipGraph (two graphs alternate the
// frontier buffers); the host reads only the next frontier's size.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <type_traits>

#include "cbx_device.h"
#include "cbx_flatten.h"
#include "cbx_semiring.h"
#include "cbx_spgemm.h"

namespace cbx {

// ----------------------------------------------- global-memory accumulation
template <int K, class T>
struct GAcc;
template <>
struct GAcc<kPlusTimes, int64_t> {
  static __device__ void acc(int64_t* p, int64_t v) { atomicAdd((unsigned long long*)p, (unsigned long long)v); }
};
template <>
struct GAcc<kPlusTimes, double> {
  static __device__ void acc(double* p, double v) { atomicAdd(p, v); }
};
template <>
struct GAcc<kPlusTimes, float> {
  static __device__ void acc(float* p, float v) { atomicAdd(p, v); }
};
template <>
struct GAcc<kMinPlus, int64_t> {
  static __device__ void acc(int64_t* p, int64_t v) { atomicMin((long long*)p, (long long)v); }
};
template <>
struct GAcc<kSelectMax, int64_t> {
  static __device__ void acc(int64_t* p, int64_t v) { atomicMax((long long*)p, (long long)v); }
};
template <class F, class U>
__device__ void gcas_minmax(F* p, F v, bool is_min) {
  U* up = reinterpret_cast<U*>(p);
  U old = __hip_atomic_load(up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const F cur = __builtin_bit_cast(F, old);
    if (!(is_min ? (v < cur) : (cur < v))) return;
    if (__hip_atomic_compare_exchange_strong(up, &old, __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}
template <>
struct GAcc<kMinPlus, double> {
  static __device__ void acc(double* p, double v) { gcas_minmax<double, unsigned long long>(p, v, true); }
};
template <>
struct GAcc<kMinPlus, float> {
  static __device__ void acc(float* p, float v) { gcas_minmax<float, unsigned int>(p, v, true); }
};
template <>
struct GAcc<kSelectMax, double> {
  static __device__ void acc(double* p, double v) { gcas_minmax<double, unsigned long long>(p, v, false); }
};
template <>
struct GAcc<kSelectMax, float> {
  static __device__ void acc(float* p, float v) { gcas_minmax<float, unsigned int>(p, v, false); }
};

namespace {

__global__ void k_sv_colmap(const int64_t* __restrict__ jc, int64_t nzc, int64_t n, int32_t* __restrict__ colmap) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) colmap[i] = -1;
}
__global__ void k_sv_colmap_scatter(const int64_t* __restrict__ jc, int64_t nzc, int32_t* __restrict__ colmap) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < nzc) colmap[jc[p]] = (int32_t)p;
}

template <class T>
__global__ void k_sv_init(T* __restrict__ acc, int64_t m, T ident, uint32_t* __restrict__ bits, int64_t nw) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < m) acc[i] = ident;
  if (i < nw) bits[i] = 0u;
}

// x entries must be valid column indices, ascending and unique.
__global__ void k_sv_check_x(const int32_t* __restrict__ xind, int32_t nx, int64_t n, int* __restrict__ bad) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nx) return;
  const int32_t k = xind[t];
  if (k < 0 || k >= n || (t + 1 < nx && xind[t + 1] <= k)) atomicOr(bad, 1);
}

// y_acc[i] (+)= A(i,k) (x) x_k for k in x; presence bit per row.
template <int K, class T, bool BOOLA, int NT, int U>
__global__ __launch_bounds__(NT) void k_sv_expand(const int64_t* __restrict__ cpA, const int64_t* __restrict__ irA,
                                                  const void* __restrict__ valA_, const int32_t* __restrict__ colmap,
                                                  const int32_t* __restrict__ xind, const T* __restrict__ xval,
                                                  int32_t nx, T* __restrict__ acc, uint32_t* __restrict__ bits) {
  using SR = Sem<K, T>;
  __shared__ SegStage<T, NT> st;
  __shared__ long long wsum[NT / kWave];
  const T* valA = (const T*)valA_;
  for (int64_t c0 = (int64_t)blockIdx.x * NT; c0 < nx; c0 += (int64_t)gridDim.x * NT) {
    const int64_t t = c0 + threadIdx.x;
    int64_t base = 0;
    int len = 0;
    T xv{};
    if (t < nx) {
      const int32_t p = colmap[xind[t]];
      if (p >= 0) {
        base = cpA[p];
        len = (int)(cpA[p + 1] - base);
        xv = xval[t];
      }
    }
    int nseg;
    const int tot = stage_chunk<T, NT>(st, wsum, base, len, xv, nseg);
    visit_staged<T, NT, U>(st, nseg, tot, [&](const int64_t(&idx)[U], const T(&xs)[U]) {
      int64_t r[U];
      T a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        r[u] = -1;
        if (idx[u] >= 0) {
          r[u] = irA[idx[u]];
          if constexpr (!BOOLA) a[u] = valA[idx[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (r[u] < 0) continue;
        T v;
        if constexpr (BOOLA) v = xs[u];
        else v = (T)SR::out(SR::mul(a[u], xs[u]));
        GAcc<K, T>::acc(&acc[r[u]], v);
        atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
      }
    });
  }
}

__global__ void k_sv_wordcnt(const uint32_t* __restrict__ bits, int64_t nw, int32_t* __restrict__ cnt) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w < nw) cnt[w] = __popc(bits[w]);
}

template <class T>
__global__ void k_sv_emit(const uint32_t* __restrict__ bits, const int32_t* __restrict__ off, int64_t nw,
                          const T* __restrict__ acc, int32_t* __restrict__ yind, T* __restrict__ yval) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= nw) return;
  uint32_t b = bits[w];
  int32_t q = off[w];
  while (b) {
    const int j = __ffs(b) - 1;
    b &= b - 1;
    const int32_t i = (int32_t)(w * 32 + j);
    yind[q] = i;
    yval[q] = acc[i];
    ++q;
  }
}

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

#define CBX_TRY_SV(x)              \
  do {                             \
    cbx_status st_ = (x);          \
    if (st_ != CBX_OK) return st_; \
  } while (0)

template <int K, class T>
T ident_of() {
  if (K == kPlusTimes) return std::is_floating_point<T>::value ? (T)-0.0 : (T)0;
  if (K == kMinPlus) return std::numeric_limits<T>::max();
  return std::numeric_limits<T>::lowest();
}

__global__ void k_sv_zero(int32_t* __restrict__ a, int32_t* __restrict__ b) {
  if (threadIdx.x == 0) {
    *a = 0;
    *b = 0;
  }
}

// Scratch of one SpMSpV: the column map of A (n), the dense accumulator (m
// values) and presence bitmap (nw words), per-word counts and their scan
// (nw + 1 each), a flag and the scan's temporary storage.
struct SvWork {
  int32_t *colmap = nullptr, *cnt = nullptr, *off = nullptr;
  void* acc = nullptr;
  uint32_t* bits = nullptr;
  int* bad = nullptr;
  void* tmp = nullptr;
  size_t tb = 0;
};

inline int64_t sv_words(int64_t m) { return m / 32 + 1; }

template <class T>
size_t sv_scan_bytes(int64_t m) {
  size_t tb = 0;
  const int64_t nw = sv_words(m);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int32_t*)nullptr, (int32_t*)nullptr, (int)(nw + 1));
  return std::max<size_t>(tb, 16);
}

// The product with its scratch in `w` (every array is initialised here, so
// the scratch may hold anything on entry).
template <int K, class T>
cbx_status spmspv_run(const cbx_dcsc* A, bool boolA, const int32_t* xind, const T* xval, int32_t nx, int32_t* yind,
                      T* yval, int32_t* ny, const SvWork& w, hipStream_t s) {
  const int64_t m = A->m, n = A->n, nw = sv_words(m);
  T* acc = (T*)w.acc;
  int hbad = 0;
  k_sv_zero<<<1, 64, 0, s>>>(w.bad, w.cnt + nw);
  k_sv_check_x<<<nb(nx, 256), 256, 0, s>>>(xind, nx, n, w.bad);
  CBX_HIP(hipMemcpyAsync(&hbad, w.bad, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  if (hbad) return set_error(CBX_ERR_INVALID, "cbx_spmspv: x indices must be ascending, unique and below ncols");
  k_sv_colmap<<<nb(n, 256), 256, 0, s>>>(A->jc, A->nzc, n, w.colmap);
  k_sv_colmap_scatter<<<nb(A->nzc, 256), 256, 0, s>>>(A->jc, A->nzc, w.colmap);
  k_sv_init<T><<<nb(std::max(m, nw), 256), 256, 0, s>>>(acc, m, ident_of<K, T>(), w.bits, nw);
  const unsigned g = std::min<unsigned>(nb(nx, 256), 2048u);
  if (boolA)
    k_sv_expand<K, T, true, 256, 4><<<g, 256, 0, s>>>(A->cp, A->ir, A->val, w.colmap, xind, xval, nx, acc, w.bits);
  else
    k_sv_expand<K, T, false, 256, 4><<<g, 256, 0, s>>>(A->cp, A->ir, A->val, w.colmap, xind, xval, nx, acc, w.bits);
  k_sv_wordcnt<<<nb(nw, 256), 256, 0, s>>>(w.bits, nw, w.cnt);
  size_t tb = w.tb;
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(w.tmp, tb, w.cnt, w.off, (int)(nw + 1), s));
  k_sv_emit<T><<<nb(nw, 256), 256, 0, s>>>(w.bits, w.off, nw, acc, yind, yval);
  CBX_HIP(hipMemcpyAsync(ny, w.off + nw, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

template <int K, class T>
cbx_status spmspv_typed(const cbx_dcsc* A, bool boolA, const int32_t* xind, const T* xval, int32_t nx, int32_t* yind,
                        T* yval, int32_t* ny, hipStream_t s) {
  const int64_t m = A->m, n = A->n, nw = sv_words(m);
  *ny = 0;
  if (nx == 0 || A->nnz == 0) return CBX_OK;
  SvWork w;
  w.tb = sv_scan_bytes<T>(m);
  auto cleanup = [&]() {
    for (void* p : {(void*)w.colmap, (void*)w.cnt, (void*)w.off, w.acc, (void*)w.bits, (void*)w.bad, w.tmp})
      if (p) (void)hipFreeAsync(p, s);
  };
  if (hipMallocAsync(&w.colmap, sizeof(int32_t) * std::max<int64_t>(n, 1), s) != hipSuccess ||
      hipMallocAsync(&w.cnt, sizeof(int32_t) * (nw + 1), s) != hipSuccess ||
      hipMallocAsync(&w.off, sizeof(int32_t) * (nw + 1), s) != hipSuccess ||
      hipMallocAsync(&w.acc, sizeof(T) * std::max<int64_t>(m, 1), s) != hipSuccess ||
      hipMallocAsync(&w.bits, sizeof(uint32_t) * nw, s) != hipSuccess ||
      hipMallocAsync(&w.bad, sizeof(int), s) != hipSuccess || hipMallocAsync(&w.tmp, w.tb, s) != hipSuccess) {
    cleanup();
    (void)hipGetLastError();
    return set_error(CBX_ERR_NOMEM, "cbx_spmspv: workspace allocation failed");
  }
  const cbx_status st = spmspv_run<K, T>(A, boolA, xind, xval, nx, yind, yval, ny, w, s);
  cleanup();
  CBX_HIP(hipStreamSynchronize(s));
  return st;
}

// Carves the scratch of one SpMSpV (the arrays of SvWork plus the structure
// check's three counters) out of one caller-given region; with base == null
// it only adds up the bytes.  Every piece starts on a 256-byte boundary.
template <class T>
size_t sv_carve(int64_t m, int64_t n, char* base, SvWork* w, unsigned long long** check_cnt) {
  const int64_t nw = sv_words(m);
  size_t used = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = base ? base + used : nullptr;
    used += (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    return p;
  };
  SvWork t;
  t.tb = sv_scan_bytes<T>(m);
  t.colmap = (int32_t*)take(sizeof(int32_t) * std::max<int64_t>(n, 1));
  t.cnt = (int32_t*)take(sizeof(int32_t) * (nw + 1));
  t.off = (int32_t*)take(sizeof(int32_t) * (nw + 1));
  t.acc = take(sizeof(T) * std::max<int64_t>(m, 1));
  t.bits = (uint32_t*)take(sizeof(uint32_t) * nw);
  t.bad = (int*)take(sizeof(int));
  t.tmp = take(t.tb);
  unsigned long long* cnt = (unsigned long long*)take(3 * sizeof(unsigned long long));
  if (w) *w = t;
  if (check_cnt) *check_cnt = cnt;
  return used;
}

// cbx_spmspv with its scratch in `region` (bytes of device memory): the
// structure check and the product, no allocation.
template <int K, class T>
cbx_status spmspv_in_region(const cbx_dcsc* A, bool boolA, const int32_t* xind, const T* xval, int32_t nx,
                            int32_t* yind, T* yval, int32_t* ny, char* region, size_t bytes, hipStream_t s) {
  *ny = 0;
  const size_t need = sv_carve<T>(A->m, A->n, nullptr, nullptr, nullptr);
  CBX_RETURN_IF(need > bytes, CBX_ERR_LIMIT,
                "cbx_spmspv_noalloc: a %lld x %lld matrix needs %zu bytes of scratch, more than the %zu available",
                (long long)A->m, (long long)A->n, need, bytes);
  SvWork w;
  unsigned long long* cnt = nullptr;
  sv_carve<T>(A->m, A->n, region, &w, &cnt);
  if (A->nzc == 0) {
    CBX_RETURN_IF(A->nnz != 0, CBX_ERR_INVALID, "cbx_dcsc_check: nnz without columns");
  } else {
    dcsc_check_launch(A, cnt, s);  // the kernels index by A's rows and columns
    const cbx_status st = dcsc_check_result(cnt, s);
    if (st != CBX_OK) return st;
  }
  if (nx == 0 || A->nnz == 0) return CBX_OK;
  return spmspv_run<K, T>(A, boolA, xind, xval, nx, yind, yval, ny, w, s);
}

}  // namespace

// ====================================================================== BFS
namespace {

__global__ void k_bfs_dense_cp(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc, int64_t nzc,
                               int64_t n, int64_t* __restrict__ beg, int64_t* __restrict__ end) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < nzc) {
    beg[jc[p]] = cp[p];
    end[jc[p]] = cp[p + 1];
  }
}

// Number of positions where two int64 arrays differ (the symmetry check of
// a plan: A's arrays against its transpose's).
__global__ void k_bfs_count_diff(const int64_t* __restrict__ a, const int64_t* __restrict__ b, int64_t n,
                                 unsigned long long* __restrict__ cnt) {
  long long d = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d += a[i] != b[i];
  d = wave_sum_ll(d);
  if (lane_id() == 0 && d) atomicAdd(cnt, (unsigned long long)d);
}

__global__ void k_bfs_rows32(const int64_t* __restrict__ ir, int64_t nnz, int32_t* __restrict__ rows) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nnz) rows[i] = (int32_t)ir[i];
}

// Frontier vertices with more than kBfsHub neighbours ("hubs") are expanded
// in pieces of kBfsHub edges by k_bfs_expand_hubs, one workgroup per piece,
// instead of by the workgroup that holds them in the frontier: R-MAT hubs have
// 10^5 neighbours and would otherwise keep one workgroup busy long after the
// others finished.  A piece is (vertex << 32 | piece index).
constexpr int64_t kBfsHub = 1024;

__device__ __forceinline__ void bfs_add_pieces(int32_t v, int64_t deg, int32_t* __restrict__ pcount,
                                               long long* __restrict__ pieces) {
  const int np = (int)((deg + kBfsHub - 1) / kBfsHub);
  const int b = atomicAdd(pcount, np);
  for (int k = 0; k < np; ++k) pieces[b + k] = ((long long)v << 32) | k;
}

__global__ void k_bfs_reset(int64_t n, int32_t* __restrict__ cand, int64_t* __restrict__ parents,
                            uint32_t* __restrict__ visited, int64_t nw, int32_t* __restrict__ f0,
                            int32_t* __restrict__ counts, int32_t* __restrict__ pcount,
                            long long* __restrict__ pieces0, const int64_t* __restrict__ beg,
                            const int64_t* __restrict__ end, unsigned long long* __restrict__ stat, int64_t root) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    cand[i] = -1;
    parents[i] = -1;
  }
  if (i < nw) visited[i] = 0u;
  if (i == 0) {
    f0[0] = (int32_t)root;
    counts[0] = 1;
    counts[1] = 0;
    pcount[0] = 0;
    pcount[1] = 0;
    stat[0] = 0ull;  // levels
    stat[1] = 0ull;  // vertices reached
    const int64_t deg = end[root] - beg[root];
    if (deg > kBfsHub) bfs_add_pieces((int32_t)root, deg, pcount, pieces0);
  }
}

// First kernel of a level: clears the next frontier's count and degree sum,
// and counts the level if its frontier is not empty (a level graph launched
// after the search ended changes nothing).
__global__ void k_bfs_level_begin(const int32_t* __restrict__ cin, int32_t* __restrict__ c,
                                  unsigned long long* __restrict__ edges, int32_t* __restrict__ pc,
                                  unsigned long long* __restrict__ stat) {
  if (threadIdx.x == 0) {
    *c = 0;
    *edges = 0ull;
    *pc = 0;
    if (*cin > 0) ++stat[0];
  }
}

// A frontier vertex offers itself (k) to neighbour r: SelectMax keeps the
// largest offer in cand[r]; the first offer appends r to the next frontier
// (one atomic per wave for the appends of this step).  All lanes call it.
__device__ __forceinline__ void bfs_offer(int32_t r, int32_t k, const uint32_t* __restrict__ visited,
                                          int32_t* __restrict__ cand, int32_t* __restrict__ fout,
                                          int32_t* __restrict__ fout_count) {
  bool push = false;
  if (r >= 0 && !((visited[r >> 5] >> (r & 31)) & 1u)) {
    // Offers only grow cand, so a value read without the atomic (possibly
    // stale, never above the true one) that already holds an offer at least
    // this large makes the atomic a no-op: skip it.  Hubs receive thousands
    // of offers per level and only the record-breaking ones reach the
    // serialising atomic.
    const int32_t seen = __hip_atomic_load(&cand[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (seen < k) {
      const int32_t old = atomicMax(&cand[r], k);
      push = old == -1;
    }
  }
  const uint64_t mask = __ballot(push);
  if (mask) {
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int basei = 0;
    if (lane_id() == leader) basei = atomicAdd(fout_count, __popcll(mask));
    basei = __shfl(basei, leader, kWave);
    if (push) fout[basei + __popcll(mask & ((1ull << lane_id()) - 1ull))] = r;
  }
}

// One level's SpMV(A, fringe) with SelectMax<bool,int64> fused with the
// EWiseMult(exclude) filter: each frontier vertex k offers k to its unvisited
// neighbours; cand[i] keeps the largest offer; the first offer appends i.
template <int NT, int U>
__global__ __launch_bounds__(NT) void k_bfs_expand(const int64_t* __restrict__ beg, const int64_t* __restrict__ end,
                                                   const int32_t* __restrict__ rows, const int32_t* __restrict__ fin,
                                                   const int32_t* __restrict__ fin_count,
                                                   const uint32_t* __restrict__ visited, int32_t* __restrict__ cand,
                                                   int32_t* __restrict__ fout, int32_t* __restrict__ fout_count) {
  __shared__ SegStage<int32_t, NT> st;
  __shared__ long long wsum[NT / kWave];
  const int nf = *fin_count;
  for (int c0 = blockIdx.x * NT; c0 < nf; c0 += gridDim.x * NT) {
    const int t = c0 + threadIdx.x;
    int64_t base = 0;
    int len = 0;
    int32_t k = 0;
    if (t < nf) {
      k = fin[t];
      base = beg[k];
      len = (int)(end[k] - base);
      if (len > kBfsHub) len = 0;  // a hub: k_bfs_expand_hubs takes its pieces
    }
    int nseg;
    const int tot = stage_chunk<int32_t, NT>(st, wsum, base, len, k, nseg);
    visit_staged<int32_t, NT, U>(st, nseg, tot, [&](const int64_t(&idx)[U], const int32_t(&ks)[U]) {
      int32_t r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) r[u] = idx[u] >= 0 ? rows[idx[u]] : -1;
#pragma unroll
      for (int u = 0; u < U; ++u) bfs_offer(r[u], ks[u], visited, cand, fout, fout_count);
    });
  }
}

// The hubs' pieces of the same level: one workgroup per piece of kBfsHub
// consecutive edges (coalesced row reads, kBfsHub / NT edges per thread).
template <int NT>
__global__ __launch_bounds__(NT) void k_bfs_expand_hubs(const int64_t* __restrict__ beg,
                                                        const int64_t* __restrict__ end,
                                                        const int32_t* __restrict__ rows,
                                                        const long long* __restrict__ pieces,
                                                        const int32_t* __restrict__ pcount,
                                                        const uint32_t* __restrict__ visited,
                                                        int32_t* __restrict__ cand, int32_t* __restrict__ fout,
                                                        int32_t* __restrict__ fout_count) {
  static_assert(kBfsHub % NT == 0, "whole rounds per piece");
  const int np = *pcount;
  for (int p = blockIdx.x; p < np; p += gridDim.x) {
    const long long pc = pieces[p];
    const int32_t v = (int32_t)(pc >> 32);
    const int64_t k = pc & 0xffffffffll;
    const int64_t e0 = beg[v] + k * kBfsHub, e1 = min(end[v], e0 + kBfsHub);
#pragma unroll
    for (int64_t e = e0 + threadIdx.x; e < e0 + kBfsHub; e += NT)  // uniform trip count (ballots inside)
      bfs_offer(e < e1 ? rows[e] : -1, v, visited, cand, fout, fout_count);
  }
}

// parents.Set(fringe) (FullyDistVec.cpp:1036) and the visited mark; also
// sums the new frontier's degrees (the host's direction choice), one atomic
// per workgroup.
constexpr int kBfsFinNT = 256;
__global__ __launch_bounds__(kBfsFinNT) void k_bfs_finalize(const int32_t* __restrict__ f,
                                                            const int32_t* __restrict__ count,
                                                            const int32_t* __restrict__ cand,
                                                            const int64_t* __restrict__ beg,
                                                            const int64_t* __restrict__ end,
                                                            int64_t* __restrict__ parents,
                                                            uint32_t* __restrict__ visited,
                                                            unsigned long long* __restrict__ edges,
                                                            unsigned long long* __restrict__ reached,
                                                            int32_t* __restrict__ pcount,
                                                            long long* __restrict__ pieces) {
  __shared__ long long ws[kBfsFinNT / kWave];
  const int nf = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && nf) atomicAdd(reached, (unsigned long long)nf);
  long long deg = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nf; t += gridDim.x * blockDim.x) {
    const int32_t i = f[t];
    parents[i] = cand[i];
    atomicOr(&visited[i >> 5], 1u << (i & 31));
    const int64_t d = end[i] - beg[i];
    deg += d;
    if (d > kBfsHub) bfs_add_pieces(i, d, pcount, pieces);  // the next level expands hubs in pieces
  }
  deg = wave_sum_ll(deg);
  if (lane_id() == 0) ws[wave_id()] = deg;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long d = 0;
    for (int w = 0; w < kBfsFinNT / kWave; ++w) d += ws[w];
    if (d) atomicAdd(edges, (unsigned long long)d);
  }
}

// Bottom-up form of the same level (for large frontiers): the frontier as a
// bitmap, then every unvisited vertex scans its neighbours from the largest
// index down and stops at the first one in the frontier -- the largest
// frontier neighbour, the value SelectMax's atomic max gives top-down, so
// both forms produce the same parents and next frontier.
__global__ void k_bfs_fbits_clear(uint32_t* __restrict__ fbits, int64_t nw) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nw) fbits[i] = 0u;
}

__global__ void k_bfs_fbits_set(const int32_t* __restrict__ f, const int32_t* __restrict__ count,
                                uint32_t* __restrict__ fbits) {
  const int nf = *count;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nf; t += gridDim.x * blockDim.x)
    atomicOr(&fbits[f[t] >> 5], 1u << (f[t] & 31));
}

// Vertices of degree above kBfsSerialDeg are scanned by the whole wave, 64
// neighbours per step from the top (hubs would stall a lane for thousands of
// steps); the first step with a hit gives the largest frontier neighbour.
constexpr int64_t kBfsSerialDeg = 64;

__global__ __launch_bounds__(256) void k_bfs_bottom_up(const int64_t* __restrict__ beg,
                                                       const int64_t* __restrict__ end,
                                                       const int32_t* __restrict__ rows, int64_t n,
                                                       const uint32_t* __restrict__ fbits,
                                                       const uint32_t* __restrict__ visited,
                                                       int32_t* __restrict__ cand, int32_t* __restrict__ fout,
                                                       int32_t* __restrict__ fout_count) {
  const int lane = lane_id();
  for (int64_t v0 = blockIdx.x * (int64_t)blockDim.x; v0 < n; v0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = v0 + threadIdx.x;  // the loop bound is uniform per wave (ballots below)
    bool push = false, heavy = false;
    int64_t b = 0, e = 0;
    if (v < n && !((visited[v >> 5] >> (v & 31)) & 1u)) {
      b = beg[v];
      e = end[v];
      if (e - b > kBfsSerialDeg) {
        heavy = true;
      } else {
        for (int64_t i = e - 1; i >= b; --i) {
          const int32_t u = rows[i];
          if ((fbits[u >> 5] >> (u & 31)) & 1u) {
            cand[v] = u;
            push = true;
            break;
          }
        }
      }
    }
    uint64_t hv = __ballot(heavy);
    while (hv) {  // the wave's high-degree vertices, one at a time
      const int h = __ffsll((unsigned long long)hv) - 1;
      hv &= hv - 1;
      const int64_t hb = __shfl(b, h, kWave), he = __shfl(e, h, kWave);
      for (int64_t top = he; top > hb; top -= kWave) {
        const int64_t i = top - 1 - lane;  // lane 0 holds the largest index of the step
        int32_t u = -1;
        if (i >= hb) {
          u = rows[i];
          if (!((fbits[u >> 5] >> (u & 31)) & 1u)) u = -1;
        }
        const uint64_t hit = __ballot(u >= 0);
        if (hit) {
          const int first = __ffsll((unsigned long long)hit) - 1;
          const int32_t par = __shfl(u, first, kWave);
          if (lane == h) {
            cand[v] = par;
            push = true;
          }
          break;
        }
      }
    }
    const uint64_t mask = __ballot(push);
    if (mask) {
      const int leader = __ffsll((unsigned long long)mask) - 1;
      int basei = 0;
      if (lane == leader) basei = atomicAdd(fout_count, __popcll(mask));
      basei = __shfl(basei, leader, kWave);
      if (push) fout[basei + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)v;
    }
  }
}

}  // namespace

}  // namespace cbx

using namespace cbx;

struct cbx_bfs_plan_s {
  int64_t n = 0, nnz = 0, nw = 0;
  hipStream_t stream = nullptr;
  int64_t *beg = nullptr, *end = nullptr, *parents = nullptr;
  int32_t *rows = nullptr, *f[2] = {nullptr, nullptr}, *counts = nullptr;
  int32_t* pcount = nullptr;                  // hub pieces of the two frontier buffers
  long long* pieces[2] = {nullptr, nullptr};
  int32_t* cand = nullptr;  // the largest offer per vertex (vertex ids fit 32 bits: n < 2^31)
  uint32_t* visited = nullptr;
  unsigned long long* edges = nullptr;  // degree sums of the two frontier buffers
  uint32_t* fbits = nullptr;             // frontier bitmap of a bottom-up level
  unsigned long long* stat = nullptr;  // device: levels, vertices reached
  // pinned, 64 B: next frontier size; [2..3] its degree sum; [4..7] root's
  // range; [8..11] the device statistics
  int32_t* host_count = nullptr;
  int direction = CBX_BFS_DIRECTION_OPTIMIZING;
  // Pattern of A equal to that of its transpose.  Bottom-up levels scan column
  // v of A for frontier rows, which is row v only for a symmetric pattern, so
  // a plan over a directed graph runs every level top-down.
  bool symmetric = true;
  // [direction (0 top-down, 1 bottom-up)][frontier parity]
  hipGraph_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraphExec_t exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  // top-down only: kBfsTdLevels consecutive levels in one graph, launched
  // until the frontier is empty (levels after the last one do nothing)
  hipGraph_t td_graph = nullptr;
  hipGraphExec_t td_exec = nullptr;
};

static void bfs_plan_free(cbx_bfs_plan_s* p) {
  if (!p) return;
  for (int d = 0; d < 2; ++d)
    for (int i = 0; i < 2; ++i) {
      if (p->exec[d][i]) (void)hipGraphExecDestroy(p->exec[d][i]);
      if (p->graph[d][i]) (void)hipGraphDestroy(p->graph[d][i]);
    }
  if (p->td_exec) (void)hipGraphExecDestroy(p->td_exec);
  if (p->td_graph) (void)hipGraphDestroy(p->td_graph);
  for (int i = 0; i < 2; ++i) {
    if (p->f[i]) (void)hipFree(p->f[i]);
    if (p->pieces[i]) (void)hipFree(p->pieces[i]);
  }
  for (void* q : {(void*)p->beg, (void*)p->end, (void*)p->parents, (void*)p->rows, (void*)p->counts, (void*)p->cand,
                  (void*)p->visited, (void*)p->edges, (void*)p->fbits, (void*)p->stat, (void*)p->pcount})
    if (q) (void)hipFree(q);
  if (p->host_count) (void)hipHostFree(p->host_count);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

constexpr int kBfsNT = 256;
constexpr unsigned kBfsGrid = 2048;
constexpr unsigned kBfsHubGrid = 2048;  // workgroups over the hubs' pieces
constexpr int64_t kBfsAlpha = 256;
constexpr int kBfsTdLevels = 8;  // even: the frontier parity is 0 again after one top-down graph

extern "C" cbx_status cbx_bfs_plan_create(const cbx_dcsc* A, cbx_bfs_plan_s** out) {
  CBX_RETURN_IF(!A || !out, CBX_ERR_INVALID, "cbx_bfs_plan_create: null argument");
  *out = nullptr;
  CBX_RETURN_IF(A->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_bfs_plan_create: expected a device matrix");
  CBX_RETURN_IF(A->m != A->n, CBX_ERR_INVALID, "cbx_bfs_plan_create: the adjacency matrix must be square");
  CBX_RETURN_IF(A->n >= (1ll << 31) || A->nnz >= (1ll << 31), CBX_ERR_LIMIT,
                "cbx_bfs_plan_create: n and nnz must be below 2^31");
  cbx_status st = cbx_dcsc_check(A);
  if (st != CBX_OK) return st;
  // symmetric pattern?  (A's structure against its transpose's)
  bool symmetric = true;
  {
    cbx_dcsc T;
    std::memset(&T, 0, sizeof(T));
    st = cbx_dcsc_transpose(A, &T, nullptr);
    if (st != CBX_OK) return st;
    symmetric = T.nzc == A->nzc && T.nnz == A->nnz;
    if (symmetric && A->nnz) {
      unsigned long long* cnt = nullptr;
      unsigned long long diff = 0;
      if (hipMalloc(&cnt, sizeof(*cnt)) != hipSuccess) {
        cbx_dcsc_free(&T);
        return set_error(CBX_ERR_NOMEM, "cbx_bfs_plan_create: allocation");
      }
      (void)hipMemset(cnt, 0, sizeof(*cnt));
      k_bfs_count_diff<<<1024, 256>>>(A->cp, T.cp, A->nzc + 1, cnt);
      k_bfs_count_diff<<<1024, 256>>>(A->jc, T.jc, A->nzc, cnt);
      k_bfs_count_diff<<<1024, 256>>>(A->ir, T.ir, A->nnz, cnt);
      const hipError_t e = hipMemcpy(&diff, cnt, sizeof(diff), hipMemcpyDeviceToHost);
      (void)hipFree(cnt);
      if (e != hipSuccess) {
        cbx_dcsc_free(&T);
        return set_error(CBX_ERR_HIP, "cbx_bfs_plan_create: symmetry check: %s", hipGetErrorString(e));
      }
      symmetric = diff == 0;
    }
    cbx_dcsc_free(&T);
  }
  auto* p = new cbx_bfs_plan_s();
  p->symmetric = symmetric;
  p->n = A->n;
  p->nnz = A->nnz;
  p->nw = A->n / 32 + 1;
  auto fail = [&](const char* what) {
    bfs_plan_free(p);
    return set_error(CBX_ERR_HIP, "cbx_bfs_plan_create: %s", what);
  };
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) return fail("stream");
  const int64_t n = p->n;
  // every hub of a frontier contributes ceil(deg / kBfsHub) <= 2 deg / kBfsHub pieces
  const int64_t max_pieces = 2 * p->nnz / kBfsHub + 2;
  if (hipMalloc(&p->beg, 8 * std::max<int64_t>(n, 1)) || hipMalloc(&p->end, 8 * std::max<int64_t>(n, 1)) ||
      hipMalloc(&p->parents, 8 * std::max<int64_t>(n, 1)) || hipMalloc(&p->rows, 4 * std::max<int64_t>(p->nnz, 1)) ||
      hipMalloc(&p->f[0], 4 * std::max<int64_t>(n, 1)) || hipMalloc(&p->f[1], 4 * std::max<int64_t>(n, 1)) ||
      hipMalloc(&p->counts, 16) || hipMalloc(&p->cand, sizeof(int32_t) * std::max<int64_t>(n, 1)) ||
      hipMalloc(&p->visited, 4 * p->nw) || hipMalloc(&p->edges, 2 * sizeof(unsigned long long)) ||
      hipMalloc(&p->fbits, 4 * p->nw) || hipMalloc(&p->stat, 2 * sizeof(unsigned long long)) ||
      hipMalloc(&p->pcount, 2 * sizeof(int32_t)) || hipMalloc(&p->pieces[0], 8 * max_pieces) ||
      hipMalloc(&p->pieces[1], 8 * max_pieces) ||
      hipHostMalloc(&p->host_count, 64, hipHostMallocDefault))
    return fail("allocation");
  hipStream_t s = p->stream;
  if (hipMemsetAsync(p->beg, 0, 8 * n, s) || hipMemsetAsync(p->end, 0, 8 * n, s)) return fail("memset");
  k_bfs_dense_cp<<<nb(A->nzc, 256), 256, 0, s>>>(A->cp, A->jc, A->nzc, n, p->beg, p->end);
  k_bfs_rows32<<<nb(A->nnz, 256), 256, 0, s>>>(A->ir, A->nnz, p->rows);
  if (hipStreamSynchronize(s) != hipSuccess) return fail("setup kernels");
  // one graph per (direction, frontier parity): clear the next count, the
  // level's kernels, finalize, copy the next count and degree sum to pinned
  // host memory
  const unsigned bu_grid = (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 8192);
  for (int dir = 0; dir < 2; ++dir)
    for (int par = 0; par < 2; ++par) {
      int32_t* cin = p->counts + par;
      int32_t* cout = p->counts + (1 - par);
      unsigned long long* eout = p->edges + (1 - par);
      if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail("capture begin");
      k_bfs_level_begin<<<1, 64, 0, s>>>(cin, cout, eout, p->pcount + (1 - par), p->stat);
      if (dir == 0) {
        k_bfs_expand<kBfsNT, 4><<<kBfsGrid, kBfsNT, 0, s>>>(p->beg, p->end, p->rows, p->f[par], cin, p->visited,
                                                           p->cand, p->f[1 - par], cout);
        k_bfs_expand_hubs<kBfsNT><<<kBfsHubGrid, kBfsNT, 0, s>>>(p->beg, p->end, p->rows, p->pieces[par],
                                                                 p->pcount + par, p->visited, p->cand,
                                                                 p->f[1 - par], cout);
      } else {
        k_bfs_fbits_clear<<<nb(p->nw, 256), 256, 0, s>>>(p->fbits, p->nw);
        k_bfs_fbits_set<<<256, 256, 0, s>>>(p->f[par], cin, p->fbits);
        k_bfs_bottom_up<<<bu_grid, 256, 0, s>>>(p->beg, p->end, p->rows, n, p->fbits, p->visited, p->cand,
                                                p->f[1 - par], cout);
      }
      k_bfs_finalize<<<256, kBfsFinNT, 0, s>>>(p->f[1 - par], cout, p->cand, p->beg, p->end, p->parents, p->visited,
                                               eout, p->stat + 1, p->pcount + (1 - par), p->pieces[1 - par]);
      (void)hipMemcpyAsync(p->host_count, cout, sizeof(int32_t), hipMemcpyDeviceToHost, s);
      (void)hipMemcpyAsync(p->host_count + 2, eout, sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
      if (hipStreamEndCapture(s, &p->graph[dir][par]) != hipSuccess) return fail("capture end");
      if (hipGraphInstantiate(&p->exec[dir][par], p->graph[dir][par], nullptr, nullptr, 0) != hipSuccess)
        return fail("graph instantiate");
    }
  // kBfsTdLevels top-down levels in one graph, no host round trip between
  // them; the host reads the frontier size once per graph
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail("capture begin");
  for (int l = 0; l < kBfsTdLevels; ++l) {
    const int par = l & 1;
    int32_t* cin = p->counts + par;
    int32_t* cout = p->counts + (1 - par);
    unsigned long long* eout = p->edges + (1 - par);
    k_bfs_level_begin<<<1, 64, 0, s>>>(cin, cout, eout, p->pcount + (1 - par), p->stat);
    k_bfs_expand<kBfsNT, 4><<<kBfsGrid, kBfsNT, 0, s>>>(p->beg, p->end, p->rows, p->f[par], cin, p->visited, p->cand,
                                                       p->f[1 - par], cout);
    k_bfs_expand_hubs<kBfsNT><<<kBfsHubGrid, kBfsNT, 0, s>>>(p->beg, p->end, p->rows, p->pieces[par], p->pcount + par,
                                                             p->visited, p->cand, p->f[1 - par], cout);
    k_bfs_finalize<<<256, kBfsFinNT, 0, s>>>(p->f[1 - par], cout, p->cand, p->beg, p->end, p->parents, p->visited,
                                             eout, p->stat + 1, p->pcount + (1 - par), p->pieces[1 - par]);
  }
  (void)hipMemcpyAsync(p->host_count, p->counts, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (hipStreamEndCapture(s, &p->td_graph) != hipSuccess) return fail("capture end");
  if (hipGraphInstantiate(&p->td_exec, p->td_graph, nullptr, nullptr, 0) != hipSuccess)
    return fail("graph instantiate");
  *out = p;
  return CBX_OK;
}

extern "C" cbx_status cbx_bfs_run(cbx_bfs_plan_s* p, int64_t root, int64_t* parents_dev, cbx_bfs_stats* stats) {
  CBX_RETURN_IF(!p || !parents_dev, CBX_ERR_INVALID, "cbx_bfs_run: null argument");
  CBX_RETURN_IF(root < 0 || root >= p->n, CBX_ERR_INVALID, "cbx_bfs_run: root %lld out of range", (long long)root);
  hipStream_t s = p->stream;
  hipEvent_t e0, e1;
  CBX_HIP(hipEventCreate(&e0));
  CBX_HIP(hipEventCreate(&e1));
  CBX_HIP(hipEventRecord(e0, s));
  k_bfs_reset<<<nb(std::max(p->n, p->nw), 256), 256, 0, s>>>(p->n, p->cand, p->parents, p->visited, p->nw, p->f[0],
                                                             p->counts, p->pcount, p->pieces[0], p->beg, p->end,
                                                             p->stat, root);
  // direction per level: bottom-up while the frontier's edges exceed 1/256 of
  // the edges still unexplored.  (Beamer's rule compares the two with a factor
  // 14 for equal per-edge costs; here a top-down edge costs a device atomic on
  // the target, some 20-40 times a bottom-up edge, which is a streamed read.)
  int bottom_up_levels = 0;
  if (p->direction == CBX_BFS_TOP_DOWN || !p->symmetric) {
    // every level top-down: graphs of kBfsTdLevels levels until the frontier
    // is empty (a BFS has at most n levels)
    for (int64_t launched = 0;; launched += kBfsTdLevels) {
      CBX_HIP(hipGraphLaunch(p->td_exec, s));
      CBX_HIP(hipStreamSynchronize(s));
      if (*(volatile int32_t*)p->host_count == 0 || launched > p->n + 1) break;
    }
  } else {
    int levels = 0;
    int64_t mu = p->nnz;
    int64_t* hdeg = reinterpret_cast<int64_t*>(p->host_count + 4);  // pinned: beg[root], end[root]
    CBX_HIP(hipMemcpyAsync(hdeg, p->beg + root, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipMemcpyAsync(hdeg + 1, p->end + root, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipStreamSynchronize(s));
    int64_t fe = hdeg[1] - hdeg[0];  // edges of the current frontier
    bool bottom_up = p->direction == CBX_BFS_BOTTOM_UP;
    for (int par = 0;; par ^= 1) {
      if (p->direction == CBX_BFS_DIRECTION_OPTIMIZING) bottom_up = fe > mu / kBfsAlpha;
      CBX_HIP(hipGraphLaunch(p->exec[bottom_up ? 1 : 0][par], s));
      CBX_HIP(hipStreamSynchronize(s));
      ++levels;
      bottom_up_levels += bottom_up;
      const int32_t next = *(volatile int32_t*)p->host_count;
      fe = (int64_t) * (volatile unsigned long long*)(p->host_count + 2);
      mu -= fe;
      if (next == 0 || levels > p->n + 1) break;
    }
  }
  unsigned long long* hstat = reinterpret_cast<unsigned long long*>(p->host_count + 8);
  CBX_HIP(hipMemcpyAsync(hstat, p->stat, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(parents_dev, p->parents, sizeof(int64_t) * p->n, hipMemcpyDeviceToDevice, s));
  CBX_HIP(hipEventRecord(e1, s));
  CBX_HIP(hipStreamSynchronize(s));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (stats) {
    stats->levels = (int32_t)hstat[0];
    stats->bottom_up_levels = bottom_up_levels;
    stats->reached = (int64_t)hstat[1];
    stats->ms = ms;
  }
  return CBX_OK;
}

extern "C" cbx_status cbx_bfs_plan_set_direction(cbx_bfs_plan_s* p, cbx_bfs_direction direction) {
  CBX_RETURN_IF(!p, CBX_ERR_INVALID, "cbx_bfs_plan_set_direction: null plan");
  CBX_RETURN_IF(direction != CBX_BFS_DIRECTION_OPTIMIZING && direction != CBX_BFS_TOP_DOWN &&
                    direction != CBX_BFS_BOTTOM_UP,
                CBX_ERR_INVALID, "cbx_bfs_plan_set_direction: unknown direction %d", (int)direction);
  CBX_RETURN_IF(direction == CBX_BFS_BOTTOM_UP && !p->symmetric, CBX_ERR_INVALID,
                "cbx_bfs_plan_set_direction: bottom-up levels need a symmetric pattern (this plan's is not)");
  p->direction = direction;
  return CBX_OK;
}

extern "C" void cbx_bfs_plan_destroy(cbx_bfs_plan_s* p) { bfs_plan_free(p); }


// ------------------------------------------------------ SpMV, dense vector
// dcsc_gespmv (Friends.h:63-77): y[i] = add(y[i], multiply(A(i,j), x[j])) for
// every stored entry, y accumulated in place; a boolean A contributes x[j]
// itself.
//
// Bucketed form (m <= 2^24): the entries are cut into tiles of kSvTile; pass 1
// counts each tile's entries per 4096-row block, a scan over (block, tile)
// gives every tile a slot range per block, pass 2 writes each product with its
// row offset into its block's bucket, and pass 3 accumulates each block's
// bucket in LDS (chunks of kSvChunk entries per workgroup) and updates y once
// per row: a plain read-modify-write when one workgroup owns the block, the
// semiring's device atomic when a hot block is split over several.  Device
// atomics per entry (the direct form below, kept for taller matrices and small
// products) serialise on the hot rows of skewed matrices.
constexpr int kSvLB = 12, kSvBR = 1 << kSvLB;  // rows per block
constexpr int kSvMaxBlocks = 4096;            // bucketed form while m <= 2^24
constexpr int kSvTile = 4096, kSvNT = 256;    // entries and threads per counting / scattering workgroup
constexpr int64_t kSvChunk = 32768;           // bucket entries per accumulating workgroup
constexpr int64_t kSvMinNnz = 1 << 16;        // below this the direct form is cheaper

template <int K, class T, bool BOOLA>
__global__ void k_spmv_dense(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc,
                             const int64_t* __restrict__ ir, const void* __restrict__ valA_, int64_t nzc,
                             const T* __restrict__ x, T* __restrict__ y) {
  using SR = Sem<K, T>;
  const int64_t q = global_wave();
  if (q >= nzc) return;
  const T xj = x[jc[q]];
  for (int64_t e = cp[q] + lane_id(); e < cp[q + 1]; e += kWave) {
    T prod;
    if constexpr (BOOLA) prod = xj;
    else prod = (T)SR::out(SR::mul(((const T*)valA_)[e], xj));
    GAcc<K, T>::acc(&y[ir[e]], prod);
  }
}

// pass 1: hist[b * ntiles + t] = entries of tile t in row block b.  It also
// does the per-entry half of the structure check: part[2t] counts the tile's
// rows outside [0, nrows) (only counted, never used as an index) and
// part[2t+1] its descents ir[e+1] <= ir[e]; k_svb_verdict sums them (one slot
// per tile: thousands of workgroups adding to one counter would serialise).
__global__ __launch_bounds__(kSvNT) void k_svb_count(const int64_t* __restrict__ ir, int64_t nnz, int64_t nrows,
                                                     int nblk, int64_t ntiles, int64_t* __restrict__ hist,
                                                     long long* __restrict__ part) {
  __shared__ int h[kSvMaxBlocks];
  __shared__ long long ws[2][kSvNT / kWave];
  const int64_t t = blockIdx.x;
  for (int b = threadIdx.x; b < nblk; b += kSvNT) h[b] = 0;
  __syncthreads();
  const int64_t e0 = t * kSvTile, e1 = min<int64_t>(nnz, e0 + kSvTile);
  long long bad = 0, desc = 0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kSvNT) {
    const int64_t r = ir[e];
    if (r < 0 || r >= nrows) ++bad;
    else atomicAdd(&h[r >> kSvLB], 1);
    if (e + 1 < nnz) desc += ir[e + 1] <= r;
  }
  bad = wave_sum_ll(bad);
  desc = wave_sum_ll(desc);
  if (lane_id() == 0) {
    ws[0][wave_id()] = bad;
    ws[1][wave_id()] = desc;
  }
  __syncthreads();  // (also orders the histogram before it is read)
  if (threadIdx.x == 0) {
    long long a = 0, d = 0;
    for (int w = 0; w < kSvNT / kWave; ++w) {
      a += ws[0][w];
      d += ws[1][w];
    }
    part[2 * t] = a;
    part[2 * t + 1] = d;
  }
  for (int b = threadIdx.x; b < nblk; b += kSvNT) hist[(int64_t)b * ntiles + t] = h[b];
}

// The structure check's verdict: *malformed = 1 when a column violates the
// rules (cnt[0], from k_check_columns), a row is out of range, or the
// descents differ from those at column starts (cnt[2]).  One workgroup.
__global__ __launch_bounds__(kSvNT) void k_svb_verdict(const unsigned long long* __restrict__ cnt,
                                                       const long long* __restrict__ part, int64_t ntiles,
                                                       int* __restrict__ malformed) {
  __shared__ long long ws[2][kSvNT / kWave];
  long long bad = 0, desc = 0;
  for (int64_t t = threadIdx.x; t < ntiles; t += kSvNT) {
    bad += part[2 * t];
    desc += part[2 * t + 1];
  }
  bad = wave_sum_ll(bad);
  desc = wave_sum_ll(desc);
  if (lane_id() == 0) {
    ws[0][wave_id()] = bad;
    ws[1][wave_id()] = desc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long a = 0, d = 0;
    for (int w = 0; w < kSvNT / kWave; ++w) {
      a += ws[0][w];
      d += ws[1][w];
    }
    *malformed = (cnt[0] != 0 || a != 0 || (unsigned long long)d != cnt[2]) ? 1 : 0;
  }
}

// largest q in [lo, hi] with c[q] <= e (c ascending, c[lo] <= e)
template <class P>
__device__ __forceinline__ int64_t last_le(const P* c, int64_t lo, int64_t hi, int64_t e) {
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (c[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// pass 2: products into their blocks' buckets (row offset in the block, value)
template <int K, class T, bool BOOLA>
__global__ __launch_bounds__(kSvNT) void k_svb_scatter(const int64_t* __restrict__ cp,
                                                       const int64_t* __restrict__ jc,
                                                       const int64_t* __restrict__ ir,
                                                       const void* __restrict__ valA_, int64_t nzc, int64_t nnz,
                                                       const T* __restrict__ x, int nblk, int64_t ntiles,
                                                       const int64_t* __restrict__ offs, uint16_t* __restrict__ brow,
                                                       T* __restrict__ bval, const int* __restrict__ malformed) {
  using SR = Sem<K, T>;
  __shared__ int h[kSvMaxBlocks];
  __shared__ int64_t scp[kSvTile + 1];  // cp of the columns the tile spans (columns are non-empty)
  __shared__ int64_t s_q[2];
  if (*malformed) return;  // the structure check failed: cp, jc or ir cannot be trusted as indices
  const int64_t t = blockIdx.x;
  const int64_t e0 = t * kSvTile, e1 = min<int64_t>(nnz, e0 + kSvTile);
  for (int b = threadIdx.x; b < nblk; b += kSvNT) h[b] = 0;
  if (threadIdx.x < 2) s_q[threadIdx.x] = last_le(cp, 0, nzc - 1, threadIdx.x == 0 ? e0 : e1 - 1);
  __syncthreads();
  const int64_t q0 = s_q[0], q1 = s_q[1];
  const int nq = (int)(q1 - q0 + 2);  // <= kSvTile + 1
  for (int i = threadIdx.x; i < nq; i += kSvNT) scp[i] = cp[q0 + i];
  __syncthreads();
  int64_t ql = 0;  // column of this thread's previous entry, relative to q0 (entries ascend)
  for (int64_t e = e0 + threadIdx.x; e < e1; e += kSvNT) {
    ql = last_le(scp, ql, nq - 2, e);
    const T xj = x[jc[q0 + ql]];
    T prod;
    if constexpr (BOOLA) prod = xj;
    else prod = (T)SR::out(SR::mul(((const T*)valA_)[e], xj));
    const int64_t r = ir[e];
    const int b = (int)(r >> kSvLB);
    const int64_t pos = offs[(int64_t)b * ntiles + t] + atomicAdd(&h[b], 1);
    brow[pos] = (uint16_t)(r & (kSvBR - 1));
    bval[pos] = prod;
  }
}

// chunk table: cstart[b] = chunks of the blocks before b (cstart[nblk] = all);
// one workgroup, kSvChunkNT threads with kSvMaxBlocks / kSvChunkNT blocks each
constexpr int kSvChunkNT = 1024, kSvChunkPT = kSvMaxBlocks / kSvChunkNT;
__global__ __launch_bounds__(kSvChunkNT) void k_svb_chunks(const int64_t* __restrict__ offs, int nblk,
                                                           int64_t ntiles, int64_t nnz,
                                                           int64_t* __restrict__ cstart) {
  __shared__ long long wsum[kSvChunkNT / kWave];
  int64_t c[kSvChunkPT];
  long long sum = 0;
#pragma unroll
  for (int k = 0; k < kSvChunkPT; ++k) {
    const int b = threadIdx.x * kSvChunkPT + k;
    c[k] = 0;
    if (b < nblk) {
      const int64_t lo = offs[(int64_t)b * ntiles], hi = b + 1 < nblk ? offs[(int64_t)(b + 1) * ntiles] : nnz;
      c[k] = (hi - lo + kSvChunk - 1) / kSvChunk;
    }
    sum += c[k];
  }
  long long total;
  long long ex = block_excl_scan_ll<kSvChunkNT>(sum, wsum, total);
#pragma unroll
  for (int k = 0; k < kSvChunkPT; ++k) {
    const int b = threadIdx.x * kSvChunkPT + k;
    if (b < nblk) cstart[b] = ex;
    ex += c[k];
  }
  if (threadIdx.x == 0) cstart[nblk] = total;
}

// pass 3: one workgroup per chunk of a block's bucket
template <int K, class T>
__global__ __launch_bounds__(kSvNT) void k_svb_accumulate(const int64_t* __restrict__ offs,
                                                          const int64_t* __restrict__ cstart, int nblk,
                                                          int64_t ntiles, int64_t nnz, int64_t m,
                                                          const uint16_t* __restrict__ brow,
                                                          const T* __restrict__ bval, T* __restrict__ y,
                                                          const int* __restrict__ malformed) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  __shared__ Acc acc[kSvBR];
  if (*malformed) return;  // y is left as it was
  const int64_t g = blockIdx.x;
  if (g >= cstart[nblk]) return;
  const int b = (int)last_le(cstart, 0, nblk - 1, g);  // blocks without chunks share cstart with the next
  const int64_t lo = offs[(int64_t)b * ntiles], hi = b + 1 < nblk ? offs[(int64_t)(b + 1) * ntiles] : nnz;
  const int64_t c = g - cstart[b], nch = cstart[b + 1] - cstart[b];
  const int64_t i0 = lo + c * kSvChunk, i1 = min<int64_t>(hi, i0 + kSvChunk);
  for (int r = threadIdx.x; r < kSvBR; r += kSvNT) acc[r] = SR::ident();
  __syncthreads();
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kSvNT) SR::acc(&acc[brow[i]], (Acc)bval[i]);
  __syncthreads();
  const int64_t row0 = (int64_t)b * kSvBR;
  using Bits = typename std::conditional<sizeof(Acc) == 8, unsigned long long, unsigned int>::type;
  const Bits id = __builtin_bit_cast(Bits, SR::ident());
  for (int r = threadIdx.x; r < kSvBR && row0 + r < m; r += kSvNT) {
    const Acc a = acc[r];
    if (nch == 1) {
      y[row0 + r] = (T)SR::out(SR::add((Acc)y[row0 + r], a));
    } else if (__builtin_bit_cast(Bits, a) != id) {  // the identity changes nothing
      GAcc<K, T>::acc(&y[row0 + r], (T)SR::out(a));
    }
  }
}

template <int K, class T>
static cbx_status spmv_dense_typed(const cbx_dcsc* A, bool boolA, const T* x, T* y, hipStream_t s) {
  if (A->nzc == 0) return CBX_OK;
  const int64_t nnz = A->nnz, m = A->m;
  const int64_t nblk64 = (m + kSvBR - 1) / kSvBR;
  if (nnz < kSvMinNnz || nblk64 > kSvMaxBlocks) {
    {  // the structure check first: the kernel indexes y by A's rows and x by its columns
      ScratchSet cs(s);
      unsigned long long* chk = nullptr;
      if (!cs.take(&chk, 3)) return set_error(CBX_ERR_NOMEM, "cbx_spmv_dense: workspace allocation failed");
      dcsc_check_launch(A, chk, s);
      const cbx_status cst = dcsc_check_result(chk, s);
      if (cst != CBX_OK) return cst;
    }
    if (boolA)
      k_spmv_dense<K, T, true><<<wave_grid(A->nzc), 256, 0, s>>>(A->cp, A->jc, A->ir, A->val, A->nzc, x, y);
    else
      k_spmv_dense<K, T, false><<<wave_grid(A->nzc), 256, 0, s>>>(A->cp, A->jc, A->ir, A->val, A->nzc, x, y);
    CBX_HIP(hipGetLastError());
    CBX_HIP(hipStreamSynchronize(s));
    return CBX_OK;
  }
  const int nblk = (int)nblk64;
  const int64_t ntiles = (nnz + kSvTile - 1) / kSvTile;
  const int64_t nh = (int64_t)nblk * ntiles;
  const int64_t max_chunks = nnz / kSvChunk + nblk;
  CBX_RETURN_IF(ntiles >= (1ll << 31) || max_chunks >= (1ll << 31) || nh >= (1ll << 31), CBX_ERR_LIMIT,
                "cbx_spmv_dense: matrix too large for one launch");
  int64_t *hist = nullptr, *offs = nullptr, *cstart = nullptr;
  uint16_t* brow = nullptr;
  T* bval = nullptr;
  void* tmp = nullptr;
  unsigned long long* chk = nullptr;  // structure-check counters, then the device-side verdict
  long long* part = nullptr;          // per-tile counts of the per-entry check
  size_t tb = 0;
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, hist, offs, (int)nh, s));
  ScratchSet ws(s);
  auto cleanup = [&]() { ws.release(); };
  if (!ws.take(&hist, nh) || !ws.take(&offs, nh) || !ws.take(&cstart, nblk + 1) || !ws.take(&brow, nnz) ||
      !ws.take(&bval, nnz) || !ws.take_bytes(&tmp, std::max<size_t>(tb, 16)) || !ws.take(&chk, 4) ||
      !ws.take(&part, 2 * ntiles)) {
    cleanup();
    return set_error(CBX_ERR_NOMEM, "cbx_spmv_dense: workspace allocation failed");
  }
  int* malformed = (int*)(chk + 3);
  // the structure check (the rules of cbx_dcsc_check) runs inside the
  // launches: its per-column half here, the per-entry half in the counting
  // pass, and the kernels that index by the structure wait for the verdict
  dcsc_check_columns_launch(A, chk, s);
  k_svb_count<<<(unsigned)ntiles, kSvNT, 0, s>>>(A->ir, nnz, m, nblk, ntiles, hist, part);
  k_svb_verdict<<<1, kSvNT, 0, s>>>(chk, part, ntiles, malformed);
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, hist, offs, (int)nh, s);
  if (e == hipSuccess) {
    if (boolA)
      k_svb_scatter<K, T, true><<<(unsigned)ntiles, kSvNT, 0, s>>>(A->cp, A->jc, A->ir, A->val, A->nzc, nnz, x, nblk,
                                                                 ntiles, offs, brow, bval, malformed);
    else
      k_svb_scatter<K, T, false><<<(unsigned)ntiles, kSvNT, 0, s>>>(A->cp, A->jc, A->ir, A->val, A->nzc, nnz, x,
                                                                  nblk, ntiles, offs, brow, bval, malformed);
    k_svb_chunks<<<1, kSvChunkNT, 0, s>>>(offs, nblk, ntiles, nnz, cstart);
    k_svb_accumulate<K, T><<<(unsigned)max_chunks, kSvNT, 0, s>>>(offs, cstart, nblk, ntiles, nnz, m, brow, bval, y,
                                                                 malformed);
    e = hipGetLastError();
  }
  int bad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, malformed, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return set_error(CBX_ERR_HIP, "cbx_spmv_dense: %s", hipGetErrorString(e));
  if (bad) return set_error(CBX_ERR_INVALID, "cbx_dcsc_check: malformed DCSC");
  return CBX_OK;
}

extern "C" cbx_status cbx_spmspv(const cbx_dcsc* A, const int32_t* xind, const void* xval, int32_t nx,
                                 cbx_semiring semiring, cbx_dtype xdtype, int32_t* yind, void* yval, int32_t* ny,
                                 void* stream) {
  CBX_RETURN_IF(!A || !ny || (nx > 0 && (!xind || !xval)) || !yind || !yval, CBX_ERR_INVALID,
                "cbx_spmspv: null argument");
  CBX_RETURN_IF(A->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_spmspv: expected a device matrix");
  CBX_RETURN_IF(A->m >= (1ll << 31) || A->n >= (1ll << 31), CBX_ERR_LIMIT, "cbx_spmspv: dimensions must be below 2^31");
  CBX_RETURN_IF(nx < 0 || nx > A->n, CBX_ERR_INVALID, "cbx_spmspv: bad vector length");
  const bool boolA = A->dtype == CBX_BOOL;
  CBX_RETURN_IF(!boolA && A->dtype != xdtype, CBX_ERR_UNSUPPORTED,
                "cbx_spmspv: A must be boolean or share the vector's dtype");
  cbx_status st = cbx_dcsc_check(A);  // the kernels index by A's rows and columns
  if (st != CBX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  switch (semiring * 8 + xdtype) {
#define CBX_SV(SRK, SRE, T, DT)                                                                                     \
  case SRE * 8 + DT:                                                                                                \
    return spmspv_typed<SRK, T>(A, boolA, xind, (const T*)xval, nx, yind, (T*)yval, ny, s);
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, int64_t, CBX_I64)
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, double, CBX_F64)
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, float, CBX_F32)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, int64_t, CBX_I64)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, double, CBX_F64)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, float, CBX_F32)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, int64_t, CBX_I64)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, double, CBX_F64)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, float, CBX_F32)
#undef CBX_SV
  }
  return set_error(CBX_ERR_UNSUPPORTED, "cbx_spmspv: semiring %d with vector dtype %d is not provided", (int)semiring,
                   (int)xdtype);
}

// ------------------------------------------------ allocation-free SpMSpV
// The scratch of cbx_spmspv_noalloc: one static array per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not device
// memory, so the host never passes it to a kernel or a copy: each device's
// address of the array is looked up once (hipGetSymbolAddress, with that
// device current) and kept in g_scratch_dev.
constexpr size_t kSpmspvScratchBytes = (size_t)CBX_SPMSPV_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kSpmspvScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call made with that device current).
cbx_status scratch_region(char** out) {
  int dev = 0;
  CBX_HIP(hipGetDevice(&dev));
  CBX_RETURN_IF(dev < 0 || dev >= kMaxScratchDevices, CBX_ERR_LIMIT, "cbx_spmspv_noalloc: device %d out of range",
                dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(CBX_ERR_HIP, "cbx_spmspv_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                       hipGetErrorString(e));
    g_scratch_dev[dev] = (char*)p;
  }
  *out = g_scratch_dev[dev];
  return CBX_OK;
}
}  // namespace

extern "C" cbx_status cbx_spmspv_noalloc(const cbx_dcsc* A, const int32_t* xind, const void* xval, int32_t nx,
                                         cbx_semiring semiring, cbx_dtype xdtype, int32_t* yind, void* yval,
                                         int32_t* ny, void* stream) {
  CBX_RETURN_IF(!A || !ny || (nx > 0 && (!xind || !xval)) || !yind || !yval, CBX_ERR_INVALID,
                "cbx_spmspv_noalloc: null argument");
  CBX_RETURN_IF(A->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_spmspv_noalloc: expected a device matrix");
  CBX_RETURN_IF(A->m < 0 || A->n < 0 || A->nnz < 0 || A->nzc < 0, CBX_ERR_INVALID, "cbx_spmspv_noalloc: negative size");
  CBX_RETURN_IF(A->m >= (1ll << 31) || A->n >= (1ll << 31), CBX_ERR_LIMIT,
                "cbx_spmspv_noalloc: dimensions must be below 2^31");
  CBX_RETURN_IF(nx < 0 || nx > A->n, CBX_ERR_INVALID, "cbx_spmspv_noalloc: bad vector length");
  const bool boolA = A->dtype == CBX_BOOL;
  CBX_RETURN_IF(!boolA && A->dtype != xdtype, CBX_ERR_UNSUPPORTED,
                "cbx_spmspv_noalloc: A must be boolean or share the vector's dtype");
  char* region = nullptr;
  const cbx_status rs = scratch_region(&region);
  if (rs != CBX_OK) return rs;
  hipStream_t s = (hipStream_t)stream;
  switch (semiring * 8 + xdtype) {
#define CBX_SV(SRK, SRE, T, DT)                                                                                      \
  case SRE * 8 + DT:                                                                                                 \
    return spmspv_in_region<SRK, T>(A, boolA, xind, (const T*)xval, nx, yind, (T*)yval, ny, region,                  \
                                    kSpmspvScratchBytes, s);
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, int64_t, CBX_I64)
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, double, CBX_F64)
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, float, CBX_F32)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, int64_t, CBX_I64)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, double, CBX_F64)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, float, CBX_F32)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, int64_t, CBX_I64)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, double, CBX_F64)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, float, CBX_F32)
#undef CBX_SV
  }
  return set_error(CBX_ERR_UNSUPPORTED, "cbx_spmspv_noalloc: semiring %d with vector dtype %d is not provided",
                   (int)semiring, (int)xdtype);
}

extern "C" cbx_status cbx_spmv_dense(const cbx_dcsc* A, const void* x, void* y, cbx_semiring semiring,
                                     cbx_dtype xdtype, void* stream) {
  CBX_RETURN_IF(!A || (A->n > 0 && !x) || (A->m > 0 && !y), CBX_ERR_INVALID, "cbx_spmv_dense: null argument");
  CBX_RETURN_IF(A->location != CBX_DEVICE, CBX_ERR_INVALID, "cbx_spmv_dense: expected a device matrix");
  CBX_RETURN_IF(A->m >= (1ll << 31) || A->n >= (1ll << 31), CBX_ERR_LIMIT,
                "cbx_spmv_dense: dimensions must be below 2^31");
  const bool boolA = A->dtype == CBX_BOOL;
  CBX_RETURN_IF(!boolA && A->dtype != xdtype, CBX_ERR_UNSUPPORTED,
                "cbx_spmv_dense: A must be boolean or share the vectors' dtype");
  // (the structure check, the rules of cbx_dcsc_check, runs with the product)
  CBX_RETURN_IF(A->nzc == 0 && A->nnz != 0, CBX_ERR_INVALID, "cbx_dcsc_check: nnz without columns");
  const cbx_status init = init_device_once();
  if (init != CBX_OK) return init;
  hipStream_t s = (hipStream_t)stream;
  switch (semiring * 8 + xdtype) {
#define CBX_SV(SRK, SRE, T, DT) \
  case SRE * 8 + DT:            \
    return spmv_dense_typed<SRK, T>(A, boolA, (const T*)x, (T*)y, s);
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, int64_t, CBX_I64)
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, double, CBX_F64)
    CBX_SV(kPlusTimes, CBX_SR_PLUS_TIMES, float, CBX_F32)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, int64_t, CBX_I64)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, double, CBX_F64)
    CBX_SV(kMinPlus, CBX_SR_MIN_PLUS, float, CBX_F32)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, int64_t, CBX_I64)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, double, CBX_F64)
    CBX_SV(kSelectMax, CBX_SR_SELECT_MAX, float, CBX_F32)
#undef CBX_SV
  }
  return set_error(CBX_ERR_UNSUPPORTED, "cbx_spmv_dense: semiring %d with vector dtype %d is not provided",
                   (int)semiring, (int)xdtype);
}
