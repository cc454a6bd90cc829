// Block-level flattening of segment products (shared by SpGEMM and SpMSpV).
//
// A chunk holds at most one segment per thread (an A column restricted to
// some rows, paired with a multiplier).  stage_chunk compacts the non-empty
// segments with one packed 64-bit block scan; visit_staged / fill_round then
// give every thread U consecutive products per round, found with one binary
// search and single steps forward (segments are non-empty).
#pragma once

#include "cbx_device.h"

namespace cbx {

// Per-chunk segment staging for block-level flattening of products.
template <class T, int NT>
struct SegStage {
  int off[NT + 1];
  int64_t base[NT];
  T bval[NT];
};

// Stages one chunk (at most one segment per thread): the non-empty segments
// are compacted in order with one packed 64-bit block scan (high half counts
// segments, low half sums lengths), so staged offsets strictly increase.
// Returns the number of products; `nseg` receives the number of segments.
// Contains barriers.
template <class T, int NT>
__device__ __forceinline__ int stage_chunk(SegStage<T, NT>& st, long long* wsum, int64_t base, int len, T bv,
                                           int& nseg) {
  const long long packed = len > 0 ? ((1ll << 32) | (long long)(unsigned)len) : 0ll;
  long long tot;
  const long long ex = block_excl_scan_ll<NT>(packed, wsum, tot);
  if (len > 0) {
    const int k = (int)(ex >> 32);
    st.off[k] = (int)(ex & 0xffffffffll);
    st.base[k] = base;
    st.bval[k] = bv;
  }
  nseg = (int)(tot >> 32);
  const int total = (int)(tot & 0xffffffffll);
  if (threadIdx.x == 0) st.off[nseg] = total;
  __syncthreads();
  return total;
}

// Visits all `tot` staged products.  Each thread takes U consecutive
// products per round: one binary search for the first, then (segments being
// non-empty) at most one step forward per product.  Loads are issued for all
// U products before f runs.  f(idx, bv) gets idx = -1 for lanes past the end.
// No barriers.
// Products xb .. xb+U-1 of a staged chunk: A-entry index (-1 past the end)
// and B value of each.
template <class T, int NT, int U>
__device__ __forceinline__ void fill_round(const SegStage<T, NT>& st, int nseg, int tot, int xb, int64_t (&idx)[U],
                                           T (&bvs)[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) idx[u] = -1;
  if (xb >= tot) return;
  int lo = 0, hi = nseg;  // largest k with off[k] <= xb (off[nseg] = tot > xb)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (st.off[mid] <= xb) lo = mid;
    else hi = mid;
  }
  int k = lo;
  int cur = st.off[k], nxt = st.off[k + 1];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int x = xb + u;
    if (x < tot) {
      if (x >= nxt) {
        ++k;
        cur = nxt;
        nxt = st.off[k + 1];
      }
      idx[u] = st.base[k] + (x - cur);
      bvs[u] = st.bval[k];
    }
  }
}

template <class T, int NT, int U, class F>
__device__ __forceinline__ void visit_staged(const SegStage<T, NT>& st, int nseg, int tot, F&& f, int first_round = 0) {
  for (int xb = (first_round * NT + (int)threadIdx.x) * U; xb < tot; xb += NT * U) {
    int64_t idx[U];
    T bvs[U];
    fill_round<T, NT, U>(st, nseg, tot, xb, idx, bvs);
    f(idx, bvs);
  }
}

}  // namespace cbx
