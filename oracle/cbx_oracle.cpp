// ORACLE -- test infrastructure only.
//
// A CPU restatement of CombBLAS's node-local kernels, used by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg, only to compare
// results against.
// Nothing in the product (combblas_amd/, include/) may call it.
//
// Every function cites the reference code it restates (paths relative to the
// CombBLAS tree).  The arithmetic order is kept wherever it decides floating
// point results: the heap branch uses std::make_heap/pop_heap/push_heap with
// the reference's swapped comparison (HeapEntry.h:50-53, a min-heap on the
// row key) so ties pop in the same order; the hash branch uses the same
// table size rule, hash (key*107 & (size-1)) and linear probing, so slot
// order and accumulation order are the same; the final sort compares keys
// only (sort_less, mtSpGEMM.h:205-209).
//
// Pinning (what fixes each function independently of this code):
//   oracle/_ref (the reference headers compiled as they lie) on generated
//   inputs for every semiring and both modes; brute-force dense products on
//   tiny inputs; the closed form nnz(C) = sum_j nnz_j; flop_j closed form;
//   see tests/test_oracle.py.
#include <omp.h>

#include <algorithm>
#include <random>
#include <functional>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../include/cbx.h"

namespace orc {

// ------------------------------------------------------------ semirings
// Semirings.h:212-232 PlusTimesSRing, :235-255 MinPlusSRing (with inf_plus
// :40-47), :165-190 SelectMaxSRing; bool PlusTimes through promote_trait.
template <class T>
struct PlusTimes {
  static T add(const T& a, const T& b) { return a + b; }
  static T multiply(const T& a, const T& b) { return a * b; }
};
template <>
struct PlusTimes<uint8_t> {  // bool + bool -> OR, bool * bool -> AND
  static uint8_t add(const uint8_t& a, const uint8_t& b) { return (uint8_t)((a + b) != 0); }
  static uint8_t multiply(const uint8_t& a, const uint8_t& b) { return (uint8_t)((a != 0) && (b != 0)); }
};
template <>
struct PlusTimes<int64_t> {  // wrap like the reference's int64 arithmetic on x86
  static int64_t add(const int64_t& a, const int64_t& b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
  static int64_t multiply(const int64_t& a, const int64_t& b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
};
template <class T>
struct MinPlus {
  static T add(const T& a, const T& b) { return std::min(a, b); }
  static T multiply(const T& a, const T& b) {
    const T inf = std::numeric_limits<T>::max();
    if (a == inf || b == inf) return inf;
    return a + b;
  }
};
template <>
struct MinPlus<int64_t> {
  static int64_t add(const int64_t& a, const int64_t& b) { return std::min(a, b); }
  static int64_t multiply(const int64_t& a, const int64_t& b) {
    const int64_t inf = std::numeric_limits<int64_t>::max();
    if (a == inf || b == inf) return inf;
    return (int64_t)((uint64_t)a + (uint64_t)b);
  }
};
template <class T>
struct SelectMax {
  static T add(const T& a, const T& b) { return std::max(a, b); }
  static T multiply(const T& a, const T& b) { return a * b; }
};
template <>
struct SelectMax<int64_t> {
  static int64_t add(const int64_t& a, const int64_t& b) { return std::max(a, b); }
  static int64_t multiply(const int64_t& a, const int64_t& b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
};

// ------------------------------------------------------------ A lookup
// Dcsc::FillColInds (dcsc.cpp:1361-1423) returns, for each row index k of a
// B column, the range [cp[p], cp[p+1]) of A's stored column p with jc[p]==k,
// or an empty range.  A dense map from column id to p gives the same ranges.
struct ColMap {
  std::vector<int64_t> pos;
  explicit ColMap(const cbx_dcsc& A) : pos(A.n, -1) {
    for (int64_t p = 0; p < A.nzc; ++p) pos[A.jc[p]] = p;
  }
  inline std::pair<int64_t, int64_t> range(const cbx_dcsc& A, int64_t k) const {
    const int64_t p = pos[k];
    if (p < 0) return {0, 0};
    return {A.cp[p], A.cp[p + 1]};
  }
};

// estimateFLOP (mtSpGEMM.h:1057-1134) for B's stored columns [c0, c1).
static void estimate_flop(const cbx_dcsc& A, const cbx_dcsc& B, const ColMap& cm, int64_t c0, int64_t c1,
                          int64_t* flop) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = c0; i < c1; ++i) {
    int64_t f = 0;
    for (int64_t e = B.cp[i]; e < B.cp[i + 1]; ++e) {
      auto r = cm.range(A, B.ir[e]);
      f += r.second - r.first;
    }
    flop[i - c0] = f;
  }
}

// estimateNNZ_Hash (mtSpGEMM.h:806-933): linear-probing table of
// 2^k >= max(16, flop_i) slots, hash = key*107 & (size-1), -1 marks empty.
static void estimate_nnz_hash(const cbx_dcsc& A, const cbx_dcsc& B, const ColMap& cm, int64_t c0, int64_t c1,
                              const int64_t* flop, int64_t* nnz) {
#pragma omp parallel
  {
    std::vector<int64_t> table;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = c0; i < c1; ++i) {
      int64_t ht = 16;
      while (ht < flop[i - c0]) ht <<= 1;
      if ((int64_t)table.size() < ht) table.resize(ht);
      std::fill(table.begin(), table.begin() + ht, (int64_t)-1);
      int64_t cnt = 0;
      for (int64_t e = B.cp[i]; e < B.cp[i + 1]; ++e) {
        auto r = cm.range(A, B.ir[e]);
        for (int64_t k = r.first; k < r.second; ++k) {
          const int64_t key = A.ir[k];
          int64_t h = (key * 107) & (ht - 1);
          while (true) {
            if (table[h] == key) break;
            if (table[h] == -1) {
              table[h] = key;
              ++cnt;
              break;
            }
            h = (h + 1) & (ht - 1);
          }
        }
      }
      nnz[i - c0] = cnt;
    }
  }
}

// HeapEntry (HeapEntry.h:36-56): ordering is reversed so that the standard
// max-heap algorithms keep the smallest row on top.
template <class T>
struct HeapItem {
  int64_t key;
  int64_t runr;
  T num;
};
template <class T>
static inline bool heap_less(const HeapItem<T>& a, const HeapItem<T>& b) {
  return a.key > b.key;
}

struct Entry {
  int64_t row;
};

// One output column of LocalHybridSpGEMM (mtSpGEMM.h:289-441) or of
// LocalSpGEMMHash (mtSpGEMM.h:556-646), written to rows/vals at `out`.
template <class SR, class T>
static void numeric_column(const cbx_dcsc& A, const cbx_dcsc& B, const ColMap& cm, int64_t i, int64_t flop_i,
                           int64_t nnz_i, bool allow_heap, bool sort, int64_t* rows, T* vals,
                           std::vector<std::pair<int64_t, int64_t>>& colinds, std::vector<HeapItem<T>>& heap,
                           std::vector<std::pair<int64_t, T>>& table) {
  const T* Av = (const T*)A.val;
  const T* Bv = (const T*)B.val;
  const int64_t b0 = B.cp[i];
  const int64_t nb = B.cp[i + 1] - b0;
  if ((int64_t)colinds.size() < nb) colinds.resize(nb);
  for (int64_t j = 0; j < nb; ++j) colinds[j] = cm.range(A, B.ir[b0 + j]);
  const double cr = (double)flop_i / (double)nnz_i;
  int64_t cur = 0;
  if (allow_heap && cr < 2.0) {  // heap merge, mtSpGEMM.h:311-360
    if ((int64_t)heap.size() < nb) heap.resize(nb);
    int64_t hsize = 0;
    for (int64_t j = 0; j < nb; ++j)
      if (colinds[j].first != colinds[j].second)
        heap[hsize++] = HeapItem<T>{A.ir[colinds[j].first], j, Av[colinds[j].first]};
    std::make_heap(heap.data(), heap.data() + hsize, heap_less<T>);
    while (hsize > 0) {
      std::pop_heap(heap.data(), heap.data() + hsize, heap_less<T>);
      HeapItem<T>& top = heap[hsize - 1];
      const int64_t locb = top.runr;
      const T mrhs = SR::multiply(top.num, Bv[b0 + locb]);
      if (cur > 0 && rows[cur - 1] == top.key) {
        vals[cur - 1] = SR::add(vals[cur - 1], mrhs);
      } else {
        rows[cur] = top.key;
        vals[cur] = mrhs;
        ++cur;
      }
      if (++colinds[locb].first != colinds[locb].second) {
        top.key = A.ir[colinds[locb].first];
        top.num = Av[colinds[locb].first];
        std::push_heap(heap.data(), heap.data() + hsize, heap_less<T>);
      } else {
        --hsize;
      }
    }
    return;
  }
  // hash, mtSpGEMM.h:362-440 (and :556-646 for LocalSpGEMMHash)
  int64_t ht = 16;
  while (ht < nnz_i) ht <<= 1;
  if ((int64_t)table.size() < ht) table.resize(ht);
  for (int64_t s = 0; s < ht; ++s) table[s].first = -1;
  for (int64_t j = 0; j < nb; ++j) {
    const T bval = Bv[b0 + j];
    for (int64_t k = colinds[j].first; k < colinds[j].second; ++k) {
      const T mrhs = SR::multiply(Av[k], bval);
      const int64_t key = A.ir[k];
      int64_t h = (key * 107) & (ht - 1);
      while (true) {
        if (table[h].first == key) {
          table[h].second = SR::add(mrhs, table[h].second);
          break;
        } else if (table[h].first == -1) {
          table[h].first = key;
          table[h].second = mrhs;
          break;
        }
        h = (h + 1) & (ht - 1);
      }
    }
  }
  if (sort) {
    int64_t w = 0;
    for (int64_t s = 0; s < ht; ++s)
      if (table[s].first != -1) table[w++] = table[s];
    std::sort(table.begin(), table.begin() + w,
              [](const std::pair<int64_t, T>& a, const std::pair<int64_t, T>& b) { return a.first < b.first; });
    for (int64_t s = 0; s < w; ++s) {
      rows[cur] = table[s].first;
      vals[cur] = table[s].second;
      ++cur;
    }
  } else {
    for (int64_t s = 0; s < ht; ++s)
      if (table[s].first != -1) {
        rows[cur] = table[s].first;
        vals[cur] = table[s].second;
        ++cur;
      }
  }
}

// The whole local product restricted to B's stored columns [c0, c1), as
// LocalHybridSpGEMM (allow_heap=true, sort=true) or LocalSpGEMMHash
// (allow_heap=false).  Output: host DCSC with the non-empty columns, like
// SpDCCols(const SpTuples&, false) (SpDCCols.cpp:109-184).
template <class SR, class T>
static cbx_status spgemm_cols(const cbx_dcsc& A, const cbx_dcsc& B, int64_t c0, int64_t c1, bool allow_heap,
                              bool sort, cbx_dcsc* C, int64_t* flops_out) {
  std::memset(C, 0, sizeof(*C));
  C->m = A.m;
  C->n = B.n;
  C->dtype = A.dtype;
  C->location = CBX_HOST;
  const int64_t nc = c1 - c0;
  ColMap cm(A);
  std::vector<int64_t> flop(nc), nnz(nc), ptr(nc + 1, 0);
  estimate_flop(A, B, cm, c0, c1, flop.data());
  estimate_nnz_hash(A, B, cm, c0, c1, flop.data(), nnz.data());
  int64_t total_flop = 0;
  for (int64_t i = 0; i < nc; ++i) {
    ptr[i + 1] = ptr[i] + nnz[i];
    total_flop += flop[i];
  }
  if (flops_out) *flops_out = total_flop;
  const int64_t nnzC = ptr[nc];
  int64_t nzc = 0;
  for (int64_t i = 0; i < nc; ++i) nzc += nnz[i] > 0;
  C->nnz = nnzC;
  C->nzc = nzc;
  C->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nzc + 1));
  C->jc = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(nzc, 1));
  C->ir = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(nnzC, 1));
  C->val = std::malloc(sizeof(T) * std::max<int64_t>(nnzC, 1));
  if (!C->cp || !C->jc || !C->ir || !C->val) return CBX_ERR_NOMEM;
  int64_t q = 0;
  for (int64_t i = 0; i < nc; ++i)
    if (nnz[i] > 0) {
      C->jc[q] = B.jc[c0 + i];
      C->cp[q] = ptr[i];
      ++q;
    }
  C->cp[nzc] = nnzC;
  T* vals = (T*)C->val;
#pragma omp parallel
  {
    std::vector<std::pair<int64_t, int64_t>> colinds;
    std::vector<HeapItem<T>> heap;
    std::vector<std::pair<int64_t, T>> table;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < nc; ++i) {
      if (nnz[i] == 0) continue;
      numeric_column<SR, T>(A, B, cm, c0 + i, flop[i], nnz[i], allow_heap, sort, C->ir + ptr[i], vals + ptr[i],
                            colinds, heap, table);
    }
  }
  return CBX_OK;
}

template <template <class> class SRT>
static cbx_status dispatch_dtype(const cbx_dcsc& A, const cbx_dcsc& B, int64_t c0, int64_t c1, bool heap, bool sort,
                                 cbx_dcsc* C, int64_t* flops) {
  switch (A.dtype) {
    case CBX_I64: return spgemm_cols<SRT<int64_t>, int64_t>(A, B, c0, c1, heap, sort, C, flops);
    case CBX_F64: return spgemm_cols<SRT<double>, double>(A, B, c0, c1, heap, sort, C, flops);
    case CBX_F32: return spgemm_cols<SRT<float>, float>(A, B, c0, c1, heap, sort, C, flops);
    default: return CBX_ERR_UNSUPPORTED;
  }
}

// ------------------------------------------------------------ SpMSpV
// SpImpl<SR,IT,NUM,IVT,OVT>::SpMXSpV, heap form for sizeof(NUM) <= sizeof(OVT)
// (SpImpl.cpp:110-157): k-way merge of the columns of A selected by x with a
// HeapEntry min-heap; equal rows fold with SR::add(previous, new).
template <class SR, class T>
static void spmxspv_heap(const cbx_dcsc& A, const ColMap& cm, const int32_t* xind, const T* xval, int32_t nx,
                         std::vector<int32_t>& yind, std::vector<T>& yval) {
  const T* Av = (const T*)A.val;
  std::vector<std::pair<int64_t, int64_t>> colinds(nx);
  for (int32_t j = 0; j < nx; ++j) colinds[j] = cm.range(A, xind[j]);
  std::vector<HeapItem<T>> heap(nx);
  int64_t hsize = 0;
  for (int32_t j = 0; j < nx; ++j)
    if (colinds[j].first != colinds[j].second)
      heap[hsize++] = HeapItem<T>{A.ir[colinds[j].first], j, Av[colinds[j].first]};
  std::make_heap(heap.data(), heap.data() + hsize, heap_less<T>);
  while (hsize > 0) {
    std::pop_heap(heap.data(), heap.data() + hsize, heap_less<T>);
    HeapItem<T>& top = heap[hsize - 1];
    const int64_t locv = top.runr;
    const T mrhs = SR::multiply(top.num, xval[locv]);
    if (!yind.empty() && yind.back() == top.key) yval.back() = SR::add(yval.back(), mrhs);
    else {
      yind.push_back((int32_t)top.key);
      yval.push_back(mrhs);
    }
    if (++colinds[locv].first != colinds[locv].second) {
      top.key = A.ir[colinds[locv].first];
      top.num = Av[colinds[locv].first];
      std::push_heap(heap.data(), heap.data() + hsize, heap_less<T>);
    } else {
      --hsize;
    }
  }
}

// SpImpl<SR,IT,bool,IVT,OVT>::SpMXSpV (SpImpl.cpp:167-213): a boolean A
// contributes x_k itself; per row the contributions fold with SR::add in
// ascending-row order.  The reference orders equal rows through a sequence
// heap; add is commutative for every supported semiring, and min/max fold
// exactly, so a dense accumulator reproduces it (FP plus-times up to order).
template <class SR, class T>
static void spmxspv_bool(const cbx_dcsc& A, const ColMap& cm, const int32_t* xind, const T* xval, int32_t nx,
                         std::vector<int32_t>& yind, std::vector<T>& yval, std::vector<T>& spa,
                         std::vector<uint8_t>& seen, std::vector<int32_t>& touched) {
  touched.clear();
  for (int32_t j = 0; j < nx; ++j) {
    auto r = cm.range(A, xind[j]);
    for (int64_t q = r.first; q < r.second; ++q) {
      const int32_t i = (int32_t)A.ir[q];
      if (!seen[i]) {
        seen[i] = 1;
        spa[i] = xval[j];
        touched.push_back(i);
      } else {
        spa[i] = SR::add(spa[i], xval[j]);
      }
    }
  }
  std::sort(touched.begin(), touched.end());
  yind.assign(touched.begin(), touched.end());
  yval.resize(touched.size());
  for (size_t t = 0; t < touched.size(); ++t) {
    yval[t] = spa[touched[t]];
    seen[touched[t]] = 0;
  }
}

// ------------------------------------------------- MCL prune / recover / select
// MCLPruneRecoverySelect (ParFriends.h:185-345) on one process, restated per
// column.  Column statistics follow SpParMat::Reduce(Column, plus, 0)
// (SpParMat.cpp:961-1010): a sequential sum in the value type, in stored
// order, `s = v + s`; counts are accumulated in the value type too.
// Kselect1 (SpParMat.cpp:1413-1690): the k-th largest entry of the
// unpruned column if it has >= k entries, else its smallest entry, and
// numeric_limits<T>::min() for an empty column.  PruneColumn with std::less
// (dcsc.cpp:759-838) keeps v with !(v < threshold) and drops columns left
// empty, keeping the order of the survivors.
template <class T>
static T kth_largest(const T* v, int64_t len, int64_t k, std::vector<T>& buf) {
  if (len == 0) return std::numeric_limits<T>::min();
  buf.assign(v, v + len);
  std::sort(buf.begin(), buf.end(), std::greater<T>());
  return len >= k ? buf[k - 1] : buf[len - 1];
}

template <class T>
static cbx_status mcl_prune_typed(const cbx_dcsc& A, T thr, int64_t select, int64_t recover, T pct,
                                  cbx_dcsc* out, T* cut_out) {
  const int64_t nzc = A.nzc;
  const T* val = (const T*)A.val;
  std::vector<T> cut(nzc);
  std::vector<int64_t> keep(nzc + 1, 0);
#pragma omp parallel
  {
    std::vector<T> buf;
#pragma omp for schedule(dynamic, 64)
    for (int64_t q = 0; q < nzc; ++q) {
      const int64_t b = A.cp[q], e = A.cp[q + 1], len = e - b;
      const T* v = val + b;
      T nP = 0, sP = 0;  // Reduce over PrunedA (values <= thr removed)
      for (int64_t i = 0; i < len; ++i)
        if (!(v[i] <= thr)) { nP = T(1) + nP; sP = v[i] + sP; }
      const T nU = (T)len;
      T c = thr;
      const bool rec = nP < (T)recover && nU > nP && sP < pct;
      if (rec) {
        c = kth_largest(v, len, recover, buf);
      } else if (select > 0 && nP > (T)select) {
        c = kth_largest(v, len, select, buf);
        if (recover > 0) {
          T n1 = 0, s1 = 0;
          for (int64_t i = 0; i < len; ++i)
            if (!(v[i] < c)) { n1 = T(1) + n1; s1 = v[i] + s1; }
          if (n1 < (T)recover && s1 < pct) c = kth_largest(v, len, recover, buf);
        }
      }
      cut[q] = c;
      int64_t k = 0;
      for (int64_t i = 0; i < len; ++i) k += !(v[i] < c);
      keep[q + 1] = k;
    }
  }
  if (cut_out) std::copy(cut.begin(), cut.end(), cut_out);
  int64_t onzc = 0;
  for (int64_t q = 0; q < nzc; ++q) onzc += keep[q + 1] > 0;
  for (int64_t q = 0; q < nzc; ++q) keep[q + 1] += keep[q];
  const int64_t onnz = keep[nzc];
  std::memset(out, 0, sizeof(*out));
  out->m = A.m; out->n = A.n; out->nnz = onnz; out->nzc = onzc;
  out->dtype = A.dtype; out->location = CBX_HOST;
  out->cp = (int64_t*)std::malloc(sizeof(int64_t) * (onzc + 1));
  out->jc = (int64_t*)std::malloc(sizeof(int64_t) * (onzc ? onzc : 1));
  out->ir = (int64_t*)std::malloc(sizeof(int64_t) * (onnz ? onnz : 1));
  T* ov = (T*)std::malloc(sizeof(T) * (onnz ? onnz : 1));
  out->val = ov;
  int64_t c = 0;
  for (int64_t q = 0; q < nzc; ++q) {
    if (keep[q + 1] == keep[q]) continue;
    out->jc[c] = A.jc[q];
    out->cp[c] = keep[q];
    ++c;
  }
  out->cp[onzc] = onnz;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < nzc; ++q) {
    int64_t w = keep[q];
    for (int64_t p = A.cp[q]; p < A.cp[q + 1]; ++p)
      if (!(val[p] < cut[q])) { out->ir[w] = A.ir[p]; ov[w] = val[p]; ++w; }
  }
  return CBX_OK;
}

}  // namespace orc

using namespace orc;

extern "C" {

void orc_free(cbx_dcsc* C);

// LocalHybridSpGEMM (mode 0) / LocalSpGEMMHash sorted (mode 1) / unsorted
// (mode 2) on B's stored columns [c0, c1) (c1 < 0 means all).
cbx_status orc_spgemm(const cbx_dcsc* A, const cbx_dcsc* B, int32_t semiring, int32_t mode, int64_t c0, int64_t c1,
                      cbx_dcsc* C, int64_t* flops) {
  if (!A || !B || !C || A->n != B->m || A->dtype != B->dtype) return CBX_ERR_INVALID;
  if (c1 < 0 || c1 > B->nzc) c1 = B->nzc;
  if (c0 < 0 || c0 > c1) return CBX_ERR_INVALID;
  const bool heap = mode == 0;
  const bool sort = mode != 2;
  switch (semiring) {
    case CBX_SR_PLUS_TIMES:
      if (A->dtype == CBX_BOOL)
        return spgemm_cols<PlusTimes<uint8_t>, uint8_t>(*A, *B, c0, c1, heap, sort, C, flops);
      return dispatch_dtype<PlusTimes>(*A, *B, c0, c1, heap, sort, C, flops);
    case CBX_SR_MIN_PLUS: return dispatch_dtype<MinPlus>(*A, *B, c0, c1, heap, sort, C, flops);
    case CBX_SR_SELECT_MAX: return dispatch_dtype<SelectMax>(*A, *B, c0, c1, heap, sort, C, flops);
  }
  return CBX_ERR_UNSUPPORTED;
}

// estimateFLOP and estimateNNZ_Hash for every stored column of B.
cbx_status orc_symbolic(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* flop, int64_t* nnz) {
  if (!A || !B || A->n != B->m) return CBX_ERR_INVALID;
  ColMap cm(*A);
  estimate_flop(*A, *B, cm, 0, B->nzc, flop);
  if (nnz) estimate_nnz_hash(*A, *B, cm, 0, B->nzc, flop, nnz);
  return CBX_OK;
}

// Times the restated LocalHybridSpGEMM on B's stored columns [c0, c1) with
// the current OpenMP thread count; the output is produced, then freed.
cbx_status orc_time_spgemm(const cbx_dcsc* A, const cbx_dcsc* B, int32_t semiring, int32_t mode, int64_t c0,
                           int64_t c1, double* secs, int64_t* nnz, int64_t* flops) {
  cbx_dcsc C;
  const double t0 = omp_get_wtime();
  const cbx_status st = orc_spgemm(A, B, semiring, mode, c0, c1, &C, flops);
  const double t1 = omp_get_wtime();
  if (st != CBX_OK) return st;
  *secs = t1 - t0;
  *nnz = C.nnz;
  orc_free(&C);
  return CBX_OK;
}

// The first `count` draws of the sample stream estimateNNZ_sampling uses
// (mtSpGEMM.h:968-979): std::exponential_distribution<float>(1) over a
// default-seeded std::default_random_engine.
void orc_exp_samples(int64_t count, float* out) {
  std::default_random_engine gen;
  std::exponential_distribution<float> exp_dist(1.0f);
  for (int64_t i = 0; i < count; ++i) out[i] = exp_dist(gen);
}

// estimateNNZ_sampling (mtSpGEMM.h:936-1054): Cohen's min-hash size estimate
// with nrounds exponential samples per row of A.
//  * samples_init[i], i < m*nrounds: the i-th draw of
//    std::exponential_distribution<float>(1) from a default-seeded
//    std::default_random_engine (mtSpGEMM.h:968-979; the loop is an
//    `omp parallel for` nested in an `omp parallel` region whose threads each
//    own a default-seeded engine, so with nested parallelism inactive -- the
//    OpenMP default -- every thread writes the same sequential stream).  The
//    standard library of this toolchain (libstdc++, GCC 11) draws them, as it
//    does for the reference: minstd_rand0 (x <- 16807 x mod 2^31-1, x0 = 1),
//    u = generate_canonical<float,24> = float(x - 1) / 2^31, -log(1 - u).
//  * samples_mid[c][k] = min over the rows of A(:,c) of samples_init[row][k]
//    (:991-1005), FLT_MAX for an empty column.
//  * per stored column i of B: samples_final[k] = min over B(:,i)'s rows r of
//    samples_mid[r][k] (:1025-1035); colest[i] = float(nrounds-1) /
//    (samples_final[0] + ... + samples_final[nrounds-1]) summed in k order in
//    float (:1037-1040).
//  * the estimate: colest summed in float in column order (the reference's
//    `reduction(+:nnzest)` is that order on one thread), cast to int64.
// colest (B.nzc floats) may be null; *total receives the float sum.
cbx_status orc_estimate_sampling(const cbx_dcsc* A, const cbx_dcsc* B, int32_t nrounds, float* colest,
                                 float* total, int64_t* estimate) {
  if (!A || !B || A->n != B->m || nrounds < 1) return CBX_ERR_INVALID;
  if (estimate) *estimate = 0;
  if (total) *total = 0.0f;
  if (A->nnz == 0 || B->nnz == 0) {
    if (colest) std::fill(colest, colest + B->nzc, 0.0f);
    return CBX_OK;
  }
  const int64_t m = A->m, p = A->n, r = nrounds;
  std::vector<float> init((size_t)(m * r));
  {
    std::default_random_engine gen;
    std::exponential_distribution<float> exp_dist(1.0f);
    for (int64_t i = 0; i < m * r; ++i) init[i] = exp_dist(gen);
  }
  std::vector<float> mid((size_t)(p * r), std::numeric_limits<float>::max());
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t q = 0; q < A->nzc; ++q) {
    float* dst = mid.data() + A->jc[q] * r;
    for (int64_t e = A->cp[q]; e < A->cp[q + 1]; ++e) {
      const float* src = init.data() + A->ir[e] * r;
      for (int64_t k = 0; k < r; ++k)
        if (src[k] < dst[k]) dst[k] = src[k];
    }
  }
  std::vector<float> est((size_t)B->nzc);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < B->nzc; ++i) {
    std::vector<float> fin((size_t)r, std::numeric_limits<float>::max());
    for (int64_t e = B->cp[i]; e < B->cp[i + 1]; ++e) {
      const float* src = mid.data() + B->ir[e] * r;
      for (int64_t k = 0; k < r; ++k)
        if (src[k] < fin[k]) fin[k] = src[k];
    }
    float c = 0.0f;
    for (int64_t k = 0; k < r; ++k) c += fin[k];
    est[i] = static_cast<float>(r - 1) / c;
  }
  float sum = 0.0f;
  for (int64_t i = 0; i < B->nzc; ++i) sum += est[i];
  if (colest) std::copy(est.begin(), est.end(), colest);
  if (total) *total = sum;
  if (estimate) *estimate = static_cast<int64_t>(sum);
  return CBX_OK;
}

// y = A x (SpMXSpV).  A boolean A -> boolean form; else heap form.  Host
// arrays; yind/yval must hold A.m entries; returns the number of entries.
int32_t orc_spmspv(const cbx_dcsc* A, const int32_t* xind, const void* xval, int32_t nx, int32_t semiring,
                   int32_t xdtype, int32_t* yind, void* yval) {
  ColMap cm(*A);
  std::vector<int32_t> yi;
  int32_t ny = -1;
  auto run = [&](auto* tag, auto sr) {
    using T = std::remove_pointer_t<decltype(tag)>;
    using SR = decltype(sr);
    std::vector<T> yv;
    if (A->dtype == CBX_BOOL) {
      std::vector<T> spa(A->m);
      std::vector<uint8_t> seen(A->m, 0);
      std::vector<int32_t> touched;
      spmxspv_bool<SR, T>(*A, cm, xind, (const T*)xval, nx, yi, yv, spa, seen, touched);
    } else {
      spmxspv_heap<SR, T>(*A, cm, xind, (const T*)xval, nx, yi, yv);
    }
    std::copy(yi.begin(), yi.end(), yind);
    std::copy(yv.begin(), yv.end(), (T*)yval);
    ny = (int32_t)yi.size();
  };
  switch (semiring * 8 + xdtype) {
    case CBX_SR_PLUS_TIMES * 8 + CBX_I64: run((int64_t*)nullptr, PlusTimes<int64_t>()); break;
    case CBX_SR_PLUS_TIMES * 8 + CBX_F64: run((double*)nullptr, PlusTimes<double>()); break;
    case CBX_SR_PLUS_TIMES * 8 + CBX_F32: run((float*)nullptr, PlusTimes<float>()); break;
    case CBX_SR_MIN_PLUS * 8 + CBX_I64: run((int64_t*)nullptr, MinPlus<int64_t>()); break;
    case CBX_SR_MIN_PLUS * 8 + CBX_F64: run((double*)nullptr, MinPlus<double>()); break;
    case CBX_SR_MIN_PLUS * 8 + CBX_F32: run((float*)nullptr, MinPlus<float>()); break;
    case CBX_SR_SELECT_MAX * 8 + CBX_I64: run((int64_t*)nullptr, SelectMax<int64_t>()); break;
    case CBX_SR_SELECT_MAX * 8 + CBX_F64: run((double*)nullptr, SelectMax<double>()); break;
    case CBX_SR_SELECT_MAX * 8 + CBX_F32: run((float*)nullptr, SelectMax<float>()); break;
  }
  return ny;
}

// Top-down BFS, the level loop of TopDownBFS.cpp:442-448 on one process:
// fringe.setNumToInd(); fringe = SpMV<SelectMaxSRing<bool,int64>>(A, fringe);
// fringe = EWiseMult(fringe, parents, true, -1) (ParFriends.h:2700, keep
// entries whose parent is -1); parents.Set(fringe) (FullyDistVec.cpp:1036).
// Returns the number of iterations; parents (host int64[n]) are overwritten.
// cut_out (optional, nzc entries of the value type): the per-column threshold
// PruneColumn applied (ParFriends.h:339).
int32_t orc_mcl_prune(const cbx_dcsc* A, double thr, int64_t select, int64_t recover, double pct, cbx_dcsc* out,
                      void* cut_out) {
  if (A->dtype == CBX_F32)
    return orc::mcl_prune_typed<float>(*A, (float)thr, select, recover, (float)pct, out, (float*)cut_out);
  if (A->dtype == CBX_F64)
    return orc::mcl_prune_typed<double>(*A, thr, select, recover, pct, out, (double*)cut_out);
  return CBX_ERR_UNSUPPORTED;
}

// Galerkin coarse operator of the 27-point stencil under 2x2x2 aggregation,
// summed directly: (R A R^T)(I,J) = sum over fine i in I, j in J of A(i,j),
// with A(j,j) = 26 and A(i,j) = -1 for the up to 26 grid neighbours i of j.
// This is what the product S*(A*T) of GalerkinNew.cpp:105-106 computes,
// obtained without any SpGEMM; every value is an integer, so FP64 is exact.
int32_t orc_galerkin_direct(int32_t g, cbx_dcsc* out) {
  const int gc = (g + 1) / 2;
  const int64_t nc = (int64_t)gc * gc * gc;
  std::vector<double> vals(nc * 27, 0.0);
  std::vector<uint8_t> hit(nc * 27, 0);
#pragma omp parallel for schedule(static)
  for (int64_t J = 0; J < nc; ++J) {
    const int X = (int)(J % gc), Y = (int)((J / gc) % gc), Z = (int)(J / ((int64_t)gc * gc));
    for (int jz = 2 * Z; jz < std::min(2 * Z + 2, g); ++jz)
      for (int jy = 2 * Y; jy < std::min(2 * Y + 2, g); ++jy)
        for (int jx = 2 * X; jx < std::min(2 * X + 2, g); ++jx)
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                const int ix = jx + dx, iy = jy + dy, iz = jz + dz;
                if (ix < 0 || iy < 0 || iz < 0 || ix >= g || iy >= g || iz >= g) continue;
                const int oz = iz / 2 - Z + 1, oy = iy / 2 - Y + 1, ox = ix / 2 - X + 1;
                const int slot = ox + 3 * (oy + 3 * oz);
                vals[J * 27 + slot] += (dx | dy | dz) ? -1.0 : 26.0;
                hit[J * 27 + slot] = 1;
              }
  }
  int64_t nnz = 0;
  for (size_t i = 0; i < hit.size(); ++i) nnz += hit[i];
  std::memset(out, 0, sizeof(*out));
  out->m = nc; out->n = nc; out->nnz = nnz; out->nzc = nc;
  out->dtype = CBX_F64; out->location = CBX_HOST;
  out->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nc + 1));
  out->jc = (int64_t*)std::malloc(sizeof(int64_t) * (nc ? nc : 1));
  out->ir = (int64_t*)std::malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  double* v = (double*)std::malloc(sizeof(double) * (nnz ? nnz : 1));
  out->val = v;
  int64_t w = 0;
  for (int64_t J = 0; J < nc; ++J) {
    const int X = (int)(J % gc), Y = (int)((J / gc) % gc), Z = (int)(J / ((int64_t)gc * gc));
    out->jc[J] = J;
    out->cp[J] = w;
    for (int slot = 0; slot < 27; ++slot) {  // slot order is ascending coarse row order
      if (!hit[J * 27 + slot]) continue;
      const int ox = slot % 3 - 1, oy = (slot / 3) % 3 - 1, oz = slot / 9 - 1;
      out->ir[w] = (X + ox) + (int64_t)gc * ((Y + oy) + (int64_t)gc * (Z + oz));
      v[w] = vals[J * 27 + slot];
      ++w;
    }
  }
  out->cp[nc] = nnz;
  return CBX_OK;
}

// dcsc_gespmv (Friends.h:63-77): for every stored column in order, every entry
// in order, y[row] = add(y[row], multiply(value, x[col])); a boolean A uses
// x[col] itself (SelectMaxSRing<bool,T>::multiply of true and x).
int32_t orc_spmv_dense(const cbx_dcsc* A, const void* x, void* y, int32_t semiring, int32_t xdtype) {
  auto run = [&](auto* tag, auto sr) {
    using T = std::remove_pointer_t<decltype(tag)>;
    using SR = decltype(sr);
    const T* xv = (const T*)x;
    T* yv = (T*)y;
    for (int64_t q = 0; q < A->nzc; ++q) {
      const T xj = xv[A->jc[q]];
      for (int64_t e = A->cp[q]; e < A->cp[q + 1]; ++e) {
        const T prod = A->dtype == CBX_BOOL ? xj : SR::multiply(((const T*)A->val)[e], xj);
        yv[A->ir[e]] = SR::add(yv[A->ir[e]], prod);
      }
    }
  };
  switch (semiring * 8 + xdtype) {
    case CBX_SR_PLUS_TIMES * 8 + CBX_I64: run((int64_t*)nullptr, PlusTimes<int64_t>()); return 0;
    case CBX_SR_PLUS_TIMES * 8 + CBX_F64: run((double*)nullptr, PlusTimes<double>()); return 0;
    case CBX_SR_PLUS_TIMES * 8 + CBX_F32: run((float*)nullptr, PlusTimes<float>()); return 0;
    case CBX_SR_MIN_PLUS * 8 + CBX_I64: run((int64_t*)nullptr, MinPlus<int64_t>()); return 0;
    case CBX_SR_MIN_PLUS * 8 + CBX_F64: run((double*)nullptr, MinPlus<double>()); return 0;
    case CBX_SR_MIN_PLUS * 8 + CBX_F32: run((float*)nullptr, MinPlus<float>()); return 0;
    case CBX_SR_SELECT_MAX * 8 + CBX_I64: run((int64_t*)nullptr, SelectMax<int64_t>()); return 0;
    case CBX_SR_SELECT_MAX * 8 + CBX_F64: run((double*)nullptr, SelectMax<double>()); return 0;
    case CBX_SR_SELECT_MAX * 8 + CBX_F32: run((float*)nullptr, SelectMax<float>()); return 0;
  }
  return CBX_ERR_UNSUPPORTED;
}

int32_t orc_bfs(const cbx_dcsc* A, int64_t root, int64_t* parents) {
  const int64_t n = A->n;
  std::fill(parents, parents + n, (int64_t)-1);
  ColMap cm(*A);
  std::vector<int32_t> find(1, (int32_t)root), yind, kept;
  std::vector<int64_t> fnum, ynum, spa(n);
  std::vector<uint8_t> seen(n, 0);
  std::vector<int32_t> touched;
  int32_t iterations = 0;
  while (!find.empty()) {
    fnum.assign(find.begin(), find.end());  // setNumToInd
    spmxspv_bool<SelectMax<int64_t>, int64_t>(*A, cm, find.data(), fnum.data(), (int32_t)find.size(), yind, ynum,
                                              spa, seen, touched);
    kept.clear();
    for (size_t t = 0; t < yind.size(); ++t)
      if (parents[yind[t]] == -1) {  // EWiseMult(..., exclude=true, -1)
        kept.push_back(yind[t]);
        parents[yind[t]] = ynum[t];  // Set
      }
    find.swap(kept);
    ++iterations;
  }
  return iterations;
}

void orc_free(cbx_dcsc* C) {
  if (!C) return;
  std::free(C->cp);
  std::free(C->jc);
  std::free(C->ir);
  std::free(C->val);
  std::memset(C, 0, sizeof(*C));
}

int32_t orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int32_t n) { omp_set_num_threads(n); }

}  // extern "C"
