// Local semiring SpGEMM C = A*B on DCSC blocks for gfx950.
//
// Replaces CombBLAS's LocalHybridSpGEMM<SR,NTO> (include/CombBLAS/mtSpGEMM.h:213-460)
// and LocalSpGEMMHash (mtSpGEMM.h:463-656).  The reference runs, per
// non-empty column j of B:
//   estimateFLOP     flop_j = sum_{k in B(:,j)} nnz(A(:,k))      (mtSpGEMM.h:1057-1134)
//   estimateNNZ_Hash nnz_j  = distinct rows of the products       (mtSpGEMM.h:806-933)
//   prefixsum        column pointers of C                          (mtSpGEMM.h:23-70)
//   numeric          heap merge or hash + sort, rows ascending     (mtSpGEMM.h:289-441)
// This file keeps those phases but maps them to the GPU as follows.
//
// Column tiers (by flop_j, after sorting columns by flop descending):
//   hash tier  flop_j <= small_flop_max: one workgroup per column with an LDS
//              linear-probing hash table of 2^ceil(log2(2*size)) slots, then
//              in-LDS compaction and a bitonic sort of the surviving rows.
//   dense tier the rest: one workgroup per column walks the rows in tiles of
//              TR rows; each tile accumulates into a dense LDS array with a
//              presence bitmap and is emitted in row order by a block scan
//              of the bitmap, so no sort is needed.  The A-entries of a tile
//              are found through a precomputed tile-pointer table
//              P[a][s] = #entries of A's a-th stored column with row < s*TR.
// Within every wave, the A-column segments of up to 64 B entries are
// flattened with a prefix sum so that all 64 lanes work on products even
// when most segments are short (R-MAT columns are).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <type_traits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cbx_device.h"
#include "cbx_flatten.h"
#include "cbx_semiring.h"
#include "cbx_spgemm.h"

namespace cbx {

// ------------------------------------------------------------------ utils
template <class T>
__device__ __forceinline__ T shfl_any(T v, int src) {
  if constexpr (sizeof(T) == 8) {
    long long x = __builtin_bit_cast(long long, v);
    x = __shfl(x, src, kWave);
    return __builtin_bit_cast(T, x);
  } else if constexpr (sizeof(T) == 4) {
    int x = __builtin_bit_cast(int, v);
    x = __shfl(x, src, kWave);
    return __builtin_bit_cast(T, x);
  } else {
    int x = (int)v;
    x = __shfl(x, src, kWave);
    return (T)x;
  }
}

// Visits the A entries of up to 64 segments, one segment per lane
// ([lo, lo+len) with multiplier bv).  All lanes of the wave must call it.
// f(idx, bv) runs once per entry; entries are spread over all lanes.
template <class T, class F>
__device__ __forceinline__ void wave_visit(int64_t lo, int len, T bv, F&& f) {
  const int lane = lane_id();
  const int incl = wave_incl_scan(len);
  const int total = __shfl(incl, kWave - 1, kWave);
  const int excl = incl - len;
  for (int tb = 0; tb < total; tb += kWave) {
    const int t = tb + lane;
    int L = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      const int cand = L + step;
      const int ex = __shfl(excl, cand, kWave);
      if (ex <= t) L = cand;
    }
    const int64_t loL = shfl_any(lo, L);
    const int exL = __shfl(excl, L, kWave);
    const T bvL = shfl_any(bv, L);
    if (t < total) f(loL + (int64_t)(t - exL), bvL);
  }
}

__device__ __forceinline__ uint32_t hash_slot(int32_t r, int lg) {
  return ((uint32_t)r * 2654435761u) >> (32 - lg);
}

__device__ __forceinline__ int ceil_log2_dev(long long x) {
  int lg = 0;
  while ((1ll << lg) < x) ++lg;
  return lg;
}

// ============================================================ unit tier
// Row-range units.  The symbolic pass builds, per column, the full presence
// bitmap of its output rows in LDS, counts the rows of every block of kBR
// rows, and cuts the column into units: runs of consecutive blocks holding
// at most KMAX outputs and spanning at most RBLK blocks.  The numeric pass
// runs one workgroup per unit.  A unit's accumulator is compact: the value of
// row r lives at rank(r) = number of output rows of the unit below r, read
// from the unit's presence bitmap with per-word prefix counts, so the unit
// is emitted in row order without sorting and LDS holds only real outputs.
constexpr int kBR = 4096;        // rows per block (also the tile-pointer granularity)
constexpr int kBRW = kBR / 32;   // 32-bit bitmap words per block
// Row lists (k_sym_units) cover the blocks with at most this many outputs:
// sixteen denser blocks already fill a 4096-output unit, so a unit holding a
// denser block spans at most 16 blocks and takes its ranks from a bitmap.
constexpr int kListBlockMax = 256;

struct Unit {
  int32_t col;     // B stored-column position
  int32_t blk_lo;  // first block
  int32_t blk_hi;  // one past the last block
  int32_t cnt;     // outputs in the unit
  int32_t off;     // offset of the unit's first output within the column
  int32_t slot;    // the column's row-map slot (its workgroup in the symbolic launch)
};

// The unit of workgroup b (of n) when runs of g consecutive units go to one
// XCD.  The hardware deals a launch's workgroups to the 8 XCDs in turn
// (workgroup b runs on XCD b % 8), so with units sorted by row block the
// identity mapping spreads every row band over all eight L2 caches; here,
// within each super-group of 8g workgroups, XCD x takes units x*g .. x*g+g-1,
// and the runs still interleave across the XCDs.  The tail after the last
// full super-group keeps the identity; g = 0 is the identity.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_unit(int b, int n, int g) {
  if (g <= 0) return b;
  const int sg = kXcds * g;
  if (b >= n / sg * sg) return b;
  const int r = b % sg;
  return b - r + (r % kXcds) * g + r / kXcds;
}

// Fused threshold of the MCL expansion (MCLPruneRecoverySelect's hard
// threshold, ParFriends.h:185-345, when no column can need recovery): with
// `on`, the numeric kernels emit only the entries with !(v < thr) of every
// output range they own, compacted at the range's start, and append the
// range (start, kept count, column) to a run list; the host code then gathers
// the runs into the product.  The epilogue on that product keeps exactly what
// it keeps on the unfiltered one (its statistics count only entries > thr and
// its cuts are never below thr), and the product's writes shrink to the kept
// entries.
template <class T>
struct Filter {
  int on = 0;
  T thr{};
  int64_t* run_src = nullptr;
  int32_t* run_kept = nullptr;
  int32_t* run_col = nullptr;
  int* run_count = nullptr;
  int run_cap = 0;
  // entries the staged tiny tier left out (its kept counts size the columns,
  // so the product's unfiltered size is the scan's total plus these)
  unsigned long long* dropped = nullptr;
};

template <class T>
__device__ __forceinline__ void filter_add_run(const Filter<T>& f, int64_t src, int kept, int col, int* err) {
  if (kept <= 0) return;
  const int k = atomicAdd(f.run_count, 1);
  if (k < f.run_cap) {
    f.run_src[k] = src;
    f.run_kept[k] = kept;
    f.run_col[k] = col;
  } else {
    atomicOr(err, 1024);
  }
}

// Segment of B entry e restricted to P tiles [t_lo, t_hi) (or the whole A
// column when `full`).
template <class T, bool VAL>
__device__ __forceinline__ void load_seg(int64_t e, int64_t b1, const int32_t* __restrict__ bpos,
                                         const int64_t* __restrict__ cpA, const int32_t* __restrict__ P, int S,
                                         int t_lo, int t_hi, bool full, const T* __restrict__ valB, int64_t& base,
                                         int& len, T& bv) {
  base = 0;
  len = 0;
  bv = T{};
  if (e < b1) {
    const int32_t p = bpos[e];
    if (p >= 0) {
      if (full) {
        base = cpA[p];
        len = (int)(cpA[p + 1] - base);
      } else {
        const int32_t* pp = P + (int64_t)p * (S + 1);
        const int a0 = pp[t_lo];
        base = cpA[p] + a0;
        len = pp[t_hi] - a0;
      }
      if constexpr (VAL) bv = valB[e];
    }
  }
}

// ------------------------------------------------------------------ setup
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_iota_i32(int32_t* p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}

// colmap[jcA[p]] = p  (dense lookup replacing Dcsc::ConstructAux/FillColInds,
// dcsc.cpp:1042-1070 and :1361-1423)
__global__ void k_colmap(const int64_t* __restrict__ jc, int64_t nzc, int32_t* __restrict__ colmap) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < nzc) colmap[jc[p]] = (int32_t)p;
}

__global__ void k_rows32(const int64_t* __restrict__ ir, int64_t nnz, int32_t* __restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < nnz) out[i] = (int32_t)ir[i];
}

// bpos[e] = stored-column position in A of B's row index irB[e] (-1 if A(:,k) is empty).
__global__ void k_bpos(const int64_t* __restrict__ irB, int64_t nnzB, const int32_t* __restrict__ colmap,
                       int32_t* __restrict__ bpos) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < nnzB) bpos[e] = colmap[irB[e]];
}

// flop_j (estimateFLOP, mtSpGEMM.h:1112-1130): one wave per column of B.
__global__ void k_flop(const int64_t* __restrict__ cpB, int64_t nzcB, const int32_t* __restrict__ bpos,
                       const int64_t* __restrict__ cpA, int64_t* __restrict__ flop) {
  const int64_t col = global_wave();
  if (col >= nzcB) return;
  const int lane = lane_id();
  long long f = 0;
  for (int64_t e = cpB[col] + lane; e < cpB[col + 1]; e += kWave) {
    const int32_t p = bpos[e];
    if (p >= 0) f += cpA[p + 1] - cpA[p];
  }
  f = wave_sum_ll(f);
  if (lane == 0) flop[col] = f;
}

// Tile pointers: P[p*(S+1)+s] = number of entries of A's p-th stored column
// with row < s*TR.  One thread per stored column.
__global__ void k_ptable(const int64_t* __restrict__ cpA, int64_t nzcA, const int32_t* __restrict__ rowsA,
                         int S, int TR, int32_t* __restrict__ P) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= nzcA) return;
  const int64_t b = cpA[p];
  const int len = (int)(cpA[p + 1] - b);
  int32_t* out = P + p * (int64_t)(S + 1);
  int t = 0;
  for (int i = 0; i < len; ++i) {
    const int tile = rowsA[b + i] / TR;
    while (t <= tile) out[t++] = i;
  }
  while (t <= S) out[t++] = len;
}

// Class boundaries over columns sorted by flop descending:
// bounds[c] = number of columns with flop > thr[c].
__global__ void k_class_bounds(const int64_t* __restrict__ flop_sorted, int64_t n, const int64_t* thr, int nthr,
                               int64_t* bounds) {
  const int c = threadIdx.x;
  if (c >= nthr) return;
  int64_t lo = 0, hi = n;  // first index with flop <= thr[c]
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (flop_sorted[mid] > thr[c]) lo = mid + 1;
    else hi = mid;
  }
  bounds[c] = lo;
}

__global__ void k_cp_finish(int64_t* nnzcol, int64_t nzcB) {
  if (threadIdx.x == 0 && blockIdx.x == 0) nnzcol[nzcB] = 0;
}

// sum of nnzcol over order[0..n) (outputs of the unit-tier columns)
__global__ void k_sum_ordered(const int32_t* __restrict__ order, int64_t n, const int64_t* __restrict__ nnzcol,
                              unsigned long long* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  long long v = i < n ? nnzcol[order[i]] : 0;
  v = wave_sum_ll(v);
  if (lane_id() == 0 && v) atomicAdd(out, (unsigned long long)v);
}

__global__ void k_nonempty_flags(const int64_t* __restrict__ nnzcol, int64_t n, uint8_t* __restrict__ flags) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) flags[i] = nnzcol[i] > 0;
}

__global__ void k_set_last(int64_t* cp, const int64_t* nsel, const int64_t* total_src, int64_t total_idx) {
  if (threadIdx.x == 0 && blockIdx.x == 0) cp[*nsel] = total_src[total_idx];
}

// Input validation (the reference's preconditions, mtSpGEMM.h / dcsc.h):
// cp starts at 0, ends at nnz and never decreases (empty columns allowed); jc
// strictly increasing and below ncols; rows of every column strictly
// increasing and below nrows.  Two fully parallel passes: per entry, rows in
// range and every adjacent descent ir[e+1] <= ir[e] counted; per column, cp
// and jc, and the descents at the positions where a column begins (each
// distinct start counted once).  With cp valid those are the only places a
// descent is allowed, so the rows ascend within columns iff the two counts
// agree.  cnt[0] violations, cnt[1] descents, cnt[2] descents at starts;
// k_validate_flag then sets bit `flagbit` of *err.
constexpr int kValNT = 256, kValBlocks = 1024;

__global__ __launch_bounds__(kValNT) void k_validate_cols(const int64_t* __restrict__ cp,
                                                          const int64_t* __restrict__ jc,
                                                          const int64_t* __restrict__ ir, int64_t nzc, int64_t nnz,
                                                          int64_t ncols, unsigned long long* __restrict__ cnt) {
  long long bad = 0, desc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nzc; q += stride) {
    const int64_t b = cp[q], e = cp[q + 1], j = jc[q];
    bool err = b < 0 || e < b || e > nnz || j < 0 || j >= ncols;
    if (q + 1 < nzc) err |= jc[q + 1] <= j;
    if (q == 0) err |= b != 0;
    if (q == nzc - 1) err |= e != nnz;
    bad += err;
    if (!err && q > 0 && b > 0 && b < nnz && cp[q - 1] < b) desc += ir[b] <= ir[b - 1];
  }
  block_add_counts<kValNT>(bad, desc, &cnt[0], &cnt[2]);
}

__global__ __launch_bounds__(kValNT) void k_validate_entries(const int64_t* __restrict__ ir, int64_t nnz,
                                                             int64_t nrows, unsigned long long* __restrict__ cnt) {
  long long bad = 0, desc = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += stride) {
    const int64_t r = ir[e];
    bad += r < 0 || r >= nrows;
    if (e + 1 < nnz) desc += ir[e + 1] <= r;
  }
  block_add_counts<kValNT>(bad, desc, &cnt[0], &cnt[1]);
}

__global__ void k_validate_flag(const unsigned long long* __restrict__ cnt, int flagbit, int* __restrict__ err) {
  if (threadIdx.x == 0 && (cnt[0] || cnt[1] != cnt[2])) atomicOr(err, 1 << flagbit);
}

// Launches the validation of X on s; cnt: 3 device counters (cleared here).
static cbx_status launch_validate(const cbx_dcsc* X, int flagbit, unsigned long long* cnt, int* err, hipStream_t s) {
  auto blocks = [](int64_t n) {
    return (unsigned)std::min<int64_t>(std::max<int64_t>((n + kValNT - 1) / kValNT, 1), kValBlocks);
  };
  CBX_HIP(hipMemsetAsync(cnt, 0, 3 * sizeof(unsigned long long), s));
  k_validate_cols<<<blocks(X->nzc), kValNT, 0, s>>>(X->cp, X->jc, X->ir, X->nzc, X->nnz, X->n, cnt);
  k_validate_entries<<<blocks(X->nnz), kValNT, 0, s>>>(X->ir, X->nnz, X->m, cnt);
  k_validate_flag<<<1, 64, 0, s>>>(cnt, flagbit, err);
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}

// --------------------------------------------------------- symbolic: hash
// estimateNNZ_Hash restated for LDS (mtSpGEMM.h:861-927): insert each product's
// row into a linear-probing table, count first insertions.
template <int HMAX, int NT>
__global__ __launch_bounds__(NT) void k_sym_hash(const int32_t* __restrict__ order, int64_t first,
                                                 const int64_t* __restrict__ flop, const int64_t* __restrict__ cpB,
                                                 const int32_t* __restrict__ bpos, const int64_t* __restrict__ cpA,
                                                 const int32_t* __restrict__ rowsA, int64_t* __restrict__ nnzcol,
                                                 int* __restrict__ err) {
  __shared__ int32_t keys[HMAX];
  __shared__ long long wsum[NT / kWave];
  __shared__ SegStage<uint8_t, NT> st;
  const int col = order[first + blockIdx.x];
  const int lg = min(ceil_log2_dev(max(2 * flop[col], 64ll)), ceil_log2_dev(HMAX));
  const int H = 1 << lg;
  for (int i = threadIdx.x; i < H; i += NT) keys[i] = -1;
  __syncthreads();
  long long cnt = 0;
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  for (int64_t eb = b0; eb < b1; eb += NT) {
    int64_t base;
    int len, nseg;
    uint8_t dummy;
    load_seg<uint8_t, false>(eb + threadIdx.x, b1, bpos, cpA, nullptr, 0, 0, 0, true, nullptr, base, len, dummy);
    const int tot = stage_chunk<uint8_t, NT>(st, wsum, base, len, 0, nseg);
    visit_staged<uint8_t, NT, 4>(st, nseg, tot, [&](const int64_t(&idx)[4], const uint8_t(&)[4]) {
      int32_t rr[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) rr[u] = idx[u] >= 0 ? rowsA[idx[u]] : -1;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int32_t r = rr[u];
        if (r < 0) continue;
        uint32_t h = hash_slot(r, lg);
        int probe = 0;
        for (; probe < H; ++probe) {  // the table is at least twice the column's flops
          const int32_t old = atomicCAS(&keys[h], -1, r);
          if (old == -1) { ++cnt; break; }
          if (old == r) break;
          h = (h + 1) & (H - 1);
        }
        if (probe == H) atomicOr(err, 1);
      }
    });
  }
  const long long tot = block_sum_ll<NT>(cnt, wsum);
  if (threadIdx.x == 0) nnzcol[col] = tot;
}

// ------------------------------------------------------------ tiny tier
// Columns with at most 256 products (most columns of multigrid triple
// products, and the sparsest R-MAT columns): one wave per column, no LDS.
// Lane l holds elements e = u*64 + l (u < 4).  The products are gathered with
// a wave scan of the B entries' A-segment lengths and a 6-step binary search
// over it by shuffles; a bitonic network sorts the 256 (row, value) pairs in
// registers (partners in another lane through shuffles, in another register
// of the same lane by swaps); equal rows are then adjacent, a segmented scan
// in element order reduces each run, and the run's last element writes the
// output at its rank.  SYM only counts the distinct rows.
constexpr int kTinyMax = 256;

template <class X>
__device__ __forceinline__ X shfl_x(X v, int src) {
  if constexpr (sizeof(X) < 4) return (X)__shfl((int)v, src, kWave);
  else return __shfl(v, src, kWave);
}

// Sort, reduce and emit the first NR registers (NR * 64 elements) of a tiny
// column; elements beyond them are empty (row kNone).
template <int K, class T, bool SYM, int NR>
__device__ __forceinline__ void tiny_finish(int32_t (&key)[kTinyMax / kWave],
                                            typename Sem<K, T>::Acc (&val)[kTinyMax / kWave], int lane, int col,
                                            int64_t* __restrict__ nnzcol, const int64_t* __restrict__ cpC,
                                            int64_t* __restrict__ irC, T* __restrict__ valC, int* __restrict__ err,
                                            const int64_t* __restrict__ stage_off, int64_t slot,
                                            const Filter<T>& flt) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int32_t kNone = 0x7fffffff;
  // bitonic sort of NR*64 elements, ascending by row
#pragma unroll
  for (int k = 2; k <= NR * kWave; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= kWave) {
        const int du = j / kWave;
#pragma unroll
        for (int u = 0; u < NR; ++u) {
          const int u2 = u ^ du;
          if (u2 > u) {
            const bool asc = (((u * kWave + lane) & k) == 0);
            if ((key[u] > key[u2]) == asc && key[u] != key[u2]) {
              const int32_t tk = key[u];
              key[u] = key[u2];
              key[u2] = tk;
              if constexpr (!SYM) {
                const Acc tv = val[u];
                val[u] = val[u2];
                val[u2] = tv;
              }
            }
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < NR; ++u) {
          const int e = u * kWave + lane;
          const bool asc = ((e & k) == 0);
          const bool lower = (lane & j) == 0;
          const int32_t pk = __shfl_xor(key[u], j, kWave);
          Acc pv{};
          if constexpr (!SYM) pv = __shfl_xor(val[u], j, kWave);
          const bool take = lower == asc ? pk < key[u] : pk > key[u];
          if (take) {
            key[u] = pk;
            if constexpr (!SYM) val[u] = pv;
          }
        }
      }
    }
  }
  // run heads (first element of each distinct row)
  bool head[NR];
  unsigned long long hm[NR];
  int count = 0;
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    int32_t prev = __shfl_up(key[u], 1, kWave);
    const int32_t carry = u > 0 ? __shfl(key[u > 0 ? u - 1 : 0], kWave - 1, kWave) : kNone;
    if (lane == 0) prev = u > 0 ? carry : -1;
    head[u] = key[u] != kNone && key[u] != prev;
    hm[u] = __ballot(head[u]);
    count += __popcll(hm[u]);
  }
  if constexpr (SYM) {
    if (lane == 0) nnzcol[col] = count;
  } else {
    // the column's allocation from the symbolic pass; or, staged (no symbolic
    // pass), its slot of flop_j entries in the staging buffer, and the count
    // is recorded for the copy into place
    int64_t have, out;
    if (stage_off) {
      out = stage_off[slot];
      have = stage_off[slot + 1] - out;
      if (lane == 0) nnzcol[col] = count;
      if (lane == 0 && count > have) atomicOr(err, 8);
    } else {
      have = nnzcol[col];
      if (lane == 0 && count != have) atomicOr(err, 8);
      out = cpC[col];
    }
    Acc carry = SR::ident();
    bool open = false;  // a run continues from the previous register
    int before = 0;     // heads in earlier registers
    int kept = 0;       // entries written (filter on)
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      // segmented inclusive scan in element order
      Acc x = val[u];
      bool f = head[u];
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const Acc y = __shfl_up(x, d, kWave);
        const bool g = __shfl_up((int)f, d, kWave) != 0;
        if (lane >= d) {
          if (!f) x = SR::add(y, x);
          f = f || g;
        }
      }
      if (!f && open) x = SR::add(carry, x);
      // the last element of a run writes it at its rank
      int32_t next = __shfl_down(key[u], 1, kWave);
      const int32_t nfirst = u + 1 < NR ? __shfl(key[u + 1 < NR ? u + 1 : u], 0, kWave) : kNone;
      if (lane == kWave - 1) next = nfirst;
      const bool end = key[u] != kNone && next != key[u];
      if (!flt.on) {
        if (end) {
          const int rank = before + __popcll(hm[u] & ((2ull << lane) - 1ull)) - 1;
          if (rank < have) {
            irC[out + rank] = key[u];
            valC[out + rank] = SR::out(x);
          }
        }
      } else {  // the fused threshold: kept entries, in row order, at out
        const T v = SR::out(x);
        const bool keep = end && !(v < flt.thr);
        const uint64_t km = __ballot(keep);
        if (keep) {
          const int rank = kept + __popcll(km & ((1ull << lane) - 1ull));
          if (rank < have) {
            irC[out + rank] = key[u];
            valC[out + rank] = v;
          }
        }
        kept += __popcll(km);
      }
      carry = __shfl(x, kWave - 1, kWave);
      open = __shfl((int)(key[u] != kNone), kWave - 1, kWave) != 0;
      before += __popcll(hm[u]);
    }
    if (flt.on) {  // staged: the kept count sizes the column; else its range joins the runs
      if (stage_off) {
        if (lane == 0) {
          nnzcol[col] = kept;
          if (count > kept) atomicAdd(flt.dropped, (unsigned long long)(count - kept));
        }
      } else if (lane == 0) {
        filter_add_run(flt, out, kept, col, err);
      }
    }
  }
}

// The products of one tiny column (at most kTinyMax of them) gathered into
// the wave's registers: element e = u * 64 + lane sits in key[u] (its row)
// and val[u] (its product); elements past the column's products hold kNone.
// Returns the number of products (wave-uniform).  All lanes must call it.
template <int K, class T, bool SYM>
__device__ __forceinline__ int tiny_gather(int64_t b0, int64_t b1, const int32_t* __restrict__ bpos,
                                           const T* __restrict__ valB, const int64_t* __restrict__ cpA,
                                           const int32_t* __restrict__ rowsA, const T* __restrict__ valA,
                                           int32_t (&key)[kTinyMax / kWave],
                                           typename Sem<K, T>::Acc (&val)[kTinyMax / kWave]) {
  using SR = Sem<K, T>;
  constexpr int kE = kTinyMax / kWave;
  constexpr int32_t kNone = 0x7fffffff;
  const int lane = lane_id();
#pragma unroll
  for (int u = 0; u < kE; ++u) {
    key[u] = kNone;
    val[u] = SR::ident();
  }
  int filled = 0;  // wave-uniform
  for (int64_t eb = b0; eb < b1; eb += kWave) {
    int64_t base;
    int len;
    T bv;
    load_seg<T, !SYM>(eb + lane, b1, bpos, cpA, nullptr, 0, 0, 0, true, valB, base, len, bv);
    const int inc = wave_incl_scan(len);
    const int off = filled + inc - len;  // first element of this lane's segment
    const int total = __shfl(inc, kWave - 1, kWave);
    // lanes that hold B entries in this chunk, and the search's first step:
    // the largest power of two below that count (all wave-uniform)
    const int nb = (int)(b1 - eb < kWave ? b1 - eb : kWave);
    int top = 0;
    if (nb > 1) top = 1 << (31 - __clz(nb - 1));
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      // registers whose 64 elements miss this chunk's range stay as they are
      if ((u + 1) * kWave <= filled || u * kWave >= filled + total) continue;
      const int e = u * kWave + lane;
      // last lane s < nb with off_s <= e (segments of length 0 share
      // offsets; the search lands on the one that holds e when e is in this
      // chunk)
      int lo = 0;
      for (int step = top; step > 0; step >>= 1) {
        const int o = __shfl(off, lo + step, kWave);
        if (lo + step < nb && o <= e) lo += step;
      }
      const int o = __shfl(off, lo, kWave);
      const int l = __shfl(len, lo, kWave);
      const int64_t bs = shfl_x(base, lo);
      const T bval = shfl_x(bv, lo);
      if (e >= filled && e < filled + total && e - o < l) {
        const int64_t idx = bs + (e - o);
        key[u] = rowsA[idx];
        if constexpr (!SYM) val[u] = SR::mul(valA[idx], bval);
      }
    }
    filled += total;
  }
  return filled;
}

template <int K, class T, bool SYM>
__global__ __launch_bounds__(256) void k_tiny(const int32_t* __restrict__ order, int64_t first, int64_t ncols,
                                              const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                              const T* __restrict__ valB, const int64_t* __restrict__ cpA,
                                              const int32_t* __restrict__ rowsA, const T* __restrict__ valA,
                                              int64_t* __restrict__ nnzcol, const int64_t* __restrict__ cpC,
                                              int64_t* __restrict__ irC, T* __restrict__ valC, int* __restrict__ err,
                                              const int64_t* __restrict__ stage_off = nullptr, Filter<T> flt = {}) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int kE = kTinyMax / kWave;  // 4 elements per lane
  constexpr int32_t kNone = 0x7fffffff;
  const int lane = lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x / kWave) + wave_id(); i < ncols; i += waves) {
    const int col = order[first + i];
    const int64_t b0 = cpB[col], b1 = cpB[col + 1];
    int32_t key[kE];
    Acc val[kE];
    const int filled = tiny_gather<K, T, SYM>(b0, b1, bpos, valB, cpA, rowsA, valA, key, val);
    if (filled > kTinyMax && lane == 0) atomicOr(err, 4);
    // sort only as many registers as the column fills (wave-uniform)
    if (filled <= kWave) tiny_finish<K, T, SYM, 1>(key, val, lane, col, nnzcol, cpC, irC, valC, err, stage_off, i, flt);
    else if (filled <= 2 * kWave)
      tiny_finish<K, T, SYM, 2>(key, val, lane, col, nnzcol, cpC, irC, valC, err, stage_off, i, flt);
    else tiny_finish<K, T, SYM, 4>(key, val, lane, col, nnzcol, cpC, irC, valC, err, stage_off, i, flt);
  }
}

// LDS form of the tiny tier: one wave per column, no sort.  The column's
// products (gathered into registers as above) are inserted into a per-wave
// linear-probing table of kTinyHash slots keyed by row (at most half full),
// accumulating each product into its row's slot with the semiring's LDS
// atomic; the occupied slots are compacted, and each distinct row's rank (its
// position in the sorted column) is the number of distinct rows below it,
// counted against the compacted rows read back four at a time (every lane
// reads the same address).  Cross-lane traffic per column: a few dozen LDS
// instructions against the several hundred lane shuffles of the register
// bitonic sort.
constexpr int kTinyHash = 2 * kTinyMax;  // slots per wave
constexpr int kTinyWaves = 4;            // waves per workgroup at most (launch bound 256)

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int K, class T, bool SYM>
__global__ __launch_bounds__(256) void k_tiny_h(const int32_t* __restrict__ order, int64_t first, int64_t ncols,
                                                const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                const T* __restrict__ valB, const int64_t* __restrict__ cpA,
                                                const int32_t* __restrict__ rowsA, const T* __restrict__ valA,
                                                int64_t* __restrict__ nnzcol, const int64_t* __restrict__ cpC,
                                                int64_t* __restrict__ irC, T* __restrict__ valC,
                                                int* __restrict__ err, const int64_t* __restrict__ stage_off = nullptr,
                                                Filter<T> flt = {}) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int kE = kTinyMax / kWave;
  constexpr int kSPL = kTinyHash / kWave;  // table slots owned by a lane in the compaction
  constexpr int32_t kNone = 0x7fffffff, kEmpty = -1;
  constexpr int kLgHash = 9;
  static_assert((1 << kLgHash) == kTinyHash, "table size");
  __shared__ int32_t s_key[kTinyWaves][kTinyHash];
  __shared__ Acc s_val[SYM ? 1 : kTinyWaves][SYM ? 1 : kTinyHash];
  __shared__ __align__(16) int32_t s_row[kTinyWaves][kTinyMax + 4];  // compacted rows, kNone-padded
  __shared__ uint16_t s_slot[SYM ? 1 : kTinyWaves][SYM ? 1 : kTinyMax];
  const int lane = lane_id(), w = wave_id();
  int32_t* hk = s_key[w];
  int32_t* crow = s_row[w];
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / kWave);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x / kWave) + w; i < ncols; i += waves) {
    const int col = order[first + i];
    const int64_t b0 = cpB[col], b1 = cpB[col + 1];
    int32_t key[kE];
    Acc val[kE];
    const int filled = tiny_gather<K, T, SYM>(b0, b1, bpos, valB, cpA, rowsA, valA, key, val);
    if (filled > kTinyMax && lane == 0) atomicOr(err, 4);
    const int nr = (min(filled, kTinyMax) + kWave - 1) / kWave;  // registers in use (wave-uniform)
#pragma unroll
    for (int q = 0; q < kSPL; ++q) {
      hk[q * kWave + lane] = kEmpty;
      if constexpr (!SYM) s_val[w][q * kWave + lane] = SR::ident();
    }
    wave_lds_sync();
    // insert: every product finds its row's slot (linear probing)
    int slot[kE];
#pragma unroll
    for (int u = 0; u < kE; ++u) {
      slot[u] = -1;
      if (u < nr && key[u] != kNone) {
        uint32_t h = hash_slot(key[u], kLgHash);
        while (true) {
          const int32_t old = atomicCAS(&hk[h], kEmpty, key[u]);
          if (old == kEmpty || old == key[u]) break;
          h = (h + 1) & (kTinyHash - 1);
        }
        slot[u] = (int)h;
      }
    }
    if constexpr (!SYM) {
#pragma unroll
      for (int u = 0; u < kE; ++u)
        if (slot[u] >= 0) SR::acc(&s_val[w][slot[u]], val[u]);
    }
    wave_lds_sync();
    // compaction of the occupied slots (lane owns slots lane*kSPL ..)
    int32_t kk[kSPL];
    int occ = 0;
#pragma unroll
    for (int q = 0; q < kSPL; ++q) {
      kk[q] = hk[lane * kSPL + q];
      occ += kk[q] != kEmpty;
    }
    const int incl = wave_incl_scan(occ);
    const int D = __shfl(incl, kWave - 1, kWave);  // distinct rows of the column
    if constexpr (SYM) {
      if (lane == 0) nnzcol[col] = D;
    } else {
      int pos = incl - occ;
#pragma unroll
      for (int q = 0; q < kSPL; ++q)
        if (kk[q] != kEmpty) {
          crow[pos] = kk[q];
          s_slot[w][pos] = (uint16_t)(lane * kSPL + q);
          ++pos;
        }
      if (lane < 4) crow[D + lane] = kNone;  // padding for the four-wide reads
      wave_lds_sync();
      int64_t have, out;
      if (stage_off) {  // staged (no symbolic pass): its slot of flop_j entries; the count goes to nnzcol
        out = stage_off[i];
        have = stage_off[i + 1] - out;
        if (lane == 0) nnzcol[col] = D;
        if (lane == 0 && D > have) atomicOr(err, 8);
      } else {
        have = nnzcol[col];
        if (lane == 0 && D != have) atomicOr(err, 8);
        out = cpC[col];
      }
      // rank of every distinct row: the distinct rows below it
      const int nc = (D + kWave - 1) / kWave;  // compacted entries per lane (wave-uniform)
      int32_t mine[kE];
      int rank[kE];
#pragma unroll
      for (int r = 0; r < kE; ++r) {
        mine[r] = r < nc && r * kWave + lane < D ? crow[r * kWave + lane] : kNone;
        rank[r] = 0;
      }
      for (int j = 0; j < D; j += 4) {
        const int4 q4 = *reinterpret_cast<const int4*>(crow + j);
#pragma unroll
        for (int r = 0; r < kE; ++r)
          if (r < nc) rank[r] += (q4.x < mine[r]) + (q4.y < mine[r]) + (q4.z < mine[r]) + (q4.w < mine[r]);
      }
      if (!flt.on) {
#pragma unroll
        for (int r = 0; r < kE; ++r) {
          const int e = r * kWave + lane;
          if (r < nc && e < D && rank[r] < have) {
            irC[out + rank[r]] = mine[r];
            valC[out + rank[r]] = SR::out(s_val[w][s_slot[w][e]]);
          }
        }
      } else {  // the fused threshold: a kept row's place is the kept rows below it
        T v[kE];
        bool keep[kE];
#pragma unroll
        for (int r = 0; r < kE; ++r) {
          const int e = r * kWave + lane;
          keep[r] = false;
          if (r < nc && e < D) {
            v[r] = SR::out(s_val[w][s_slot[w][e]]);
            keep[r] = !(v[r] < flt.thr);
          }
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < kE; ++r)  // kept rows replace the compacted rows (others: kNone)
          if (r < nc && r * kWave + lane < D) crow[r * kWave + lane] = keep[r] ? mine[r] : kNone;
        wave_lds_sync();
        int kr[kE];
#pragma unroll
        for (int r = 0; r < kE; ++r) kr[r] = 0;
        for (int j = 0; j < D; j += 4) {
          const int4 q4 = *reinterpret_cast<const int4*>(crow + j);
#pragma unroll
          for (int r = 0; r < kE; ++r)
            if (r < nc) kr[r] += (q4.x < mine[r]) + (q4.y < mine[r]) + (q4.z < mine[r]) + (q4.w < mine[r]);
        }
        int kept = 0;
#pragma unroll
        for (int r = 0; r < kE; ++r) {
          kept += __popcll(__ballot(keep[r]));
          if (keep[r] && kr[r] < have) {
            irC[out + kr[r]] = mine[r];
            valC[out + kr[r]] = v[r];
          }
        }
        if (lane == 0) {
          if (stage_off) {
            nnzcol[col] = kept;
            if (D > kept) atomicAdd(flt.dropped, (unsigned long long)(D - kept));
          } else {
            filter_add_run(flt, out, kept, col, err);
          }
        }
      }
    }
    wave_lds_sync();  // the next column reuses the wave's table
  }
}

// Staged tiny columns into place: one wave per column copies its nnzcol[col]
// entries from its staging slot to cpC[col].
template <class T>
__global__ void k_tiny_place(const int32_t* __restrict__ order, int64_t first, int64_t ncols,
                             const int64_t* __restrict__ stage_off, const int64_t* __restrict__ nnzcol,
                             const int64_t* __restrict__ cpC, const int64_t* __restrict__ sir,
                             const T* __restrict__ sval, int64_t* __restrict__ irC, T* __restrict__ valC,
                             Filter<T> flt = {}, int* __restrict__ err = nullptr) {
  const int64_t i = global_wave();
  if (i >= ncols) return;
  const int col = order[first + i];
  const int64_t src = stage_off[i], dst = cpC[col], cnt = nnzcol[col];
  for (int64_t e = lane_id(); e < cnt; e += kWave) {
    irC[dst + e] = sir[src + e];
    valC[dst + e] = sval[src + e];
  }
  if (flt.on && lane_id() == 0) filter_add_run(flt, dst, (int)cnt, col, err);  // staged entries are the kept ones
}

// -------------------------------------------------------- symbolic: dense
// Distinct-row count with an LDS bitmap of RBW*32 rows per pass; a pass
// covers G row tiles of the tile-pointer table.
template <int NT, int RBW>
__global__ __launch_bounds__(NT) void k_sym_dense(const int32_t* __restrict__ order, int64_t first,
                                                  const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                  const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                  const int32_t* __restrict__ P, int S, int G, int TR,
                                                  int64_t* __restrict__ nnzcol) {
  __shared__ uint32_t bits[RBW];
  __shared__ long long wsum[NT / kWave];
  const int col = order[first + blockIdx.x];
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  long long cnt = 0;
  const int npass = (S + G - 1) / G;
  for (int sp = 0; sp < npass; ++sp) {
    for (int i = threadIdx.x; i < RBW; i += NT) bits[i] = 0u;
    __syncthreads();
    const int s_lo = sp * G, s_hi = min(S, s_lo + G);
    const int32_t row0 = s_lo * TR;
    for (int64_t eb = b0 + wave_id() * kWave; eb < b1; eb += NT) {
      const int64_t e = eb + lane_id();
      int64_t lo = 0;
      int len = 0;
      if (e < b1) {
        const int32_t p = bpos[e];
        if (p >= 0) {
          if (npass == 1) {
            lo = cpA[p];
            len = (int)(cpA[p + 1] - lo);
          } else {
            const int32_t* pp = P + (int64_t)p * (S + 1);
            const int a = pp[s_lo];
            lo = cpA[p] + a;
            len = pp[s_hi] - a;
          }
        }
      }
      wave_visit(lo, len, 0, [&](int64_t idx, int) {
        const int32_t r = rowsA[idx] - row0;
        const uint32_t bit = 1u << (r & 31);
        const uint32_t old = atomicOr(&bits[r >> 5], bit);
        if (!(old & bit)) ++cnt;
      });
    }
    __syncthreads();
  }
  const long long tot = block_sum_ll<NT>(cnt, wsum);
  if (threadIdx.x == 0) nnzcol[col] = tot;
}

// ---------------------------------------------------------- numeric: hash
// Hash accumulation (mtSpGEMM.h:362-423 / :595-622), compaction (:425-432)
// and sort by row (:434); SORT=false keeps hash order like LocalSpGEMMHash
// with sort=false (:624-634).
template <int K, class T, int HMAX, int NT, bool SORT>
__global__ __launch_bounds__(NT) void k_num_hash(const int32_t* __restrict__ order, int64_t first,
                                                 const int64_t* __restrict__ nnzcol, const int64_t* __restrict__ cpB,
                                                 const int32_t* __restrict__ bpos, const T* __restrict__ valB,
                                                 const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                 const T* __restrict__ valA, const int64_t* __restrict__ cpC,
                                                 int64_t* __restrict__ irC, T* __restrict__ valC,
                                                 int* __restrict__ err, Filter<T> flt = {}) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  __shared__ int32_t keys[HMAX];
  __shared__ Acc acc[HMAX];
  __shared__ long long wsum[NT / kWave];
  __shared__ SegStage<T, NT> st;
  const int col = order[first + blockIdx.x];
  const int64_t nz = nnzcol[col];
  const int lg = min(ceil_log2_dev(max(2 * nz, 64ll)), ceil_log2_dev(HMAX));
  const int H = 1 << lg;
  for (int i = threadIdx.x; i < H; i += NT) {
    keys[i] = -1;
    acc[i] = SR::ident();
  }
  __syncthreads();
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  for (int64_t eb = b0; eb < b1; eb += NT) {
    int64_t base;
    int len, nseg;
    T bv;
    load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, nullptr, 0, 0, 0, true, valB, base, len, bv);
    const int tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
    visit_staged<T, NT, 4>(st, nseg, tot, [&](const int64_t(&idx)[4], const T(&bvs)[4]) {
      int32_t rr[4];
      T aa[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        rr[u] = -1;
        if (idx[u] >= 0) {
          rr[u] = rowsA[idx[u]];
          aa[u] = valA[idx[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int32_t r = rr[u];
        if (r < 0) continue;
        const Acc v = SR::mul(aa[u], bvs[u]);
        uint32_t h = hash_slot(r, lg);
        int probe = 0;
        for (; probe < H; ++probe) {  // the table is at least twice the column's nnz
          const int32_t old = atomicCAS(&keys[h], -1, r);
          if (old == -1 || old == r) break;
          h = (h + 1) & (H - 1);
        }
        if (probe < H) SR::acc(&acc[h], v);
        else atomicOr(err, 2);
      }
    });
  }
  __syncthreads();
  // in-place compaction, NT slots at a time
  int base = 0;
  for (int c0 = 0; c0 < H; c0 += NT) {
    const int i = c0 + threadIdx.x;
    const int32_t k = i < H ? keys[i] : -1;
    const Acc a = i < H ? acc[i] : SR::ident();
    long long tot;
    const int ex = (int)block_excl_scan_ll<NT>(k != -1 ? 1 : 0, wsum, tot);
    if (k != -1) {
      keys[base + ex] = k;
      acc[base + ex] = a;
    }
    base += tot;
    __syncthreads();
  }
  if (base != nz && threadIdx.x == 0) atomicOr(err, 8);
  const int cnt = min(base, (int)nz);
  if constexpr (SORT) {
    int N = 1;
    while (N < cnt) N <<= 1;
    for (int i = cnt + threadIdx.x; i < N; i += NT) keys[i] = 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= N; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < N; i += NT) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const int32_t ki = keys[i], kx = keys[ixj];
            const bool up = (i & k) == 0;
            if ((ki > kx) == up) {
              keys[i] = kx;
              keys[ixj] = ki;
              const Acc t = acc[i];
              acc[i] = acc[ixj];
              acc[ixj] = t;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  const int64_t out = cpC[col];
  if (!flt.on) {
    for (int i = threadIdx.x; i < cnt; i += NT) {
      irC[out + i] = keys[i];
      valC[out + i] = SR::out(acc[i]);
    }
    return;
  }
  int kept = 0;  // the fused threshold: the kept entries, in order, at out
  for (int c0 = 0; c0 < cnt; c0 += NT) {
    const int i = c0 + threadIdx.x;
    const T v = i < cnt ? SR::out(acc[i]) : T{};
    const bool keep = i < cnt && !(v < flt.thr);
    long long tot;
    const int ex = (int)block_excl_scan_ll<NT>(keep ? 1 : 0, wsum, tot);
    if (keep) {
      irC[out + kept + ex] = keys[i];
      valC[out + kept + ex] = v;
    }
    kept += (int)tot;
  }
  if (threadIdx.x == 0) filter_add_run(flt, out, kept, col, err);
}

// --------------------------------------------------------- numeric: dense
// Row-tiled dense accumulation: TR rows per tile, TR = 32*NT so that each
// thread owns one 32-bit word of the presence bitmap during emission.
template <int K, class T, int NT>
__global__ __launch_bounds__(NT) void k_num_dense(const int32_t* __restrict__ order, int64_t first,
                                                  const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                  const T* __restrict__ valB, const int64_t* __restrict__ cpA,
                                                  const int32_t* __restrict__ rowsA, const T* __restrict__ valA,
                                                  const int32_t* __restrict__ P, int S,
                                                  const int64_t* __restrict__ cpC, int64_t* __restrict__ irC,
                                                  T* __restrict__ valC, int* __restrict__ err) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int TR = 32 * NT;
  __shared__ Acc acc[TR];
  __shared__ uint32_t bits[NT];
  __shared__ int wsum[NT / kWave];
  const int col = order[first + blockIdx.x];
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  int64_t out = cpC[col];
  const int64_t out_end = cpC[col + 1];
  for (int i = threadIdx.x; i < TR; i += NT) acc[i] = SR::ident();
  bits[threadIdx.x] = 0u;
  __syncthreads();
  for (int s = 0; s < S; ++s) {
    const int32_t row0 = s * TR;
    for (int64_t eb = b0 + wave_id() * kWave; eb < b1; eb += NT) {
      const int64_t e = eb + lane_id();
      int64_t lo = 0;
      int len = 0;
      T bv{};
      if (e < b1) {
        const int32_t p = bpos[e];
        if (p >= 0) {
          const int32_t* pp = P + (int64_t)p * (S + 1) + s;
          const int a = pp[0];
          lo = cpA[p] + a;
          len = pp[1] - a;
          bv = valB[e];
        }
      }
      wave_visit(lo, len, bv, [&](int64_t idx, T b) {
        const int32_t r = rowsA[idx] - row0;
        SR::acc(&acc[r], SR::mul(valA[idx], b));
        atomicOr(&bits[r >> 5], 1u << (r & 31));
      });
    }
    __syncthreads();
    uint32_t w = bits[threadIdx.x];
    int tot;
    const int ex = block_excl_scan<NT>(__popc(w), wsum, tot);
    int64_t o = out + ex;
    while (w) {
      const int b = __ffs(w) - 1;
      w &= w - 1;
      const int r = threadIdx.x * 32 + b;
      if (o < out_end) {  // symbolic and numeric agree; guard anyway
        irC[o] = (int64_t)row0 + r;
        valC[o] = SR::out(acc[r]);
      } else {
        atomicOr(err, 4);
      }
      acc[r] = SR::ident();
      ++o;
    }
    bits[threadIdx.x] = 0u;
    out += tot;
    __syncthreads();
  }
}

// Symbolic pass of the unit tier: one workgroup per column.  Units are cut
// over passes of 256 blocks (2^20 rows); within a pass the presence bitmap is
// built IBLK blocks at a time (each sub-pass gathers only the products in its
// rows, through the tile table P), so the LDS bitmap stays small enough for
// several workgroups per CU.  Writes nnz_j and the column's units.
//
// With `lists`, the column's sorted output rows in blocks of at most
// kListBlockMax outputs are also written, as 12-bit offsets within their
// 4096-row blocks, at their output positions in the column's region
// lists[list_off[slot] .. list_off[slot + 1]) (slot = the workgroup), and its
// per-block counts into bcnt16[slot * S ..].  The small units of a pass whose
// rows fit the region are then cut by output count alone over runs of such
// blocks: those wider than rblk_l blocks go to units_s, whose numeric pass
// takes ranks from the list instead of a bitmap.
//
// Eight waves per SIMD: the pass is latency-bound on its gathers, so the
// kernel asks for 64 VGPRs (six of its values spill to scratch) and keeps
// its LDS under 40 KB (the unit staging lives in the bitmap, idle while a
// pass is cut), four workgroups per CU.  At the compiler's own choice (80
// VGPRs, 47 KB: six waves) the pass took 64 ms at R-MAT s20, here 55.7 ms.
template <int NT, int U, int IBLK>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_sym_units(
    const int32_t* __restrict__ order, int64_t first,
                                                  const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                  const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                  const int32_t* __restrict__ P, int S, int kmax_l, int rblk_l,
                                                  int kmax_h, int rblk_h, int64_t heavy_thr,
                                                  int64_t* __restrict__ nnzcol, Unit* __restrict__ units_l,
                                                  int* __restrict__ n_l, int cap_l, Unit* __restrict__ units_h,
                                                  int* __restrict__ n_h, int cap_h, uint32_t* __restrict__ rmap,
                                                  int64_t rm_cols, int* __restrict__ err,
                                                  uint16_t* __restrict__ lists = nullptr,
                                                  const int64_t* __restrict__ list_off = nullptr,
                                                  uint16_t* __restrict__ bcnt16 = nullptr,
                                                  Unit* __restrict__ units_s = nullptr, int* __restrict__ n_s = nullptr,
                                                  int cap_s = 0) {
  constexpr int BPP = 256;             // blocks per cutting pass (2^20 rows)
  constexpr int BMW = IBLK * kBRW;     // bitmap words of one sub-pass
  static_assert(BPP % IBLK == 0, "sub-passes tile the cutting pass");
  static_assert(IBLK <= kWave, "a wave scans the block counts of a sub-pass");
  static_assert(sizeof(SegStage<uint8_t, NT>) >= BPP * sizeof(Unit), "list units are staged in the gather stage");
  __shared__ uint32_t bits[BMW];
  __shared__ int bpre[BPP];           // block counts, then their inclusive prefix
  __shared__ SegStage<uint8_t, NT> st;
  __shared__ long long wsum[NT / kWave];
  __shared__ int s_nu, s_slot, s_ns, s_sslot;
  // the units of a pass, staged while it is cut: in the bitmap, which every
  // thread has finished reading (barrier) and the next pass clears only after
  // the end-of-pass barrier
  static_assert(sizeof(uint32_t) * BMW >= BPP * sizeof(Unit), "the bitmap holds a pass's units");
  Unit* su = reinterpret_cast<Unit*>(bits);
  __shared__ unsigned long long s_dense[BPP / kWave];  // blocks of the pass above kListBlockMax outputs
  Unit* su2 = reinterpret_cast<Unit*>(&st);  // list units of a pass (the stage is idle while cutting)
  const int col = order[first + blockIdx.x];
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  const int npass = (S + BPP - 1) / BPP;
  const bool whole = S <= IBLK;  // one sub-pass covers every row: no tile look-ups
  int64_t list_base = 0, region = 0;
  if (lists) {
    list_base = list_off[blockIdx.x];
    region = list_off[blockIdx.x + 1] - list_base;
  }
  int64_t colcnt = 0;
  int unit_off = 0;  // meaningful in thread 0 only
  for (int sp = 0; sp < npass; ++sp) {
    const int t_lo = sp * BPP, t_hi = min(S, t_lo + BPP), nblk = t_hi - t_lo;
    bool listok = lists != nullptr;  // every row of this pass written so far (uniform)
    int64_t run = 0;                 // this pass's rows before the current sub-pass (uniform)
    for (int i_lo = t_lo; i_lo < t_hi; i_lo += IBLK) {
      const int i_hi = min(t_hi, i_lo + IBLK), inb = i_hi - i_lo;
      const int32_t row0 = i_lo * kBR;
      for (int i = threadIdx.x; i < inb * kBRW; i += NT) bits[i] = 0u;
      // (the staging barrier below orders the clear before any insertion)
      for (int64_t eb = b0; eb < b1; eb += NT) {
        int64_t base;
        int len;
        uint8_t dummy;
        load_seg<uint8_t, false>(eb + threadIdx.x, b1, bpos, cpA, P, S, i_lo, i_hi, whole, nullptr, base, len,
                                 dummy);
        int nseg;
        const int tot = stage_chunk<uint8_t, NT>(st, wsum, base, len, 0, nseg);
        visit_staged<uint8_t, NT, U>(st, nseg, tot, [&](const int64_t(&idx)[U], const uint8_t(&)[U]) {
          int32_t r[U];
#pragma unroll
          for (int u = 0; u < U; ++u) r[u] = idx[u] >= 0 ? rowsA[idx[u]] - row0 : -1;
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (r[u] >= 0) atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
        });
      }
      __syncthreads();
      // the column's row map (its presence bitmap, S blocks of kBRW words)
      // for the numeric pass, which then need not rebuild it; kept for the
      // first rm_cols columns of the launch (the only ones that can form
      // large units)
      if (rmap && blockIdx.x < rm_cols) {
        uint32_t* dst = rmap + (int64_t)blockIdx.x * S * kBRW + (int64_t)i_lo * kBRW;
        for (int i = threadIdx.x; i < inb * kBRW; i += NT) dst[i] = bits[i];
      }
      // per-block counts: one wave per block, two words per lane
      for (int b = wave_id(); b < inb; b += NT / kWave) {
        const uint32_t* w = bits + b * kBRW;
        int c = __popc(w[lane_id()]) + __popc(w[lane_id() + kWave]);
        c = (int)wave_sum_ll(c);
        if (lane_id() == 0) {
          bpre[i_lo - t_lo + b] = c;
          if (lists) bcnt16[(int64_t)blockIdx.x * S + i_lo + b] = (uint16_t)c;
        }
      }
      __syncthreads();  // the counts are read, and the next sub-pass clears the bitmap
      if (lists) {
        // the sub-pass's rows in order, one wave per block as for the counts
        // (two words, 64 rows, per lane): every wave scans the sub-pass's
        // block counts (IBLK <= 64, one per lane) for the block offsets and
        // the total, and its lanes place their rows by a wave scan, so the
        // stores of a wave cover one contiguous range
        const int bc = lane_id() < inb ? bpre[i_lo - t_lo + lane_id()] : 0;
        const int bincl = wave_incl_scan(bc);
        const int sub_total = __shfl(bincl, kWave - 1, kWave);
        const int64_t have = colcnt + run;
        listok = listok && have + sub_total <= region;
        if (listok) {
          for (int b = wave_id(); b < inb; b += NT / kWave) {
            const int boff = __shfl(bincl - bc, b, kWave);
            if (__shfl(bc, b, kWave) > kListBlockMax) continue;  // dense block: its units use the bitmap
            const uint32_t* w = bits + b * kBRW;
            uint64_t x = (uint64_t)w[2 * lane_id()] | ((uint64_t)w[2 * lane_id() + 1] << 32);
            const int c = __popcll(x);
            const int ex = wave_incl_scan(c) - c;
            uint16_t* dst = lists + list_base + have + boff + ex;
            int k = 0;
            while (x) {
              const int bit = __ffsll((unsigned long long)x) - 1;
              x &= x - 1;
              dst[k++] = (uint16_t)(lane_id() * 64 + bit);
            }
          }
        }
        run += sub_total;
        __syncthreads();  // the bitmap is read before the next sub-pass clears it
      }
    }
    if (wave_id() == 0) {  // inclusive prefix over blocks, and the mask of the blocks without a list
      int carry = 0;
      for (int b0i = 0; b0i < nblk; b0i += kWave) {
        const int b = b0i + lane_id();
        const int v = b < nblk ? bpre[b] : 0;
        const int inc = wave_incl_scan(v) + carry;
        if (b < nblk) bpre[b] = inc;
        carry = __shfl(inc, kWave - 1, kWave);
        const unsigned long long dense = __ballot(b < nblk && v > kListBlockMax);
        if (lane_id() == 0) s_dense[b0i / kWave] = dense;
      }
    }
    __syncthreads();
    const int pass_total = bpre[nblk - 1];
    if (threadIdx.x == 0) {
      // greedy cut into units (a single block never exceeds kBR <= kmax);
      // a small-unit pass with its list written is cut by count alone
      const bool heavy = pass_total > heavy_thr;
      const bool by_list = !heavy && listok;
      const int kmax = heavy ? kmax_h : kmax_l, rblk = heavy ? rblk_h : rblk_l;
      // first block at or after `from` whose rows were not listed (nblk if none)
      auto next_dense = [&](int from) {
        for (int q = from / kWave; q * kWave < nblk; ++q) {
          unsigned long long msk = s_dense[q];
          if (q == from / kWave) msk &= ~0ull << (from % kWave);
          if (msk) return min(nblk, q * kWave + __ffsll(msk) - 1);
        }
        return nblk;
      };
      int start = 0, before = 0, nu = 0, ns = 0;
      while (start < nblk) {
        const int lim = before + kmax;
        // a unit wider than rblk blocks takes its ranks from the list, so it
        // stops before the next block whose rows were not listed
        const int endmax = by_list ? min(nblk, max(start + rblk, next_dense(start))) : min(nblk, start + rblk);
        int lo = start, hi = endmax - 1;  // last b in [start, endmax) with bpre[b] <= lim
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (bpre[mid] <= lim) lo = mid;
          else hi = mid - 1;
        }
        // a unit whose first blocks alone fill half the value window stops
        // at the window's rows, so the numeric pass handles it in one sweep
        const int dl = kmax / kBR;
        if (lo - start + 1 > dl && bpre[start + dl - 1] - before >= kmax / 2) lo = start + dl - 1;
        const int c = bpre[lo] - before;
        if (c > 0) {
          Unit& u = (by_list && lo + 1 - start > rblk) ? su2[ns++] : su[nu++];
          u.col = col;
          u.blk_lo = t_lo + start;
          u.blk_hi = t_lo + lo + 1;
          u.cnt = c;
          u.off = unit_off;
          u.slot = (int32_t)blockIdx.x;
          unit_off += c;
        }
        before = bpre[lo];
        start = lo + 1;
      }
      s_nu = nu;
      s_ns = ns;
      int* cnt = heavy ? n_h : n_l;
      s_slot = nu ? atomicAdd(cnt, nu) : 0;
      if (s_slot + nu > (heavy ? cap_h : cap_l)) atomicOr(err, 32);
      s_sslot = ns ? atomicAdd(n_s, ns) : 0;
      if (ns && s_sslot + ns > cap_s) atomicOr(err, 32);
    }
    __syncthreads();
    {
      const bool heavy = pass_total > heavy_thr;
      Unit* dst = heavy ? units_h : units_l;
      const int cap = heavy ? cap_h : cap_l;
      for (int i = threadIdx.x; i < s_nu; i += NT)
        if (s_slot + i < cap) dst[s_slot + i] = su[i];
      for (int i = threadIdx.x; i < s_ns; i += NT)
        if (s_sslot + i < cap_s) units_s[s_sslot + i] = su2[i];
    }
    colcnt += pass_total;
    __syncthreads();
  }
  if (threadIdx.x == 0) nnzcol[col] = colcnt;
}

// Banded form of the same symbolic pass, in two kernels.  k_sym_band runs one
// workgroup per (column, band of IBLK row blocks) with the band as the slow
// grid dimension, so the workgroups in flight at any moment gather only the
// A entries of one band of rows (IBLK * 4096 rows: about nnz(A) * IBLK / S
// entries, which the L2 caches can hold), where the per-column pass sweeps
// every band of a few columns.  It builds the band's presence bitmap, keeps
// the row-map slice, and writes the band's per-block counts.  k_sym_cut then
// cuts every column into units from its block counts, as k_sym_units does.
template <int NT, int U, int IBLK>
__global__ __launch_bounds__(NT) void k_sym_band(const int32_t* __restrict__ order, int64_t first,
                                                 const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                 const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                 const int32_t* __restrict__ P, int S, int32_t* __restrict__ bcnt,
                                                 uint32_t* __restrict__ rmap, int64_t rm_cols) {
  constexpr int BMW = IBLK * kBRW;  // bitmap words of one band
  __shared__ uint32_t bits[BMW];
  __shared__ SegStage<uint8_t, NT> st;
  __shared__ long long wsum[NT / kWave];
  const int64_t slot = blockIdx.x;  // the column's position in the launch
  const int col = order[first + slot];
  const int i_lo = (int)blockIdx.y * IBLK, i_hi = min(S, i_lo + IBLK), inb = i_hi - i_lo;
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  const int32_t row0 = i_lo * kBR;
  const bool whole = S <= IBLK;  // one band covers every row: no tile look-ups
  for (int i = threadIdx.x; i < inb * kBRW; i += NT) bits[i] = 0u;
  // (the staging barrier below orders the clear before any insertion)
  for (int64_t eb = b0; eb < b1; eb += NT) {
    int64_t base;
    int len;
    uint8_t dummy;
    load_seg<uint8_t, false>(eb + threadIdx.x, b1, bpos, cpA, P, S, i_lo, i_hi, whole, nullptr, base, len, dummy);
    int nseg;
    const int tot = stage_chunk<uint8_t, NT>(st, wsum, base, len, 0, nseg);
    visit_staged<uint8_t, NT, U>(st, nseg, tot, [&](const int64_t(&idx)[U], const uint8_t(&)[U]) {
      int32_t r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) r[u] = idx[u] >= 0 ? rowsA[idx[u]] - row0 : -1;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (r[u] >= 0) atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
    });
  }
  __syncthreads();
  if (rmap && slot < rm_cols) {
    uint32_t* dst = rmap + slot * S * kBRW + (int64_t)i_lo * kBRW;
    for (int i = threadIdx.x; i < inb * kBRW; i += NT) dst[i] = bits[i];
  }
  for (int b = wave_id(); b < inb; b += NT / kWave) {  // one wave per block, two words per lane
    const uint32_t* wv = bits + b * kBRW;
    int c = __popc(wv[lane_id()]) + __popc(wv[lane_id() + kWave]);
    c = (int)wave_sum_ll(c);
    if (lane_id() == 0) bcnt[slot * S + i_lo + b] = c;
  }
}

// Unit cutting of the banded symbolic pass: one workgroup of 256 threads per
// column, over passes of 256 blocks, with the greedy rule of k_sym_units.
__global__ __launch_bounds__(256) void k_sym_cut(const int32_t* __restrict__ order, int64_t first,
                                                 const int32_t* __restrict__ bcnt, int S, int kmax_l, int rblk_l,
                                                 int kmax_h, int rblk_h, int64_t heavy_thr,
                                                 int64_t* __restrict__ nnzcol, Unit* __restrict__ units_l,
                                                 int* __restrict__ n_l, int cap_l, Unit* __restrict__ units_h,
                                                 int* __restrict__ n_h, int cap_h, int* __restrict__ err) {
  constexpr int BPP = 256;
  __shared__ int bpre[BPP];
  __shared__ Unit su[BPP];
  __shared__ long long wsum[256 / kWave];
  __shared__ int s_nu, s_slot;
  const int64_t slot = blockIdx.x;
  const int col = order[first + slot];
  const int npass = (S + BPP - 1) / BPP;
  int64_t colcnt = 0;
  int unit_off = 0;  // meaningful in thread 0 only
  for (int sp = 0; sp < npass; ++sp) {
    const int t_lo = sp * BPP, nblk = min(S, t_lo + BPP) - t_lo;
    const int c = (int)threadIdx.x < nblk ? bcnt[slot * S + t_lo + threadIdx.x] : 0;
    long long total;
    const long long ex = block_excl_scan_ll<256>(c, wsum, total);
    bpre[threadIdx.x] = (int)(ex + c);  // inclusive prefix over the pass's blocks
    __syncthreads();
    const int pass_total = (int)total;
    if (threadIdx.x == 0) {
      const bool heavy = pass_total > heavy_thr;
      const int kmax = heavy ? kmax_h : kmax_l, rblk = heavy ? rblk_h : rblk_l;
      int start = 0, before = 0, nu = 0;
      while (start < nblk) {
        const int lim = before + kmax;
        const int endmax = min(nblk, start + rblk);
        int lo = start, hi = endmax - 1;  // last b in [start, endmax) with bpre[b] <= lim
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (bpre[mid] <= lim) lo = mid;
          else hi = mid - 1;
        }
        const int dl = kmax / kBR;
        if (lo - start + 1 > dl && bpre[start + dl - 1] - before >= kmax / 2) lo = start + dl - 1;
        const int cu = bpre[lo] - before;
        if (cu > 0) {
          su[nu].col = col;
          su[nu].blk_lo = t_lo + start;
          su[nu].blk_hi = t_lo + lo + 1;
          su[nu].cnt = cu;
          su[nu].off = unit_off;
          su[nu].slot = (int32_t)slot;
          unit_off += cu;
          ++nu;
        }
        before = bpre[lo];
        start = lo + 1;
      }
      s_nu = nu;
      int* cnt = heavy ? n_h : n_l;
      s_slot = nu ? atomicAdd(cnt, nu) : 0;
      if (s_slot + nu > (heavy ? cap_h : cap_l)) atomicOr(err, 32);
    }
    __syncthreads();
    {
      const bool heavy = pass_total > heavy_thr;
      Unit* dst = heavy ? units_h : units_l;
      const int cap = heavy ? cap_h : cap_l;
      for (int i = threadIdx.x; i < s_nu; i += 256)
        if (s_slot + i < cap) dst[s_slot + i] = su[i];
    }
    colcnt += pass_total;
    __syncthreads();
  }
  if (threadIdx.x == 0) nnzcol[col] = colcnt;
}

// Emission of a unit's cnt outputs at `out`, rows ascending.  direct: the
// value of window row r sits at vals[r] (rows [0, ninit), presence in bits);
// otherwise at vals[rank], with the rank of a row given by its bitmap word's
// prefix count pre[] and the rows staged in rank order in the value array's
// LDS after the values have left (every Acc is at least 4 bytes), so both
// leave in full, coalesced lines.  With the filter on, only entries with
// !(v < thr) are written, compacted at `out`, and the unit's range is
// appended to the run list; `scratch` (LDS no longer needed by the caller,
// cnt bits) holds the kept flags between the value and row passes.
template <int K, class T, int NT>
__device__ __forceinline__ void unit_emit(bool direct, int ninit, int cnt, int nw64, const uint32_t* bits,
                                          const uint16_t* pre, typename Sem<K, T>::Acc* vals, int32_t row0,
                                          int64_t out, int64_t* __restrict__ irC, T* __restrict__ valC,
                                          const Filter<T>& flt, int col, uint32_t* scratch, long long* wsum,
                                          int* err) {
  using SR = Sem<K, T>;
  auto rank_of = [&](int32_t r) {
    const int w = r >> 6, b = r & 63;
    const uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
    return (int)pre[w] + __popcll(word & ((1ull << b) - 1ull));
  };
  auto present = [&](int r) { return ((bits[r >> 5] >> (r & 31)) & 1u) != 0; };
  if (direct) {
    if (!flt.on) {  // one thread per row, stores stay coalesced
      for (int r = threadIdx.x; r < ninit; r += NT) {
        if (!present(r)) continue;
        const int q = rank_of(r);
        if (q < cnt) {
          irC[out + q] = (int64_t)row0 + r;
          valC[out + q] = SR::out(vals[r]);
        }
      }
      return;
    }
    int kept = 0;
    for (int c0 = 0; c0 < ninit; c0 += NT) {
      const int r = c0 + threadIdx.x;
      T v{};
      bool keep = false;
      if (r < ninit && present(r)) {
        v = SR::out(vals[r]);
        keep = !(v < flt.thr);
      }
      long long tot;
      const int ex = (int)block_excl_scan_ll<NT>(keep ? 1 : 0, wsum, tot);
      if (keep && kept + ex < cnt) {
        irC[out + kept + ex] = (int64_t)row0 + r;
        valC[out + kept + ex] = v;
      }
      kept += (int)tot;
    }
    if (threadIdx.x == 0) filter_add_run(flt, out, min(kept, cnt), col, err);
    return;
  }
  int kept = 0;
  if (!flt.on) {
    for (int i = threadIdx.x; i < cnt; i += NT) valC[out + i] = SR::out(vals[i]);
  } else {
    for (int i = threadIdx.x; i <= cnt / 32; i += NT) scratch[i] = 0u;
    __syncthreads();
    for (int c0 = 0; c0 < cnt; c0 += NT) {
      const int i = c0 + threadIdx.x;
      const T v = i < cnt ? SR::out(vals[i]) : T{};
      const bool keep = i < cnt && !(v < flt.thr);
      long long tot;
      const int ex = (int)block_excl_scan_ll<NT>(keep ? 1 : 0, wsum, tot);
      if (keep) {
        valC[out + kept + ex] = v;
        atomicOr(&scratch[i >> 5], 1u << (i & 31));
      }
      kept += (int)tot;
    }
  }
  __syncthreads();
  uint32_t* rstage = reinterpret_cast<uint32_t*>(vals);
  for (int w = threadIdx.x; w < nw64; w += NT) {
    uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
    int q = pre[w];
    while (word) {
      const int b = __ffsll((unsigned long long)word) - 1;
      word &= word - 1;
      if (q < cnt) rstage[q] = (uint32_t)(w * 64 + b);
      ++q;
    }
  }
  __syncthreads();
  if (!flt.on) {
    for (int i = threadIdx.x; i < cnt; i += NT) irC[out + i] = (int64_t)row0 + rstage[i];
    return;
  }
  int kept2 = 0;
  for (int c0 = 0; c0 < cnt; c0 += NT) {
    const int i = c0 + threadIdx.x;
    const bool keep = i < cnt && ((scratch[i >> 5] >> (i & 31)) & 1u);
    long long tot;
    const int ex = (int)block_excl_scan_ll<NT>(keep ? 1 : 0, wsum, tot);
    if (keep) irC[out + kept2 + ex] = (int64_t)row0 + rstage[i];
    kept2 += (int)tot;
  }
  if (threadIdx.x == 0) filter_add_run(flt, out, kept, col, err);
}

// Numeric pass of the unit tier: one workgroup per unit.  Sweep 1 sets the
// unit's presence bitmap, a block scan turns it into per-word ranks, sweep 2
// accumulates every product at its row's rank.  When the column's B entries
// fit one chunk, the first CR rounds of products (row, product value) stay in
// registers between the sweeps, so sweep 2 re-reads nothing for them.
// Waves per SIMD the small-unit kernel asks for: eight where its LDS allows
// four 512-thread workgroups per CU (a value window of at most 16 KB: 4-byte
// accumulators), which then caps it at 64 VGPRs; otherwise the compiler's
// choice (the LDS limits 8-byte windows to six waves anyway).
template <class Acc, int NT, int KMAX>
struct UnitWaves {
  static constexpr int value = (NT == 512 && sizeof(Acc) * KMAX <= 16384) ? 8 : 1;
};

template <int K, class T, int NT, int U, int KMAX, int RBLK, int CR>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(
    UnitWaves<typename Sem<K, T>::Acc, NT, KMAX>::value))) void k_num_units(const Unit* __restrict__ units,
                                                                            const int64_t* __restrict__ cpB,
                                                  const int32_t* __restrict__ bpos, const T* __restrict__ valB,
                                                  const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                  const T* __restrict__ valA, const int32_t* __restrict__ P, int S,
                                                  const int64_t* __restrict__ cpC, int64_t* __restrict__ irC,
                                                  T* __restrict__ valC, int* __restrict__ err, Filter<T> flt = {}, int xg = 0) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int NW32 = RBLK * kBRW;   // bitmap words
  constexpr int NW64 = NW32 / 2;
  constexpr int WPT = (NW64 + NT - 1) / NT;  // 64-bit words per thread in the prefix scan
  __shared__ Acc vals[KMAX];
  __shared__ uint32_t bits[NW32];
  __shared__ uint16_t pre[NW64];
  __shared__ SegStage<T, NT> st;
  __shared__ long long wsum[NT / kWave];
  const Unit un = units[xcd_unit((int)blockIdx.x, (int)gridDim.x, xg)];
  const int col = un.col;
  const int t_lo = un.blk_lo, t_hi = un.blk_hi;
  const int32_t row0 = t_lo * kBR;
  const int cnt = un.cnt;
  const int64_t out = cpC[col] + un.off;
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  // A unit whose rows fit the value window (dense output) accumulates at
  // vals[row - row0] in a single sweep; the others count first and
  // accumulate at the rank in a second sweep.
  const bool direct = (t_hi - t_lo) * kBR <= KMAX;
  const bool single_chunk = !direct && (b1 - b0) <= NT;
  if (cnt > KMAX || t_hi - t_lo > RBLK) {  // units are cut to fit; guard anyway
    if (threadIdx.x == 0) atomicOr(err, 64);
    return;
  }
  const int ninit = direct ? (t_hi - t_lo) * kBR : cnt;
  for (int i = threadIdx.x; i < ninit; i += NT) vals[i] = SR::ident();
  const int nw32 = (t_hi - t_lo) * kBRW;
  for (int i = threadIdx.x; i < nw32; i += NT) bits[i] = 0u;

  auto set_bits = [&](const int64_t(&idx)[U], const T(&)[U]) {
    int32_t r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = idx[u] >= 0 ? rowsA[idx[u]] - row0 : -1;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r[u] >= 0) atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
  };
  auto rank_of = [&](int32_t r) {
    const int w = r >> 6, b = r & 63;
    const uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
    return (int)pre[w] + __popcll(word & ((1ull << b) - 1ull));
  };
  auto accumulate = [&](const int64_t(&idx)[U], const T(&bvs)[U]) {
    int32_t r[U];
    T a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = -1;
      if (idx[u] >= 0) {
        r[u] = rowsA[idx[u]] - row0;
        a[u] = valA[idx[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (r[u] < 0) continue;
      const int rank = rank_of(r[u]);
      if (rank < cnt) SR::acc(&vals[rank], SR::mul(a[u], bvs[u]));
    }
  };

  int32_t crow[CR][U];
  Acc cprod[CR][U];
  int tot = 0, nseg = 0;
  if (single_chunk) {
    int64_t base;
    int len;
    T bv;
    load_seg<T, true>(b0 + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
    tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
    // sweep 1, cached rounds
#pragma unroll
    for (int rr = 0; rr < CR; ++rr) {
      int64_t idx[U];
      T bvs[U];
      fill_round<T, NT, U>(st, nseg, tot, (rr * NT + (int)threadIdx.x) * U, idx, bvs);
      T a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        crow[rr][u] = -1;
        if (idx[u] >= 0) {
          crow[rr][u] = rowsA[idx[u]] - row0;
          a[u] = valA[idx[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (crow[rr][u] >= 0) {
          atomicOr(&bits[crow[rr][u] >> 5], 1u << (crow[rr][u] & 31));
          cprod[rr][u] = SR::mul(a[u], bvs[u]);
        }
    }
    visit_staged<T, NT, U>(st, nseg, tot, set_bits, CR);
  } else {
    auto acc_direct = [&](const int64_t(&idx)[U], const T(&bvs)[U]) {
      int32_t r[U];
      T a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        r[u] = -1;
        if (idx[u] >= 0) {
          r[u] = rowsA[idx[u]] - row0;
          a[u] = valA[idx[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (r[u] >= 0) {
          atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
          SR::acc(&vals[r[u]], SR::mul(a[u], bvs[u]));
        }
    };
    if (direct) __syncthreads();  // the value window is initialised before anyone accumulates
    for (int64_t eb = b0; eb < b1; eb += NT) {
      int64_t base;
      int len;
      T bv;
      load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
      tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
      if (direct) visit_staged<T, NT, U>(st, nseg, tot, acc_direct);
      else visit_staged<T, NT, U>(st, nseg, tot, set_bits);
    }
  }
  __syncthreads();
  // exclusive prefix popcount per 64-bit word
  const int nw64 = nw32 / 2;
  {
    int local[WPT];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      const int w = threadIdx.x * WPT + q;
      const int c = w < nw64 ? __popc(bits[2 * w]) + __popc(bits[2 * w + 1]) : 0;
      local[q] = sum;
      sum += c;
    }
    long long total;
    const int ex = (int)block_excl_scan_ll<NT>(sum, wsum, total);
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      const int w = threadIdx.x * WPT + q;
      if (w < nw64) pre[w] = (uint16_t)(ex + local[q]);
    }
    if (threadIdx.x == 0 && total != cnt) atomicOr(err, 128);
  }
  __syncthreads();
  if (direct) {  // values already sit at their rows
    unit_emit<K, T, NT>(true, ninit, cnt, nw64, bits, pre, vals, row0, out, irC, valC, flt, col,
                        reinterpret_cast<uint32_t*>(&st), wsum, err);
    return;
  }
  // sweep 2: accumulate values at their ranks
  if (single_chunk) {
#pragma unroll
    for (int rr = 0; rr < CR; ++rr)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (crow[rr][u] >= 0) {
          const int rank = rank_of(crow[rr][u]);
          if (rank < cnt) SR::acc(&vals[rank], cprod[rr][u]);
        }
    visit_staged<T, NT, U>(st, nseg, tot, accumulate, CR);
  } else {
    for (int64_t eb = b0; eb < b1; eb += NT) {
      int64_t base;
      int len;
      T bv;
      load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
      int ns2;
      const int t2 = stage_chunk<T, NT>(st, wsum, base, len, bv, ns2);
      visit_staged<T, NT, U>(st, ns2, t2, accumulate);
    }
  }
  __syncthreads();
  unit_emit<K, T, NT>(false, ninit, cnt, nw64, bits, pre, vals, row0, out, irC, valC, flt, col,
                      reinterpret_cast<uint32_t*>(&st), wsum, err);
}

// Numeric pass of the unit tier with the symbolic pass's row map: the unit's
// presence bitmap is loaded (coalesced) instead of rebuilt, ranks come from
// its prefix popcounts, and one sweep gathers every product once and
// accumulates it at its rank (at its row for a unit whose rows fit the value
// window).
template <int K, class T, int NT, int U, int KMAX, int RBLK>
__global__ __launch_bounds__(NT) void k_num_units_rm(const Unit* __restrict__ units, const uint32_t* __restrict__ rmap,
                                                     const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                     const T* __restrict__ valB, const int64_t* __restrict__ cpA,
                                                     const int32_t* __restrict__ rowsA, const T* __restrict__ valA,
                                                     const int32_t* __restrict__ P, int S, int64_t rm_cols,
                                                     const int64_t* __restrict__ cpC, int64_t* __restrict__ irC,
                                                     T* __restrict__ valC, int* __restrict__ err,
                                                     Filter<T> flt = {}, int xg = 0) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int NW32 = RBLK * kBRW;
  constexpr int NW64 = NW32 / 2;
  constexpr int WPT = (NW64 + NT - 1) / NT;
  __shared__ Acc vals[KMAX];
  __shared__ uint32_t bits[NW32];
  __shared__ uint16_t pre[NW64];
  __shared__ SegStage<T, NT> st;
  __shared__ long long wsum[NT / kWave];
  const Unit un = units[xcd_unit((int)blockIdx.x, (int)gridDim.x, xg)];
  const int col = un.col;
  const int t_lo = un.blk_lo, t_hi = un.blk_hi;
  const int32_t row0 = t_lo * kBR;
  const int cnt = un.cnt;
  const int64_t out = cpC[col] + un.off;
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  const bool direct = (t_hi - t_lo) * kBR <= KMAX;
  if (cnt > KMAX || t_hi - t_lo > RBLK || t_lo < 0 || t_hi > S || un.slot < 0 || un.slot >= rm_cols) {
    if (threadIdx.x == 0) atomicOr(err, 64);
    return;
  }
  const int ninit = direct ? (t_hi - t_lo) * kBR : cnt;
  for (int i = threadIdx.x; i < ninit; i += NT) vals[i] = SR::ident();
  const int nw32 = (t_hi - t_lo) * kBRW;
  const uint32_t* src = rmap + (int64_t)un.slot * S * kBRW + (int64_t)t_lo * kBRW;
  for (int i = threadIdx.x; i < nw32; i += NT) bits[i] = src[i];
  __syncthreads();
  const int nw64 = nw32 / 2;
  {
    int local[WPT];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      const int w = threadIdx.x * WPT + q;
      const int c = w < nw64 ? __popc(bits[2 * w]) + __popc(bits[2 * w + 1]) : 0;
      local[q] = sum;
      sum += c;
    }
    long long total;
    const int ex = (int)block_excl_scan_ll<NT>(sum, wsum, total);
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      const int w = threadIdx.x * WPT + q;
      if (w < nw64) pre[w] = (uint16_t)(ex + local[q]);
    }
    if (threadIdx.x == 0 && total != cnt) atomicOr(err, 128);
  }
  __syncthreads();
  auto rank_of = [&](int32_t r) {
    const int w = r >> 6, b = r & 63;
    const uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
    return (int)pre[w] + __popcll(word & ((1ull << b) - 1ull));
  };
  auto accumulate = [&](const int64_t(&idx)[U], const T(&bvs)[U]) {
    int32_t r[U];
    T a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = -1;
      if (idx[u] >= 0) {
        r[u] = rowsA[idx[u]] - row0;
        a[u] = valA[idx[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (r[u] < 0) continue;
      const int q = direct ? r[u] : rank_of(r[u]);
      if (q < (direct ? ninit : cnt)) SR::acc(&vals[q], SR::mul(a[u], bvs[u]));
    }
  };
  for (int64_t eb = b0; eb < b1; eb += NT) {
    int64_t base;
    int len;
    T bv;
    load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
    int nseg;
    const int tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
    visit_staged<T, NT, U>(st, nseg, tot, accumulate);
  }
  __syncthreads();
  unit_emit<K, T, NT>(direct, ninit, cnt, nw64, bits, pre, vals, row0, out, irC, valC, flt, col,
                      reinterpret_cast<uint32_t*>(&st), wsum, err);
}

// Numeric pass of the list units: at most KMAX outputs of one column over
// any number of the row blocks of one 256-block pass.  Their ranks come from
// the column's sorted row list that the symbolic pass wrote (12-bit offsets
// grouped by block, with per-block counts), so one sweep gathers every
// product once: a product's rank is its block's first list position plus a
// binary search among that block's offsets.  Units this sparse would need a
// bitmap of up to 256 blocks (128 KB), or many narrow units that each
// re-read the B column and gather one- or two-entry A segments.
template <int K, class T, int NT, int U, int KMAX>
__global__ __launch_bounds__(NT) void k_num_list(const Unit* __restrict__ units, const uint16_t* __restrict__ lists,
                                                 const int64_t* __restrict__ list_off,
                                                 const uint16_t* __restrict__ bcnt16, const int64_t* __restrict__ cpB,
                                                 const int32_t* __restrict__ bpos, const T* __restrict__ valB,
                                                 const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                 const T* __restrict__ valA, const int32_t* __restrict__ P, int S,
                                                 const int64_t* __restrict__ cpC, int64_t* __restrict__ irC,
                                                 T* __restrict__ valC, int* __restrict__ err, Filter<T> flt = {}, int xg = 0) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int BPP = 256;  // blocks of one symbolic cutting pass
  static_assert(NT > BPP, "one block count per thread, plus the total");
  __shared__ Acc vals[KMAX];
  __shared__ uint16_t lst[KMAX];
  __shared__ int dir[BPP + 1];  // list position of each block's first row; dir[nb] = cnt
  __shared__ SegStage<T, NT> st;
  __shared__ long long wsum[NT / kWave];
  const Unit un = units[xcd_unit((int)blockIdx.x, (int)gridDim.x, xg)];
  const int col = un.col;
  const int t_lo = un.blk_lo, t_hi = un.blk_hi, nb = t_hi - t_lo;
  const int32_t row0 = t_lo * kBR;
  const int cnt = un.cnt;
  const int64_t out = cpC[col] + un.off;
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  if (cnt > KMAX || nb <= 0 || nb > BPP) {  // units are cut to fit; guard anyway
    if (threadIdx.x == 0) atomicOr(err, 64);
    return;
  }
  const uint16_t* src = lists + list_off[un.slot] + un.off;
  for (int i = threadIdx.x; i < cnt; i += NT) {
    lst[i] = src[i];
    vals[i] = SR::ident();
  }
  {
    const int c = threadIdx.x < nb ? (int)bcnt16[(int64_t)un.slot * S + t_lo + threadIdx.x] : 0;
    long long tot;
    const int ex = (int)block_excl_scan_ll<NT>(c, wsum, tot);  // barriers: lst and vals are ready after it
    if (tot != cnt) {  // the counts and the unit disagree (uniform): positions would leave the list
      if (threadIdx.x == 0) atomicOr(err, 128);
      return;
    }
    if (threadIdx.x <= nb) dir[threadIdx.x] = ex;
  }
  __syncthreads();
  auto accumulate = [&](const int64_t(&idx)[U], const T(&bvs)[U]) {
    int32_t r[U];
    T a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = -1;
      if (idx[u] >= 0) {
        r[u] = rowsA[idx[u]] - row0;
        a[u] = valA[idx[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (r[u] < 0) continue;
      const int b = r[u] >> 12;  // kBR = 4096 rows per block
      const uint16_t o = (uint16_t)(r[u] & (kBR - 1));
      const int end = dir[b + 1];
      int lo = dir[b], hi = end;  // first position in the block with offset >= o
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lst[mid] < o) lo = mid + 1;
        else hi = mid;
      }
      if (lo < end && lst[lo] == o) SR::acc(&vals[lo], SR::mul(a[u], bvs[u]));
      else atomicOr(err, 4);  // symbolic and numeric agree; guard anyway
    }
  };
  for (int64_t eb = b0; eb < b1; eb += NT) {
    int64_t base;
    int len;
    T bv;
    load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
    int nseg;
    const int tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
    visit_staged<T, NT, U>(st, nseg, tot, accumulate);
  }
  __syncthreads();
  auto row_of = [&](int i) {  // the block holding list position i: largest b with dir[b] <= i
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (dir[mid] <= i) lo = mid;
      else hi = mid;
    }
    return (int64_t)row0 + (int64_t)lo * kBR + lst[i];
  };
  if (!flt.on) {
    for (int i = threadIdx.x; i < cnt; i += NT) {
      irC[out + i] = row_of(i);
      valC[out + i] = SR::out(vals[i]);
    }
    return;
  }
  int kept = 0;
  for (int c0 = 0; c0 < cnt; c0 += NT) {
    const int i = c0 + threadIdx.x;
    T v{};
    bool keep = false;
    if (i < cnt) {
      v = SR::out(vals[i]);
      keep = !(v < flt.thr);
    }
    long long tot;
    const int ex = (int)block_excl_scan_ll<NT>(keep ? 1 : 0, wsum, tot);
    if (keep) {
      irC[out + kept + ex] = row_of(i);
      valC[out + kept + ex] = v;
    }
    kept += (int)tot;
  }
  if (threadIdx.x == 0) filter_add_run(flt, out, kept, col, err);
}

// Predictor features per stored column of B (one wave per column); see
// cbx_spgemm_features in cbx.h for the list.
__global__ void k_features(const int64_t* __restrict__ cpB, int64_t nzcB, const int32_t* __restrict__ bpos,
                           const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA, int64_t m,
                           float* __restrict__ feats) {
  const int64_t col = global_wave();
  if (col >= nzcB) return;
  const int lane = lane_id();
  long long f = 0, mx = 0, lo = m, hi = -1;
  for (int64_t e = cpB[col] + lane; e < cpB[col + 1]; e += kWave) {
    const int32_t p = bpos[e];
    if (p < 0) continue;
    const int64_t a0 = cpA[p], a1 = cpA[p + 1];
    if (a1 == a0) continue;
    const long long len = a1 - a0;
    f += len;
    mx = len > mx ? len : mx;
    const long long r0 = rowsA[a0], r1 = rowsA[a1 - 1];
    lo = r0 < lo ? r0 : lo;
    hi = r1 > hi ? r1 : hi;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    f += __shfl_xor(f, d, kWave);
    const long long m2 = __shfl_xor(mx, d, kWave);
    mx = m2 > mx ? m2 : mx;
    const long long l2 = __shfl_xor(lo, d, kWave);
    lo = l2 < lo ? l2 : lo;
    const long long h2 = __shfl_xor(hi, d, kWave);
    hi = h2 > hi ? h2 : hi;
  }
  if (lane != 0) return;
  const double b = (double)(cpB[col + 1] - cpB[col]);
  const double span = f > 0 ? (double)(hi - lo + 1) : 0.0;
  const double est = span > 0 ? span * (1.0 - exp(-(double)f / span)) : 0.0;
  float* o = feats + col * CBX_MLP_FEATURES;
  o[0] = (float)log2(1.0 + b);
  o[1] = (float)log2(1.0 + (double)f);
  o[2] = (float)log2(1.0 + (double)mx);
  o[3] = (float)log2(1.0 + (double)f / (b > 0 ? b : 1.0));
  o[4] = (float)(span / (double)(m > 0 ? m : 1));
  o[5] = (float)log2(1.0 + est);
}

// ============================================================ learned tier
// Single pass: units are cut from predicted column sizes, and each unit's
// exact size (from its presence bitmap) is published to a decoupled
// look-back chain that yields its output offset before its values are
// accumulated and written.
struct UnitL {
  int32_t col;
  int32_t blk_lo;
  int32_t blk_hi;
};

// Row histogram of A over kBR-row blocks (the prior that spreads a
// predicted column size over the blocks).
__global__ void k_row_hist(const int32_t* __restrict__ rowsA, int64_t nnzA, int S, unsigned long long* __restrict__ hist) {
  extern __shared__ unsigned int lh[];
  for (int b = threadIdx.x; b < S; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnzA; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&lh[rowsA[i] / kBR], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < S; b += blockDim.x)
    if (lh[b]) atomicAdd(&hist[b], (unsigned long long)lh[b]);
}

// Lowest and highest output row block a column can touch (from the first and
// last row of every contributing A column).  One wave per column.
__global__ void k_col_blocks(const int64_t* __restrict__ cpB, int64_t nzcB, const int32_t* __restrict__ bpos,
                             const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                             int32_t* __restrict__ blo, int32_t* __restrict__ bhi) {
  const int64_t col = global_wave();
  if (col >= nzcB) return;
  int lo = 0x7fffffff, hi = -1;
  for (int64_t e = cpB[col] + lane_id(); e < cpB[col + 1]; e += kWave) {
    const int32_t p = bpos[e];
    if (p < 0) continue;
    const int64_t a0 = cpA[p], a1 = cpA[p + 1];
    if (a1 == a0) continue;
    lo = min(lo, rowsA[a0] / kBR);
    hi = max(hi, rowsA[a1 - 1] / kBR);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    lo = min(lo, __shfl_xor(lo, d, kWave));
    hi = max(hi, __shfl_xor(hi, d, kWave));
  }
  if (lane_id() == 0) {
    blo[col] = lo;
    bhi[col] = hi;
  }
}

// Greedy cut of a column into units from expected block counts.  Block b of
// column j is expected to hold flop_j * w[b] products (w = A's row
// histogram) and, its 4096 rows each being hit at rate
// lambda = flop_j * w[b] / 4096, g_b = 4096 (1 - e^-lambda) distinct rows
// (saturating for dense columns, proportional to w for sparse ones); the
// predicted column size is spread over the blocks in proportion to g_b.  A
// block expected to hold at least direct_flop products forms a unit of its
// own (its rows fit the value window: one sweep, no counting sweep).  Other
// units close when their expected outputs would pass `target`, their
// expected products would pass `fcap`, or their span would pass rblk blocks.
// Capping the products keeps every unit's counting sweep short, so no unit
// holds its successors' look-back for long.  Pass 1 counts (out == null),
// pass 2 writes at `base`.
__device__ __forceinline__ float distinct_rows(float flop, float wb) {
  return (float)kBR * (1.f - __expf(-flop * wb * (1.f / (float)kBR)));
}

__device__ __forceinline__ int cut_units(int col, float pred, float flop, int lo, int hi,
                                         const float* __restrict__ w, float target, float fcap, int rblk,
                                         float direct_flop, UnitL* out) {
  float gsum = 0.f;
  for (int b = lo; b <= hi; ++b) gsum += distinct_rows(flop, w[b]);
  const float scale = gsum > 0.f ? pred / gsum : 0.f;
  int n = 0, start = lo;
  float acc = 0.f, accf = 0.f;
  for (int b = lo; b <= hi; ++b) {
    const float e = scale * distinct_rows(flop, w[b]), ef = flop * w[b];
    if (ef >= direct_flop) {  // more products than a unit holds in registers: a block of its own, one sweep
      if (b > start) {
        if (out) out[n] = UnitL{col, start, b};
        ++n;
      }
      if (out) out[n] = UnitL{col, b, b + 1};
      ++n;
      start = b + 1;
      acc = 0.f;
      accf = 0.f;
      continue;
    }
    if (b > start && (acc + e > target || accf + ef > fcap || b - start >= rblk)) {
      if (out) out[n] = UnitL{col, start, b};
      ++n;
      start = b;
      acc = 0.f;
      accf = 0.f;
    }
    acc += e;
    accf += ef;
  }
  if (start <= hi) {
    if (out) out[n] = UnitL{col, start, hi + 1};
    ++n;
  }
  return n;
}

__global__ void k_learned_units(const int64_t* __restrict__ flop, int64_t nzcB, int64_t small_max,
                                const float* __restrict__ logpred, const int32_t* __restrict__ blo,
                                const int32_t* __restrict__ bhi, const float* __restrict__ w, float target,
                                float heavy_pred, float fcap, int rblk, float direct_flop, int64_t m,
                                const int64_t* __restrict__ ubase, int64_t* __restrict__ ucount,
                                UnitL* __restrict__ units, float* __restrict__ pred_nnz,
                                uint8_t* __restrict__ heavy_flag) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= nzcB) return;
  if (!units) heavy_flag[j] = 0;
  if (flop[j] <= small_max || bhi[j] < 0) {
    if (!units) {
      ucount[j] = 0;
      pred_nnz[j] = 0.f;
    }
    return;
  }
  float pred = exp2f(logpred[j]) - 1.f;
  const float cap = (float)min<int64_t>(flop[j], m);
  pred = fminf(fmaxf(pred, 1.f), cap);
  if (pred > heavy_pred) {  // predicted heavy: exact symbolic and the large two-phase units
    if (!units) {
      heavy_flag[j] = 1;
      ucount[j] = 0;
      pred_nnz[j] = 0.f;
    }
    return;
  }
  if (!units) {
    ucount[j] = cut_units((int)j, pred, (float)flop[j], blo[j], bhi[j], w, target, fcap, rblk, direct_flop, nullptr);
    pred_nnz[j] = pred;
  } else {
    cut_units((int)j, pred, (float)flop[j], blo[j], bhi[j], w, target, fcap, rblk, direct_flop, units + ubase[j]);
  }
}

constexpr unsigned long long kAggFlag = 1ull << 62, kInclFlag = 2ull << 62, kValMask = (1ull << 62) - 1;

// One workgroup per unit, units taken in order through a ticket counter so
// that every predecessor of a unit is already resident when it waits on it.
template <int K, class T, int NT, int U, int KMAX, int RBLK, int CR>
__global__ __launch_bounds__(NT) void k_num_learned(
    const UnitL* __restrict__ units, int* __restrict__ ticket, unsigned long long* __restrict__ status,
    const int64_t* __restrict__ hcs, const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
    const T* __restrict__ valB, const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
    const T* __restrict__ valA, const int32_t* __restrict__ P, int S, int64_t cap, int64_t* __restrict__ irC,
    T* __restrict__ valC, unsigned long long* __restrict__ col_total, int* __restrict__ ovf_units,
    int* __restrict__ err) {
  using SR = Sem<K, T>;
  using Acc = typename SR::Acc;
  constexpr int NW32 = RBLK * kBRW;
  constexpr int NW64 = NW32 / 2;
  constexpr int WPT = (NW64 + NT - 1) / NT;
  __shared__ Acc vals[KMAX];
  __shared__ uint32_t bits[NW32];
  // exclusive prefix counts: 16 bits while 64 * NW64 <= 65536
  using PreT = typename std::conditional<(64 * NW64 <= 65536), uint16_t, int32_t>::type;
  __shared__ PreT pre[NW64];
  __shared__ SegStage<T, NT> st;
  __shared__ long long wsum[NT / kWave];
  __shared__ int s_t;
  __shared__ long long s_excl;
  if (threadIdx.x == 0) s_t = atomicAdd(ticket, 1);
  __syncthreads();
  const int t = s_t;
  const UnitL un = units[t];
  const int col = un.col, t_lo = un.blk_lo, t_hi = un.blk_hi;
  const int32_t row0 = t_lo * kBR;
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  const bool single_chunk = (b1 - b0) <= NT;
  const int nw32 = min(t_hi - t_lo, RBLK) * kBRW;
  for (int i = threadIdx.x; i < nw32; i += NT) bits[i] = 0u;

  auto set_bits = [&](const int64_t(&idx)[U], const T(&)[U]) {
    int32_t r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = idx[u] >= 0 ? rowsA[idx[u]] - row0 : -1;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r[u] >= 0) atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
  };
  auto rank_of = [&](int32_t r) {
    const int w = r >> 6, b = r & 63;
    const uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
    return (int)pre[w] + __popcll(word & ((1ull << b) - 1ull));
  };
  int q0 = 0;  // first rank of the current value window
  auto accumulate = [&](const int64_t(&idx)[U], const T(&bvs)[U]) {
    int32_t r[U];
    T a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = -1;
      if (idx[u] >= 0) {
        r[u] = rowsA[idx[u]] - row0;
        a[u] = valA[idx[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (r[u] < 0) continue;
      const int q = rank_of(r[u]) - q0;
      if (q >= 0 && q < KMAX) SR::acc(&vals[q], SR::mul(a[u], bvs[u]));
    }
  };

  int32_t crow[CR][U];
  Acc cprod[CR][U];
  int tot = 0, nseg = 0;
  const bool valid = (t_hi - t_lo) <= RBLK;
  if (!valid && threadIdx.x == 0) atomicOr(err, 64);
  // A unit whose rows fit the value window accumulates at its rows in a
  // single sweep (its size, known after that sweep, never exceeds the window)
  const bool direct = valid && (t_hi - t_lo) * kBR <= KMAX;
  const int ndirect = direct ? (t_hi - t_lo) * kBR : 0;
  if (direct) {
    for (int i = threadIdx.x; i < ndirect; i += NT) vals[i] = SR::ident();
    // (the staging barriers below order the window's initialisation before any accumulation)
    auto acc_direct = [&](const int64_t(&idx)[U], const T(&bvs)[U]) {
      int32_t r[U];
      T a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        r[u] = -1;
        if (idx[u] >= 0) {
          r[u] = rowsA[idx[u]] - row0;
          a[u] = valA[idx[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (r[u] >= 0) {
          atomicOr(&bits[r[u] >> 5], 1u << (r[u] & 31));
          SR::acc(&vals[r[u]], SR::mul(a[u], bvs[u]));
        }
    };
    for (int64_t eb = b0; eb < b1; eb += NT) {
      int64_t base;
      int len;
      T bv;
      load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
      tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
      visit_staged<T, NT, U>(st, nseg, tot, acc_direct);
    }
  } else if (single_chunk) {
    int64_t base;
    int len;
    T bv;
    load_seg<T, true>(b0 + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
    if (!valid) len = 0;
    tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
#pragma unroll
    for (int rr = 0; rr < CR; ++rr) {
      int64_t idx[U];
      T bvs[U];
      fill_round<T, NT, U>(st, nseg, tot, (rr * NT + (int)threadIdx.x) * U, idx, bvs);
      T a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        crow[rr][u] = -1;
        if (idx[u] >= 0) {
          crow[rr][u] = rowsA[idx[u]] - row0;
          a[u] = valA[idx[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (crow[rr][u] >= 0) {
          atomicOr(&bits[crow[rr][u] >> 5], 1u << (crow[rr][u] & 31));
          cprod[rr][u] = SR::mul(a[u], bvs[u]);
        }
    }
    visit_staged<T, NT, U>(st, nseg, tot, set_bits, CR);
  } else {
    for (int64_t eb = b0; eb < b1; eb += NT) {
      int64_t base;
      int len;
      T bv;
      load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
      if (!valid) len = 0;
      tot = stage_chunk<T, NT>(st, wsum, base, len, bv, nseg);
      visit_staged<T, NT, U>(st, nseg, tot, set_bits);
    }
  }
  __syncthreads();
  const int nw64 = nw32 / 2;
  int cnt;
  {
    int local[WPT];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      const int w = threadIdx.x * WPT + q;
      const int c = w < nw64 ? __popc(bits[2 * w]) + __popc(bits[2 * w + 1]) : 0;
      local[q] = sum;
      sum += c;
    }
    long long total;
    const int ex = (int)block_excl_scan_ll<NT>(sum, wsum, total);
#pragma unroll
    for (int q = 0; q < WPT; ++q) {
      const int w = threadIdx.x * WPT + q;
      if (w < nw64) pre[w] = (PreT)(ex + local[q]);
    }
    cnt = (int)total;
  }
  // publish the size, then look back for the offset with the whole first
  // wave: 64 predecessors per step, summing back to the nearest one that
  // holds an inclusive prefix (bounded wait on predecessors still counting).
  // A status word carries its flag and value together and no other data is
  // passed between units, so the atomics are relaxed: at agent scope an
  // acquire or release would invalidate or write back this XCD's whole L2.
  if (threadIdx.x < kWave) {
    const int lane = (int)threadIdx.x;
    if (lane == 0)
      __hip_atomic_store(&status[t], kAggFlag | (unsigned long long)cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long long excl = 0;
    int spins = 0;
    for (int i = t - 1; i >= 0;) {
      const int j = i - lane;
      const unsigned long long v =
          j >= 0 ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kInclFlag;
      const unsigned long long incl = __ballot((v >> 62) == 2);
      const unsigned long long none = __ballot((v >> 62) == 0);
      const int first = incl ? __ffsll((unsigned long long)incl) - 1 : kWave;  // nearest inclusive prefix
      const unsigned long long need = first >= kWave - 1 ? ~0ull : ((2ull << first) - 1ull);
      if (none & need) {  // a predecessor in range has not published yet
        if (++spins > (1 << 22)) {
          if (lane == 0) atomicOr(err, 256);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      excl += wave_sum_ll(lane <= first ? (long long)(v & kValMask) : 0ll);
      if (first < kWave) break;
      i -= kWave;
    }
    if (lane == 0) {
      __hip_atomic_store(&status[t], kInclFlag | (unsigned long long)(excl + cnt), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      s_excl = excl;
      if (cnt) atomicAdd(&col_total[col], (unsigned long long)cnt);
      if (cnt > KMAX) atomicAdd(ovf_units, 1);
    }
  }
  __syncthreads();
  const int64_t out = s_excl + hcs[col];
  const bool can_write = out + cnt <= cap;
  if (!can_write && threadIdx.x == 0) atomicOr(err, 512);
  if (direct) {  // values sit at their rows: one thread per row, stores stay coalesced
    if (can_write)
      for (int r = threadIdx.x; r < ndirect; r += NT) {
        if (!((bits[r >> 5] >> (r & 31)) & 1u)) continue;
        const int q = rank_of(r);
        irC[out + q] = (int64_t)row0 + r;
        valC[out + q] = SR::out(vals[r]);
      }
    return;
  }
  // values, in windows of KMAX ranks (one window unless the unit overflowed)
  for (q0 = 0; q0 < cnt; q0 += KMAX) {
    const int wn = min(KMAX, cnt - q0);
    for (int i = threadIdx.x; i < wn; i += NT) vals[i] = SR::ident();
    __syncthreads();
    if (single_chunk) {
#pragma unroll
      for (int rr = 0; rr < CR; ++rr)
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (crow[rr][u] >= 0) {
            const int q = rank_of(crow[rr][u]) - q0;
            if (q >= 0 && q < KMAX) SR::acc(&vals[q], cprod[rr][u]);
          }
      visit_staged<T, NT, U>(st, nseg, tot, accumulate, CR);
    } else {
      for (int64_t eb = b0; eb < b1; eb += NT) {
        int64_t base;
        int len;
        T bv;
        load_seg<T, true>(eb + threadIdx.x, b1, bpos, cpA, P, S, t_lo, t_hi, false, valB, base, len, bv);
        if (!valid) len = 0;
        int ns2;
        const int t2 = stage_chunk<T, NT>(st, wsum, base, len, bv, ns2);
        visit_staged<T, NT, U>(st, ns2, t2, accumulate);
      }
    }
    __syncthreads();
    if (can_write)
      for (int i = threadIdx.x; i < wn; i += NT) valC[out + q0 + i] = SR::out(vals[i]);
    __syncthreads();
  }
  if (can_write) {
    // rows in rank order; staged in the value array's LDS when they fit in
    // one window (the common case), so that they leave in coalesced lines
    const bool staged = cnt <= KMAX;
    uint32_t* rstage = reinterpret_cast<uint32_t*>(vals);
    for (int w = threadIdx.x; w < nw64; w += NT) {
      uint64_t word = (uint64_t)bits[2 * w] | ((uint64_t)bits[2 * w + 1] << 32);
      int q = pre[w];
      while (word) {
        const int b = __ffsll((unsigned long long)word) - 1;
        word &= word - 1;
        if (staged) rstage[q] = (uint32_t)(w * 64 + b);
        else irC[out + q] = (int64_t)row0 + w * 64 + b;
        ++q;
      }
    }
    if (staged) {
      __syncthreads();
      for (int i = threadIdx.x; i < cnt; i += NT) irC[out + i] = (int64_t)row0 + rstage[i];
    }
  }
}

// Tile pointers at kBR granularity, one wave per stored A column:
// P[p*(S+1)+t] = #entries of column p with row < t*kBR (binary search).
__global__ void k_ptable_wave(const int64_t* __restrict__ cpA, int64_t nzcA, const int32_t* __restrict__ rowsA,
                              int S, int32_t* __restrict__ P) {
  const int64_t p = global_wave();
  if (p >= nzcA) return;
  const int64_t b = cpA[p];
  const int len = (int)(cpA[p + 1] - b);
  const int32_t* rows = rowsA + b;
  int32_t* out = P + p * (int64_t)(S + 1);
  for (int t = lane_id(); t <= S; t += kWave) {
    const int64_t key = (int64_t)t * kBR;
    int lo = 0, hi = len;  // first i with rows[i] >= key
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (rows[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    out[t] = lo;
  }
}

// w[b] = hist[b] / nnzA
__global__ void k_hist_to_w(const unsigned long long* __restrict__ hist, int S, double inv, float* __restrict__ w) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < S) w[b] = (float)((double)hist[b] * inv);
}

__global__ void k_add_col_totals(int64_t* __restrict__ nnzcol, const unsigned long long* __restrict__ col_total,
                                 int64_t n) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < n) nnzcol[j] += (int64_t)col_total[j];
}

// logpred[j] = log2(1 + est[j]): an estimated column size in the network's output form.
__global__ void k_log2p1(const float* __restrict__ est, int64_t n, float* __restrict__ logpred) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j < n) logpred[j] = log2f(1.0f + fmaxf(est[j], 0.0f));
}

__global__ void k_sum_f32(const float* __restrict__ v, int64_t n, double* __restrict__ out) {
  double a = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a += v[i];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) a += __shfl_xor(a, d, kWave);
  if (lane_id() == 0) atomicAdd(out, a);
}

// ------------------------------------------------------------------ driver
namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipStream_t s = nullptr;
  ~DevBuf() { scratch_free(p, bytes, s); }
};

// Large per-call arrays whose size repeats from product to product (the unit
// lists) come from the caching allocator: a stream-pool allocation of several
// GB has, now and then, taken over a second when the pool had to grow.  The
// block returns to the cache at scope exit, once the stream's kernels that
// read it are done (the synchronisation is free on the normal path, which
// has already synchronised; an early error return waits for them).
struct CachedBuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  ~CachedBuf() {
    if (!p) return;
    (void)hipStreamSynchronize(s);
    dev_free(p);
  }
};

// Row maps of the symbolic pass (from the caching allocator): returned to
// the cache once the stream's unit kernels that read them are done.
struct RowMap {
  uint32_t* p = nullptr;
  hipStream_t s = nullptr;
  ~RowMap() {
    if (!p) return;
    (void)hipStreamSynchronize(s);
    dev_free(p);
  }
};

template <class X>
cbx_status calloc_cached(CachedBuf& b, size_t count, hipStream_t s, X** out) {
  b.s = s;
  const cbx_status st = dev_alloc(&b.p, std::max<size_t>(count * sizeof(X), 256));
  if (st != CBX_OK) {
    b.p = nullptr;
    return st;
  }
  *out = (X*)b.p;
  return CBX_OK;
}

template <class X>
cbx_status dalloc(DevBuf& b, size_t count, hipStream_t s, X** out) {
  b.s = s;
  b.bytes = std::max<size_t>(count * sizeof(X), 256);
  const cbx_status st = scratch_alloc(&b.p, b.bytes, s);
  if (st != CBX_OK) {
    b.p = nullptr;
    return st;
  }
  *out = (X*)b.p;
  return CBX_OK;
}

#define CBX_TRY(x)                 \
  do {                             \
    cbx_status st_ = (x);          \
    if (st_ != CBX_OK) return st_; \
  } while (0)

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Unit order for the numeric pass: by first row block, so that the units in
// flight at any moment gather the A entries of a narrow band of rows (a few
// hundred KB of A), which the XCDs' L2 caches can hold, instead of every row
// range of a few heavy columns.  Units are independent (their output offsets
// are known), so the order changes nothing but the cache footprint.
__global__ void k_unit_keys(const Unit* __restrict__ u, int n, int32_t* __restrict__ key, int32_t* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    key[i] = u[i].blk_lo;
    idx[i] = i;
  }
}

__global__ void k_unit_permute(const Unit* __restrict__ src, const int32_t* __restrict__ idx, int n,
                               Unit* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// A copy of units[0, n) sorted by blk_lo (stable), kept in `keep`.
__global__ void k_gather_flop(const int32_t* __restrict__ cols, int64_t n, const int64_t* __restrict__ flop,
                              int64_t* __restrict__ key) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) key[i] = flop[cols[i]];
}

// Reorders cols[0..n) by descending flop (into `keep`, returned in *out), so a
// per-column launch over them starts its longest columns first, as the
// two-phase path's flop-sorted order does.
cbx_status sort_cols_by_flop(const int32_t* cols, int64_t n, const int64_t* flop, hipStream_t s, DevBuf& keep,
                             int32_t** out) {
  DevBuf b_k, b_k2, b_t;
  int64_t *k1, *k2;
  uint8_t* t;
  CBX_TRY(dalloc(b_k, n, s, &k1));
  CBX_TRY(dalloc(b_k2, n, s, &k2));
  CBX_TRY(dalloc(keep, n, s, out));
  size_t tb = 0;
  CBX_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb, k1, k2, cols, *out, (int)n, 0, 64, s));
  CBX_TRY(dalloc(b_t, std::max<size_t>(tb, 16), s, &t));
  k_gather_flop<<<nblk(n, 256), 256, 0, s>>>(cols, n, flop, k1);
  CBX_HIP(hipcub::DeviceRadixSort::SortPairsDescending(t, tb, k1, k2, cols, *out, (int)n, 0, 64, s));
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}

cbx_status sort_units_by_block(const Unit* units, int n, int S, hipStream_t s, DevBuf& keep, Unit** out) {
  DevBuf b_k, b_k2, b_i, b_i2, b_t;
  int32_t *k1, *k2, *i1, *i2;
  uint8_t* t;
  CBX_TRY(dalloc(b_k, n, s, &k1));
  CBX_TRY(dalloc(b_k2, n, s, &k2));
  CBX_TRY(dalloc(b_i, n, s, &i1));
  CBX_TRY(dalloc(b_i2, n, s, &i2));
  CBX_TRY(dalloc(keep, n, s, out));
  int bits = 1;
  while ((1 << bits) <= S) ++bits;
  size_t tb = 0;
  CBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k1, k2, i1, i2, n, 0, bits, s));
  CBX_TRY(dalloc(b_t, std::max<size_t>(tb, 16), s, &t));
  k_unit_keys<<<nblk(n, 256), 256, 0, s>>>(units, n, k1, i1);
  CBX_HIP(hipcub::DeviceRadixSort::SortPairs(t, tb, k1, k2, i1, i2, n, 0, bits, s));
  k_unit_permute<<<nblk(n, 256), 256, 0, s>>>(units, i2, n, *out);
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}

struct Timer {
  hipEvent_t ev[6];
  bool on;
  explicit Timer(bool enable) : on(enable) {
    if (on)
      for (auto& e : ev) (void)hipEventCreate(&e);
  }
  ~Timer() {
    if (on)
      for (auto& e : ev) (void)hipEventDestroy(e);
  }
  void mark(int i, hipStream_t s) {
    if (on) (void)hipEventRecord(ev[i], s);
  }
  double ms(int a, int b) {
    float f = 0;
    if (on) (void)hipEventElapsedTime(&f, ev[a], ev[b]);
    return f;
  }
};

constexpr int kHashNT = 256;
constexpr int kDenseNT = 512;                 // dense numeric threads
constexpr int kDenseTR = 32 * kDenseNT;       // rows per numeric tile (16384)
constexpr int kSymNT = 512;
constexpr int kSymRBW = 32768;                // bitmap words per symbolic pass (2^20 rows)
// unit tier
constexpr int kSymUnitNT = 512, kSymUnitIBLK = 64;  // 32 KB bitmap: three workgroups per CU
constexpr int kSymBandNT = 256;                     // banded symbolic pass
// light units span at most 16 blocks: ~53 KB of LDS, three workgroups per CU
constexpr int kLightNT = 512, kLightK = 4096, kLightRB = 16;    // <= 4096 outputs, <= 64K rows
constexpr int kHeavyNT = 1024, kHeavyK = 16384, kHeavyRB = 16;  // <= 16384 outputs, <= 64K rows
// The large units' value window: 16384 outputs, or 12288 for 4-byte
// accumulators, whose large-unit kernel then takes about 76 KB of LDS
// instead of 92 KB: two workgroups per CU, eight waves per SIMD, not four.
template <int K, class T>
constexpr int heavy_k() {
  return sizeof(typename Sem<K, T>::Acc) <= 4 ? 12288 : kHeavyK;
}
// learned units: one workgroup shape for all (512 threads, 4096-entry value
// window, 32 row blocks), cut to target * 4096 predicted outputs
constexpr int kLearnNT = 512, kLearnK = 4096, kLearnRB = 32, kLearnCR = 2;
constexpr int kLearnFlopCap = 131072;  // expected products per learned unit
constexpr int64_t kLearnChunk = 1 << 21;  // workgroups per learned launch
// Tiny tier launch: `waves` columns (one wave each) per workgroup, grid-stride
// beyond 256K columns in flight.
inline unsigned tiny_grid(int64_t ncols, int waves) {
  return (unsigned)std::min<int64_t>((ncols + waves - 1) / waves, (int64_t)262144 / waves);
}

// The tiny tier in the configured form (register sort or LDS table) and
// workgroup shape.
template <int K, class T, bool SYM>
void launch_tiny(const SpgemmConfig& cfg, const int32_t* order, int64_t first, int64_t ncols, const int64_t* cpB,
                 const int32_t* bpos, const T* valB, const int64_t* cpA, const int32_t* rowsA, const T* valA,
                 int64_t* nnzcol, const int64_t* cpC, int64_t* irC, T* valC, int* err, const int64_t* stage_off,
                 hipStream_t s, const Filter<T>& flt = {}) {
  const dim3 g(tiny_grid(ncols, cfg.tiny_wg_waves)), b(kWave * cfg.tiny_wg_waves);
  if (cfg.tiny_algo == 1)
    k_tiny_h<K, T, SYM><<<g, b, 0, s>>>(order, first, ncols, cpB, bpos, valB, cpA, rowsA, valA, nnzcol, cpC, irC,
                                        valC, err, stage_off, flt);
  else
    k_tiny<K, T, SYM><<<g, b, 0, s>>>(order, first, ncols, cpB, bpos, valB, cpA, rowsA, valA, nnzcol, cpC, irC, valC,
                                      err, stage_off, flt);
}

// Row lists of the small units (k_sym_units); lists == nullptr: none.
struct UnitLists {
  uint16_t* lists = nullptr;
  int64_t* off = nullptr;  // per symbolic slot, ncols + 1 entries
  uint16_t* bcnt = nullptr;
  Unit* units = nullptr;
  int* n = nullptr;
  int cap = 0;
};

// Region of symbolic slot i in the row lists: the column's outputs in the
// small-unit passes are at most min(flop_j, unit_split_nnz, m).  Slot i is
// column order[i] (flop indexed by column), or column i of flop when order
// is null (flop already in slot order).
__global__ void k_list_regions(const int32_t* __restrict__ order, const int64_t* __restrict__ flop, int64_t n,
                               int64_t cap, int64_t* __restrict__ sz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const int64_t f = flop[order ? order[i] : i];
    sz[i] = f < cap ? f : cap;
  } else if (i == n) {
    sz[i] = 0;
  }
}

// Region offsets, row lists, block counts and list-unit array for the
// symbolic slots order[0..n) (order null: slot i is flop[i]), when they fit a
// fifth of the free memory; otherwise ul.lists stays null.  The buffers are
// owned by the caller's DevBufs / CachedBuf.
static cbx_status alloc_unit_lists(const int32_t* order, const int64_t* flop, int64_t n, int S, int64_t cap_units,
                                   int64_t region_cap, int* counter, void* tmp, size_t tmp_bytes, DevBuf& b_lsz,
                                   DevBuf& b_loff, DevBuf& b_lists, DevBuf& b_lbc, CachedBuf& b_us, hipStream_t s,
                                   UnitLists& ul) {
  ul = UnitLists{};
  int64_t *lsz, *off;
  CBX_TRY(dalloc(b_lsz, n + 1, s, &lsz));
  CBX_TRY(dalloc(b_loff, n + 1, s, &off));
  k_list_regions<<<nblk(n + 1, 256), 256, 0, s>>>(order, flop, n, region_cap, lsz);
  size_t need = 0;
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, lsz, off, (int)(n + 1), s));
  DevBuf b_t3;
  void* t3 = tmp;
  if (need > tmp_bytes) CBX_TRY(dalloc(b_t3, need, s, (uint8_t**)&t3));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(t3, need, lsz, off, (int)(n + 1), s));
  int64_t total = 0;
  CBX_HIP(hipMemcpyAsync(&total, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  size_t freeb = 0, totb = 0;
  const double bytes = 2.0 * (double)total + 2.0 * (double)n * S + (double)sizeof(Unit) * cap_units;
  uint16_t *lists = nullptr, *bcnt = nullptr;
  Unit* units = nullptr;
  if (hipMemGetInfo(&freeb, &totb) == hipSuccess && bytes <= 0.2 * (double)(freeb + cache_idle_bytes()) &&
      dalloc(b_lists, std::max<int64_t>(total, 1), s, &lists) == CBX_OK &&
      dalloc(b_lbc, n * (int64_t)S, s, &bcnt) == CBX_OK && calloc_cached(b_us, cap_units, s, &units) == CBX_OK) {
    ul.lists = lists;
    ul.off = off;
    ul.bcnt = bcnt;
    ul.units = units;
    ul.n = counter;
    ul.cap = (int)cap_units;
  }
  (void)hipGetLastError();
  return CBX_OK;
}

// The unit tier's symbolic pass: per column (sym_band 0, optionally with the
// small units' row lists) or banded (the band height in row blocks, 16 or 64;
// bcnt_buf receives the block counts).
cbx_status launch_sym_units(const SpgemmConfig& cfg, const int32_t* order, int64_t ncols, const cbx_dcsc* A,
                            const cbx_dcsc* B, const int32_t* bpos, const int32_t* rowsA, const int32_t* P, int S,
                            int64_t* nnzcol, Unit* units_l, int* n_l, int cap_l, Unit* units_h, int* n_h, int cap_h,
                            uint32_t* rmap, int64_t rm_cols, int* err, DevBuf& bcnt_buf, int kmax_h, hipStream_t s,
                            const UnitLists& ul = {}) {
  if (cfg.sym_band == 0) {
    k_sym_units<kSymUnitNT, 4, kSymUnitIBLK><<<(unsigned)ncols, kSymUnitNT, 0, s>>>(
        order, 0, B->cp, bpos, A->cp, rowsA, P, S, kLightK, kLightRB, kmax_h, kHeavyRB, cfg.unit_split_nnz, nnzcol,
        units_l, n_l, cap_l, units_h, n_h, cap_h, rmap, rm_cols, err, ul.lists, ul.off, ul.bcnt, ul.units, ul.n,
        ul.cap);
    return CBX_OK;
  }
  int32_t* bcnt;
  CBX_TRY(dalloc(bcnt_buf, (size_t)ncols * S, s, &bcnt));
  const int band = cfg.sym_band;
  const dim3 g((unsigned)ncols, (unsigned)((S + band - 1) / band));  // band-major dispatch order
  if (band == 16)
    k_sym_band<kSymBandNT, 4, 16><<<g, kSymBandNT, 0, s>>>(order, 0, B->cp, bpos, A->cp, rowsA, P, S, bcnt, rmap,
                                                          rm_cols);
  else
    k_sym_band<kSymBandNT, 4, 64><<<g, kSymBandNT, 0, s>>>(order, 0, B->cp, bpos, A->cp, rowsA, P, S, bcnt, rmap,
                                                          rm_cols);
  k_sym_cut<<<(unsigned)ncols, 256, 0, s>>>(order, 0, bcnt, S, kLightK, kLightRB, kmax_h, kHeavyRB,
                                            cfg.unit_split_nnz, nnzcol, units_l, n_l, cap_l, units_h, n_h, cap_h, err);
  return CBX_OK;
}

}  // namespace

struct LearnedInfo {
  int64_t nnzC = 0, units = 0, overflow_units = 0, hash_total = 0, heavy_units = 0;
  int64_t exact_total = 0;  // outputs of the predicted-heavy columns (computed exactly)
  double pred_sum = 0;
  float kernel_ms = 0;  // the single-pass numeric kernel alone
  bool over_capacity = false;
};

// Learned single pass for the unit-tier columns (see the learned tier above).
// On entry nnzcol holds the exact sizes of the hash-tier columns (0 for the
// rest).  On success C->ir / C->val are allocated and the unit-tier outputs
// written, nnzcol holds every column's size and cpC its exclusive scan.
template <int K, class T>
static cbx_status learned_numeric(const cbx_dcsc* A, const cbx_dcsc* B, const SpgemmConfig& cfg, hipStream_t s,
                                  const int32_t* rowsA, const int32_t* bpos, const int64_t* flop, const int32_t* P,
                                  int S, int64_t* nnzcol, int64_t* cpC, void* tmp, size_t tmp_bytes, cbx_dcsc* C,
                                  int* err, LearnedInfo& info) {
  const int64_t nzcB = B->nzc, m = A->m;
  DevBuf b_feat, b_lp, b_hist, b_w, b_blo, b_bhi, b_uc, b_ub, b_pn, b_units, b_status, b_tick, b_ct, b_ovf, b_hcs, b_sum,
      b_hflag, b_hcols, b_nh, b_hu;
  CachedBuf b_ul, b_uh, b_us;
  DevBuf b_bcnt, b_lsz, b_loff, b_lists, b_lbc;
  UnitLists ul;
  uint8_t* heavy_flag;
  int32_t* hcols;
  int64_t* nh_dev;
  int* hucnt;
  Unit *units_l = nullptr, *units_h = nullptr;
  float *feats, *logpred, *w, *pred_nnz;
  unsigned long long *hist, *status, *col_total;
  int32_t *blo, *bhi;
  int64_t *ucount, *ubase, *hcs;
  int *ticket, *ovf;
  double* dsum;
  UnitL* units;
  CBX_TRY(dalloc(b_feat, nzcB * CBX_MLP_FEATURES, s, &feats));
  CBX_TRY(dalloc(b_lp, nzcB, s, &logpred));
  CBX_TRY(dalloc(b_hist, S, s, &hist));
  CBX_TRY(dalloc(b_w, S, s, &w));
  CBX_TRY(dalloc(b_blo, nzcB, s, &blo));
  CBX_TRY(dalloc(b_bhi, nzcB, s, &bhi));
  CBX_TRY(dalloc(b_uc, nzcB + 1, s, &ucount));
  CBX_TRY(dalloc(b_ub, nzcB + 1, s, &ubase));
  CBX_TRY(dalloc(b_pn, nzcB, s, &pred_nnz));
  CBX_TRY(dalloc(b_hcs, nzcB + 1, s, &hcs));
  CBX_TRY(dalloc(b_sum, 2, s, &dsum));
  CBX_TRY(dalloc(b_tick, 2, s, &ticket));
  CBX_TRY(dalloc(b_ovf, 2, s, &ovf));
  CBX_TRY(dalloc(b_ct, nzcB, s, &col_total));
  CBX_TRY(dalloc(b_hflag, nzcB, s, &heavy_flag));
  CBX_TRY(dalloc(b_hcols, nzcB, s, &hcols));
  CBX_TRY(dalloc(b_nh, 1, s, &nh_dev));
  CBX_TRY(dalloc(b_hu, 3, s, &hucnt));
  // prediction: the network on the columns' features, or (without it) the
  // min-hash estimates of estimateNNZ_sampling
  if (cfg.sampling_rounds > 0) {
    DevBuf b_samp, b_est;
    float *samp, *est;
    CBX_TRY(dalloc(b_samp, sampling_scratch_floats(A, cfg.sampling_rounds), s, &samp));
    CBX_TRY(dalloc(b_est, nzcB, s, &est));
    CBX_TRY(sampling_estimate_launch(A, B, cfg.sampling_rounds, rowsA, bpos, samp, est, s));
    k_log2p1<<<nblk(nzcB, 256), 256, 0, s>>>(est, nzcB, logpred);
  } else {
    k_features<<<wave_grid(nzcB), 256, 0, s>>>(B->cp, nzcB, bpos, A->cp, rowsA, m, feats);
    mlp_infer_launch(cfg.predictor, feats, nzcB, logpred, s);
  }
  // block prior and per-column block range
  CBX_HIP(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * S, s));
  k_row_hist<<<256, 256, sizeof(unsigned int) * S, s>>>(rowsA, A->nnz, S, hist);
  k_hist_to_w<<<nblk(S, 256), 256, 0, s>>>(hist, S, 1.0 / (double)std::max<int64_t>(A->nnz, 1), w);
  k_col_blocks<<<wave_grid(nzcB), 256, 0, s>>>(B->cp, nzcB, bpos, A->cp, rowsA, blo, bhi);
  // units in column order: count, scan, write
  const float target = cfg.learned_target * (float)kLearnK;
  const float heavy_pred = (float)cfg.heavy_nnz;
  const float fcap = (float)kLearnFlopCap;
  // a block expected to hold more products than a unit keeps in registers
  // between its sweeps becomes a single-block unit, accumulated in one sweep
  const float direct_flop = (float)(kLearnCR * 4 * kLearnNT);
  k_learned_units<<<nblk(nzcB, 256), 256, 0, s>>>(flop, nzcB, cfg.small_flop_max, logpred, blo, bhi, w, target,
                                                  heavy_pred, fcap, kLearnRB, direct_flop, m, nullptr, ucount,
                                                  nullptr, pred_nnz, heavy_flag);
  CBX_HIP(hipMemsetAsync(ucount + nzcB, 0, sizeof(int64_t), s));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, ucount, ubase, (int)(nzcB + 1), s));
  // columns predicted heavy: exact sizes from the symbolic unit pass (their
  // units then go through the two-phase unit kernels), so the single pass
  // covers the many light columns and its look-back sees the heavy columns'
  // sizes up front, like the hash tier's
  int64_t hash_only = 0;
  {
    size_t need = 0, need2 = 0;
    CBX_HIP(hipcub::DeviceReduce::Sum(nullptr, need, nnzcol, nh_dev, (int)nzcB, s));
    CBX_HIP(hipcub::DeviceSelect::Flagged(nullptr, need2, hipcub::CountingInputIterator<int32_t>(0), heavy_flag,
                                          hcols, nh_dev, (int)nzcB, s));
    need = std::max(need, need2);
    DevBuf b_tmp2;
    void* t2 = tmp;
    if (need > tmp_bytes) CBX_TRY(dalloc(b_tmp2, need, s, (uint8_t**)&t2));
    const size_t tb = std::max(need, tmp_bytes);
    size_t tb1 = tb, tb2 = tb;
    CBX_HIP(hipcub::DeviceReduce::Sum(t2, tb1, nnzcol, nh_dev, (int)nzcB, s));
    CBX_HIP(hipMemcpyAsync(&hash_only, nh_dev, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipcub::DeviceSelect::Flagged(t2, tb2, hipcub::CountingInputIterator<int32_t>(0), heavy_flag, hcols,
                                          nh_dev, (int)nzcB, s));
    CBX_HIP(hipStreamSynchronize(s));  // b_tmp2 is released at scope exit
  }
  int64_t nheavy = 0;
  CBX_HIP(hipMemcpyAsync(&nheavy, nh_dev, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  int64_t cap_l = 0, cap_h = 0;
  // row maps of the exact columns, as in the two-phase path: their large
  // units then load them instead of a counting sweep (when they fit the
  // share of free memory; from the caching allocator)
  RowMap rm;
  rm.s = s;
  int64_t rm_cols = 0;
  DevBuf b_hsorted;
  if (nheavy > 1) {  // longest columns first
    int32_t* hs = nullptr;
    CBX_TRY(sort_cols_by_flop(hcols, nheavy, flop, s, b_hsorted, &hs));
    hcols = hs;
  }
  if (nheavy) {
    cap_l = std::max<int64_t>(1, std::min<int64_t>(nheavy * (int64_t)S, 1ll << 30));
    cap_h = cap_l;
    CBX_TRY(calloc_cached(b_ul, cap_l, s, &units_l));
    CBX_TRY(calloc_cached(b_uh, cap_h, s, &units_h));
    CBX_HIP(hipMemsetAsync(hucnt, 0, 3 * sizeof(int), s));
    if (cfg.row_map_share > 0.0f) {
      const int64_t rm_words = nheavy * (int64_t)S * kBRW;
      size_t freeb = 0, totb = 0;
      if (hipMemGetInfo(&freeb, &totb) == hipSuccess &&
          4.0 * (double)rm_words <= (double)cfg.row_map_share * (double)(freeb + cache_idle_bytes()) &&
          dev_alloc((void**)&rm.p, sizeof(uint32_t) * rm_words) == CBX_OK)
        rm_cols = nheavy;
      (void)hipGetLastError();
    }
    // the small units' row lists, as in the two-phase path
    if (cfg.unit_lists && cfg.sym_band == 0)
      CBX_TRY(alloc_unit_lists(hcols, flop, nheavy, S, cap_l, std::min<int64_t>(cfg.unit_split_nnz, m), hucnt + 2, tmp,
                               tmp_bytes, b_lsz, b_loff, b_lists, b_lbc, b_us, s, ul));
    CBX_TRY(launch_sym_units(cfg, hcols, nheavy, A, B, bpos, rowsA, P, S, nnzcol, units_l, hucnt, (int)cap_l, units_h,
                             hucnt + 1, (int)cap_h, rm.p, rm_cols, err, b_bcnt, heavy_k<K, T>(), s, ul));
    CBX_HIP(hipGetLastError());
  }
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, nnzcol, hcs, (int)(nzcB + 1), s));
  CBX_HIP(hipMemsetAsync(dsum, 0, sizeof(double), s));
  k_sum_f32<<<256, 256, 0, s>>>(pred_nnz, nzcB, dsum);
  int64_t nunits = 0, hash_total = 0;
  double pred_sum = 0;
  CBX_HIP(hipMemcpyAsync(&nunits, ubase + nzcB, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&hash_total, hcs + nzcB, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&pred_sum, dsum, sizeof(double), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  info.units = nunits;
  info.pred_sum = pred_sum;
  info.hash_total = hash_only;
  info.exact_total = hash_total - hash_only;
  CBX_TRY(dalloc(b_units, std::max<int64_t>(nunits, 1), s, &units));
  CBX_TRY(dalloc(b_status, std::max<int64_t>(nunits, 1), s, &status));
  if (nunits)
    k_learned_units<<<nblk(nzcB, 256), 256, 0, s>>>(flop, nzcB, cfg.small_flop_max, logpred, blo, bhi, w, target,
                                                    heavy_pred, fcap, kLearnRB, direct_flop, m, ubase, nullptr,
                                                    units, nullptr, nullptr);
  // output capacity from the predictions (plus the exact hash-tier sizes)
  const int64_t cap = hash_total + (int64_t)(pred_sum * (1.0 + cfg.capacity_margin)) + (1 << 20);
  auto alloc_out = [&]() -> cbx_status {
    if (!C->ir) CBX_TRY(dev_alloc((void**)&C->ir, sizeof(int64_t) * cap));
    if (!C->val) CBX_TRY(dev_alloc(&C->val, sizeof(T) * cap));
    return CBX_OK;
  };
  cbx_status ast = alloc_out();
  if (ast != CBX_OK && rm.p) {  // give the row maps back; the large units then count in a sweep
    CBX_HIP(hipStreamSynchronize(s));
    dev_free(rm.p);
    rm.p = nullptr;
    release_cache();
    ast = alloc_out();
  }
  if (ast == CBX_ERR_NOMEM) {  // predicted capacity does not fit: the exact two-phase path sizes the output
    info.over_capacity = true;
    return CBX_OK;
  }
  CBX_TRY(ast);
  CBX_HIP(hipMemsetAsync(status, 0, sizeof(unsigned long long) * std::max<int64_t>(nunits, 1), s));
  CBX_HIP(hipMemsetAsync(ticket, 0, sizeof(int) * 2, s));
  CBX_HIP(hipMemsetAsync(ovf, 0, sizeof(int) * 2, s));
  CBX_HIP(hipMemsetAsync(col_total, 0, sizeof(unsigned long long) * nzcB, s));
  hipEvent_t kev[3] = {nullptr, nullptr, nullptr};
  for (auto& e : kev) (void)hipEventCreate(&e);
  (void)hipEventRecord(kev[0], s);
  // launches of at most 2^21 workgroups (a dispatch holds fewer than 2^32
  // work-items); the ticket counter runs on across them, and a later launch's
  // units look back into statuses the earlier launches completed
  for (int64_t done = 0; done < nunits; done += kLearnChunk) {
    const int64_t nb = std::min<int64_t>(kLearnChunk, nunits - done);
    k_num_learned<K, T, kLearnNT, 4, kLearnK, kLearnRB, kLearnCR><<<(unsigned)nb, kLearnNT, 0, s>>>(
        units, ticket, status, hcs, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, cap, C->ir,
        (T*)C->val, col_total, ovf, err);
  }
  (void)hipEventRecord(kev[1], s);
  CBX_HIP(hipGetLastError());
  k_add_col_totals<<<nblk(nzcB, 256), 256, 0, s>>>(nnzcol, col_total, nzcB);
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, nnzcol, cpC, (int)(nzcB + 1), s));
  int herr = 0, hovf = 0;
  CBX_HIP(hipMemcpyAsync(&info.nnzC, cpC + nzcB, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  info.overflow_units = hovf;
  (void)hipEventElapsedTime(&info.kernel_ms, kev[0], kev[1]);
  info.over_capacity = (herr & 512) != 0 || info.nnzC > cap;
  if (herr & ~512 || info.over_capacity || !nheavy) {
    for (auto& e : kev) (void)hipEventDestroy(e);
    if (herr & ~512) return set_error(CBX_ERR_HIP, "cbx_spgemm: learned pass inconsistency (flag %d)", herr);
    return CBX_OK;
  }
  // the heavy columns' values, at the offsets the scan above fixed
  int hu[3] = {0, 0, 0};  // small (bitmap), large, and small list units
  CBX_HIP(hipMemcpy(hu, hucnt, sizeof(hu), hipMemcpyDeviceToHost));
  (void)hipEventRecord(kev[2], s);  // numeric of the heavy columns, after the offsets
  CBX_RETURN_IF(hu[0] > cap_l || hu[1] > cap_h || hu[2] > cap_l, CBX_ERR_HIP, "cbx_spgemm: unit list overflow");
  CBX_RETURN_IF((int64_t)hu[0] * kLightNT >= (1ll << 32) || (int64_t)hu[1] * kHeavyNT >= (1ll << 32) ||
                    (int64_t)hu[2] * kLightNT >= (1ll << 32),
                CBX_ERR_LIMIT, "cbx_spgemm: too many units for a single launch");
  info.heavy_units = hu[0] + hu[1] + hu[2];
  DevBuf b_ul_sorted, b_uh_sorted, b_us_sorted;
  if (cfg.unit_order == 1) {
    if (hu[0] > 1) CBX_TRY(sort_units_by_block(units_l, hu[0], S, s, b_ul_sorted, &units_l));
    if (hu[1] > 1) CBX_TRY(sort_units_by_block(units_h, hu[1], S, s, b_uh_sorted, &units_h));
    if (hu[2] > 1) CBX_TRY(sort_units_by_block(ul.units, hu[2], S, s, b_us_sorted, &ul.units));
  }
  if (hu[1] && rm.p)
    k_num_units_rm<K, T, kHeavyNT, 4, heavy_k<K, T>(), kHeavyRB><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
        units_h, rm.p, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, rm_cols, cpC, C->ir,
        (T*)C->val, err, Filter<T>{}, cfg.unit_xcd_group);
  else if (hu[1])
    k_num_units<K, T, kHeavyNT, 4, heavy_k<K, T>(), kHeavyRB, 4><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
        units_h, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, cpC, C->ir, (T*)C->val, err,
        Filter<T>{}, cfg.unit_xcd_group);
  if (hu[0])
    k_num_units<K, T, kLightNT, 4, kLightK, kLightRB, 2><<<(unsigned)hu[0], kLightNT, 0, s>>>(
        units_l, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, cpC, C->ir, (T*)C->val, err,
        Filter<T>{}, cfg.unit_xcd_group);
  if (hu[2])
    k_num_list<K, T, kLightNT, 4, kLightK><<<(unsigned)hu[2], kLightNT, 0, s>>>(
        ul.units, ul.lists, ul.off, ul.bcnt, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, cpC,
        C->ir, (T*)C->val, err, Filter<T>{}, cfg.unit_xcd_group);
  CBX_HIP(hipGetLastError());
  (void)hipEventRecord(kev[1], s);
  CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  float heavy_ms = 0;
  (void)hipEventElapsedTime(&heavy_ms, kev[2], kev[1]);
  info.kernel_ms += heavy_ms;
  for (auto& e : kev) (void)hipEventDestroy(e);
  if (herr) return set_error(CBX_ERR_HIP, "cbx_spgemm: learned pass, large columns inconsistent (flag %d)", herr);
  return CBX_OK;
}

// Gather of the fused threshold's runs: run k (in start order) moves its kept
// entries from its start in the unfiltered layout to dst[k].
template <class T>
__global__ __launch_bounds__(256) void k_run_copy(const int32_t* __restrict__ idx, const int64_t* __restrict__ src,
                                                  const int32_t* __restrict__ kept, const int64_t* __restrict__ dst,
                                                  const int64_t* __restrict__ ir, const T* __restrict__ val,
                                                  int64_t* __restrict__ ir2, T* __restrict__ val2) {
  const int i = idx[blockIdx.x];
  const int64_t a = src[i], d = dst[blockIdx.x];
  const int n = kept[i];
  for (int e = threadIdx.x; e < n; e += 256) {
    ir2[d + e] = ir[a + e];
    val2[d + e] = val[a + e];
  }
}

__global__ void k_run_sizes(const int32_t* __restrict__ idx, const int32_t* __restrict__ kept,
                            const int32_t* __restrict__ col, int n, int64_t* __restrict__ kept_sorted,
                            unsigned long long* __restrict__ colkept) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = idx[k];
  kept_sorted[k] = kept[i];
  atomicAdd(&colkept[col[i]], (unsigned long long)kept[i]);
}

__global__ void k_iota32(int32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// The product after a fused threshold: the runs the numeric kernels appended
// (sorted by start, which is column then row order in the unfiltered layout)
// gathered into new arrays; nnzcol and cpC become the kept counts and their
// scan.  C keeps its other fields.
template <class T>
static cbx_status filter_compact(const Filter<T>& flt, int64_t nzcB, int64_t* nnzcol, int64_t* cpC, void* tmp,
                                 size_t tmp_bytes, cbx_dcsc* C, hipStream_t s) {
  int nrun = 0;
  CBX_HIP(hipMemcpyAsync(&nrun, flt.run_count, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  CBX_RETURN_IF(nrun > flt.run_cap, CBX_ERR_HIP, "cbx_spgemm: run list overflow (%d > %d)", nrun, flt.run_cap);
  DevBuf b_k2, b_i, b_i2, b_ks, b_dst, b_t;
  int64_t *k2, *ks, *dst;
  int32_t *i1, *i2;
  uint8_t* t;
  const int nr = std::max(nrun, 1);
  CBX_TRY(dalloc(b_k2, nr, s, &k2));
  CBX_TRY(dalloc(b_i, nr, s, &i1));
  CBX_TRY(dalloc(b_i2, nr, s, &i2));
  CBX_TRY(dalloc(b_ks, nr + 1, s, &ks));
  CBX_TRY(dalloc(b_dst, nr + 1, s, &dst));
  size_t tb = 0, tb2 = 0;
  CBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, flt.run_src, k2, i1, i2, nr, 0, 64, s));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, ks, dst, nr + 1, s));
  tb = std::max(tb, tb2);
  CBX_TRY(dalloc(b_t, std::max<size_t>(tb, 16), s, &t));
  CBX_HIP(hipMemsetAsync(nnzcol, 0, sizeof(int64_t) * (nzcB + 1), s));
  int64_t total = 0;
  if (nrun) {
    k_iota32<<<nblk(nrun, 256), 256, 0, s>>>(i1, nrun);
    size_t tb1 = tb;
    CBX_HIP(hipcub::DeviceRadixSort::SortPairs(t, tb1, flt.run_src, k2, i1, i2, nrun, 0, 64, s));
    k_run_sizes<<<nblk(nrun, 256), 256, 0, s>>>(i2, flt.run_kept, flt.run_col, nrun, ks,
                                                (unsigned long long*)nnzcol);
    CBX_HIP(hipMemsetAsync(ks + nrun, 0, sizeof(int64_t), s));
    size_t tb3 = tb;
    CBX_HIP(hipcub::DeviceScan::ExclusiveSum(t, tb3, ks, dst, nrun + 1, s));
    CBX_HIP(hipMemcpyAsync(&total, dst + nrun, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipStreamSynchronize(s));
  }
  int64_t* ir2 = nullptr;
  void* val2 = nullptr;
  CBX_TRY(dev_alloc((void**)&ir2, sizeof(int64_t) * std::max<int64_t>(total, 1)));
  if (dev_alloc(&val2, sizeof(T) * std::max<int64_t>(total, 1)) != CBX_OK) {
    dev_free(ir2);
    return set_error(CBX_ERR_NOMEM, "cbx_spgemm: allocation of the thresholded product failed");
  }
  if (nrun) k_run_copy<T><<<(unsigned)nrun, 256, 0, s>>>(i2, flt.run_src, flt.run_kept, dst, C->ir, (const T*)C->val,
                                                        ir2, (T*)val2);
  size_t tb4 = tmp_bytes;
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb4, nnzcol, cpC, (int)(nzcB + 1), s));
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipStreamSynchronize(s));
  dev_free(C->ir);
  dev_free(C->val);
  C->ir = ir2;
  C->val = val2;
  C->nnz = total;
  return CBX_OK;
}

template <int K, class T>
static cbx_status spgemm_typed_once(const cbx_dcsc* A, const cbx_dcsc* B, cbx_spgemm_mode mode,
                                    const SpgemmConfig& cfg, cbx_dcsc* C, cbx_spgemm_stats* stats, hipStream_t s,
                                    int64_t* symbolic_out, bool* lists_used) {
  const int64_t m = A->m, n = B->n;
  const int64_t nzcA = A->nzc, nzcB = B->nzc, nnzA = A->nnz, nnzB = B->nnz;
  Timer tm(stats != nullptr);
  tm.mark(0, s);

  std::memset(C, 0, sizeof(*C));
  C->m = m;
  C->n = n;
  C->dtype = A->dtype;
  C->location = CBX_DEVICE;
  if (stats) std::memset(stats, 0, sizeof(*stats));

  if (nnzA == 0 || nnzB == 0 || nzcA == 0 || nzcB == 0) {  // mtSpGEMM.h:224-227
    if (symbolic_out) {
      if (nzcB) CBX_HIP(hipMemsetAsync(symbolic_out, 0, sizeof(int64_t) * nzcB, s));
      CBX_HIP(hipStreamSynchronize(s));
      return CBX_OK;
    }
    CBX_TRY(dev_alloc((void**)&C->cp, sizeof(int64_t)));
    CBX_HIP(hipMemsetAsync(C->cp, 0, sizeof(int64_t), s));
    CBX_HIP(hipStreamSynchronize(s));
    return CBX_OK;
  }

  // Unit tier needs the tile-pointer table at kBR-row granularity; it is used
  // when that table stays modest, otherwise the streaming dense tier runs.
  const int S4 = (int)((m + kBR - 1) / kBR);
  const bool use_units = m <= (1ll << 24) && (double)nzcA * (S4 + 1) * 4.0 <= 8.0e9;
  // the fused threshold (two-phase, unit or hash tiers; the streaming tier has no filtered emission)
  Filter<T> flt;
  flt.on = cfg.filter && mode == CBX_SPGEMM_TWO_PHASE && use_units && symbolic_out == nullptr;
  flt.thr = (T)cfg.filter_thr;
  DevBuf b_drop;
  if (flt.on) {
    CBX_TRY(dalloc(b_drop, 1, s, &flt.dropped));
    CBX_HIP(hipMemsetAsync(flt.dropped, 0, sizeof(unsigned long long), s));
  }
  const int TR = kDenseTR;
  const int S16 = (int)((m + TR - 1) / TR);
  const int S = use_units ? S4 : S16;
  const int G = kSymRBW * 32 / TR;  // tiles per symbolic pass of the streaming tier

  DevBuf b_colmap, b_rows, b_bpos, b_flop, b_order, b_flop_s, b_order_s, b_P, b_nnz, b_cp, b_tmp, b_thr, b_bounds,
      b_flags, b_nsel, b_err, b_vcnt, b_ucnt, b_soff, b_sir, b_sval, b_bcnt;
  CachedBuf b_ul, b_uh, b_us;
  int32_t *colmap, *rowsA, *bpos, *order, *order_s, *P;
  int *err, *ucnt;
  unsigned long long* vcnt;
  int64_t *stage_off = nullptr, *stage_ir = nullptr;  // staged tiny tier (see the symbolic pass)
  T* stage_val = nullptr;
  bool tiny_staged = false;
  int64_t *flop, *flop_s, *nnzcol, *cpC, *thr, *bounds, *nsel;
  uint8_t* flags;
  Unit *units_l = nullptr, *units_h = nullptr;
  UnitLists ul;
  DevBuf b_lsz, b_loff, b_lists, b_lbc;
  CBX_TRY(dalloc(b_colmap, A->n, s, &colmap));
  CBX_TRY(dalloc(b_rows, nnzA, s, &rowsA));
  CBX_TRY(dalloc(b_bpos, nnzB, s, &bpos));
  CBX_TRY(dalloc(b_flop, nzcB, s, &flop));
  CBX_TRY(dalloc(b_flop_s, nzcB, s, &flop_s));
  CBX_TRY(dalloc(b_order, nzcB, s, &order));
  CBX_TRY(dalloc(b_order_s, nzcB, s, &order_s));
  CBX_TRY(dalloc(b_nnz, nzcB + 1, s, &nnzcol));
  CBX_TRY(dalloc(b_cp, nzcB + 1, s, &cpC));
  CBX_TRY(dalloc(b_thr, 8, s, &thr));
  CBX_TRY(dalloc(b_bounds, 8, s, &bounds));
  CBX_TRY(dalloc(b_flags, nzcB, s, &flags));
  CBX_TRY(dalloc(b_nsel, 2, s, &nsel));
  CBX_TRY(dalloc(b_err, 1, s, &err));
  CBX_TRY(dalloc(b_vcnt, 6, s, &vcnt));
  CBX_TRY(dalloc(b_ucnt, 3, s, &ucnt));
  CBX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  CBX_HIP(hipMemsetAsync(ucnt, 0, 3 * sizeof(int), s));

  const T* valA = (const T*)A->val;
  const T* valB = (const T*)B->val;

  // --- validate both operands before any indexed access
  if (!cfg.trusted_a) CBX_TRY(launch_validate(A, 4, vcnt, err, s));
  if (!cfg.trusted_b) CBX_TRY(launch_validate(B, 5, vcnt + 3, err, s));
  {
    int herr = 0;
    CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipStreamSynchronize(s));
    if (herr)
      return set_error(CBX_ERR_INVALID,
                       "cbx_spgemm: %s is not a valid DCSC (cp/jc/ir out of range, or rows not strictly "
                       "increasing within a column)", (herr & 16) ? "A" : "B");
  }

  // --- setup: column map, 32-bit rows, B->A positions, flops, tile pointers
  k_fill_i32<<<nblk(A->n, 256), 256, 0, s>>>(colmap, A->n, -1);
  k_colmap<<<nblk(nzcA, 256), 256, 0, s>>>(A->jc, nzcA, colmap);
  k_rows32<<<nblk(nnzA, 256), 256, 0, s>>>(A->ir, nnzA, rowsA);
  k_bpos<<<nblk(nnzB, 256), 256, 0, s>>>(B->ir, nnzB, colmap, bpos);
  k_flop<<<wave_grid(nzcB), 256, 0, s>>>(B->cp, nzcB, bpos, A->cp, flop);
  k_iota_i32<<<nblk(nzcB, 256), 256, 0, s>>>(order, nzcB);

  // sort columns by flop, descending (largest work first)
  size_t tmp_bytes = 0, tmp2 = 0;
  CBX_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp_bytes, flop, flop_s, order, order_s, (int)nzcB, 0,
                                                       64, s));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, nnzcol, cpC, (int)(nzcB + 1), s));
  tmp_bytes = std::max(tmp_bytes, tmp2);
  CBX_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp2, B->jc, flags, (int64_t*)nullptr, nsel, (int)nzcB, s));
  tmp_bytes = std::max(tmp_bytes, tmp2);
  CBX_HIP(hipcub::DeviceReduce::Sum(nullptr, tmp2, flop, bounds + 5, (int)nzcB, s));
  tmp_bytes = std::max(tmp_bytes, tmp2);
  void* tmp;
  CBX_TRY(dalloc(b_tmp, tmp_bytes, s, (uint8_t**)&tmp));
  CBX_HIP(hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, flop, flop_s, order, order_s, (int)nzcB, 0, 64,
                                                       s));
  CBX_HIP(hipcub::DeviceReduce::Sum(tmp, tmp_bytes, flop, bounds + 5, (int)nzcB, s));

  // class thresholds: units > small_max ; hash8k > mid ; hash2k > tiny ; tiny (one wave) > 0 ;
  // large units possible > unit_split_nnz
  int64_t hthr[5] = {cfg.small_flop_max, cfg.hash_mid_flop, cfg.tiny_flop, 0, cfg.unit_split_nnz};
  CBX_HIP(hipMemcpyAsync(thr, hthr, sizeof(hthr), hipMemcpyHostToDevice, s));
  k_class_bounds<<<1, 64, 0, s>>>(flop_s, nzcB, thr, 5, bounds);
  int64_t hb[6];
  CBX_HIP(hipMemcpyAsync(hb, bounds, sizeof(hb), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemsetAsync(nnzcol, 0, sizeof(int64_t) * (nzcB + 1), s));
  CBX_HIP(hipStreamSynchronize(s));
  const int64_t n_dense = hb[0], n_h8 = hb[1] - hb[0], n_h2 = hb[2] - hb[1], n_h256 = hb[3] - hb[2];
  const int64_t n_heavy_max = hb[4];  // columns whose flops (an upper bound of nnz) exceed unit_split_nnz
  const int64_t total_flop = hb[5];
  // one workgroup per column (hash tiers, symbolic units): a dispatch holds
  // fewer than 2^32 work-items
  CBX_RETURN_IF(n_dense * 1024 >= (1ll << 32) || n_h8 * 1024 >= (1ll << 32) || n_h2 * kHashNT >= (1ll << 32),
                CBX_ERR_LIMIT, "cbx_spgemm: too many columns in one tier for a single launch");
  // tile pointers only when some column is beyond the hash tier (the table is
  // nzc(A) x m/4096 entries, large for tall packed batches of small columns)
  P = nullptr;
  if (n_dense) {
    CBX_TRY(dalloc(b_P, (size_t)nzcA * (S + 1), s, &P));
    if (use_units)
      k_ptable_wave<<<wave_grid(nzcA), 256, 0, s>>>(A->cp, nzcA, rowsA, S, P);
    else
      k_ptable<<<nblk(nzcA, 256), 256, 0, s>>>(A->cp, nzcA, rowsA, S, TR, P);
  }
  tm.mark(1, s);

  // --- learned single pass (unit-tier columns), with the exact hash-tier
  //     symbolic for the small columns
  bool learned = mode == CBX_SPGEMM_LEARNED && use_units && n_dense > 0 &&
                 (cfg.predictor != nullptr || cfg.sampling_rounds > 0);
  bool learned_done = false;
  LearnedInfo linfo;
  int64_t nnzC = 0, nnz_main = 0;
  int hu[3] = {0, 0, 0}, herr = 0;  // small (bitmap), large, and small list units
  if (learned) {
    if (n_h8) k_sym_hash<8192, kHashNT><<<(unsigned)n_h8, kHashNT, 0, s>>>(order_s, hb[0], flop, B->cp, bpos, A->cp, rowsA, nnzcol, err);
    if (n_h2) k_sym_hash<2048, kHashNT><<<(unsigned)n_h2, kHashNT, 0, s>>>(order_s, hb[1], flop, B->cp, bpos, A->cp, rowsA, nnzcol, err);
    if (n_h256)
      launch_tiny<K, T, true>(cfg, order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA, valA, nnzcol, nullptr,
                              nullptr, nullptr, err, nullptr, s);
    tm.mark(2, s);
    tm.mark(3, s);
    CBX_TRY((learned_numeric<K, T>(A, B, cfg, s, rowsA, bpos, flop, P, S, nnzcol, cpC, tmp, tmp_bytes, C, err, linfo)));
    if (linfo.over_capacity) {  // predictions too small overall: redo exactly
      dev_free(C->ir);
      dev_free(C->val);
      C->ir = nullptr;
      C->val = nullptr;
      CBX_HIP(hipMemsetAsync(nnzcol, 0, sizeof(int64_t) * (nzcB + 1), s));
      CBX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    } else {
      learned_done = true;
      nnzC = linfo.nnzC;
      nnz_main = linfo.nnzC - linfo.hash_total;
      C->nnz = nnzC;
      CBX_TRY(dev_alloc((void**)&C->cp, sizeof(int64_t) * (nzcB + 1)));
      CBX_TRY(dev_alloc((void**)&C->jc, sizeof(int64_t) * std::max<int64_t>(nzcB, 1)));
      tm.mark(4, s);
    }
  }
  int64_t nnz_unfiltered = 0;
  DevBuf b_rsrc, b_rkept, b_rcol, b_rcnt;  // the fused threshold's run list
  if (!learned_done) {
  // --- symbolic
  int64_t cap_l = 0, cap_h = 0;
  // row maps (presence bitmaps, 512 B per 4096 rows) of the columns that can
  // form large units (flop above unit_split_nnz; they lead the flop order),
  // when they fit a share of the free memory: the large units then load them
  // instead of rebuilding them in a counting sweep.  From the caching
  // allocator, so repeated products reuse the block.  (cfg.row_map_share of
  // the free memory at most; 0 turns them off.)
  RowMap rm;
  rm.s = s;
  uint32_t*& rmap = rm.p;
  const int64_t rm_cols = std::min<int64_t>(n_heavy_max, n_dense);
  if (use_units && rm_cols > 0 && cfg.row_map_share > 0.0f) {
    const int64_t rm_words = rm_cols * (int64_t)S * kBRW;
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) == hipSuccess &&
        4.0 * (double)rm_words <= (double)cfg.row_map_share * (double)(freeb + cache_idle_bytes()) &&
        dev_alloc((void**)&rmap, sizeof(uint32_t) * rm_words) != CBX_OK)
      rmap = nullptr;
    (void)hipGetLastError();
  }
  if (use_units && n_dense) {
    cap_l = std::max<int64_t>(1, std::min<int64_t>(n_dense * (int64_t)S, 1ll << 30));
    cap_h = std::max<int64_t>(1, std::min<int64_t>(n_heavy_max * (int64_t)S, 1ll << 30));
    CBX_TRY(calloc_cached(b_ul, cap_l, s, &units_l));
    CBX_TRY(calloc_cached(b_uh, cap_h, s, &units_h));
    // the small units' row lists (per-column symbolic form only): one region
    // of min(flop_j, unit_split_nnz, m) 2-byte entries per column, plus the
    // per-block counts, when they fit a fifth of the free memory
    if (cfg.unit_lists && cfg.sym_band == 0 && symbolic_out == nullptr) {
      CBX_TRY(alloc_unit_lists(nullptr, flop_s, n_dense, S, cap_l, std::min<int64_t>(cfg.unit_split_nnz, m), ucnt + 2,
                               tmp, tmp_bytes, b_lsz, b_loff, b_lists, b_lbc, b_us, s, ul));
      if (ul.lists && lists_used) *lists_used = true;
    }
    CBX_TRY(launch_sym_units(cfg, order_s, n_dense, A, B, bpos, rowsA, P, S, nnzcol, units_l, ucnt, (int)cap_l, units_h,
                             ucnt + 1, (int)cap_h, rmap, rm_cols, err, b_bcnt, heavy_k<K, T>(), s, ul));
  } else if (n_dense) {
    k_sym_dense<kSymNT, kSymRBW><<<(unsigned)n_dense, kSymNT, 0, s>>>(order_s, 0, B->cp, bpos, A->cp, rowsA, P, S, G,
                                                                       TR, nnzcol);
  }
  if (n_h8) k_sym_hash<8192, kHashNT><<<(unsigned)n_h8, kHashNT, 0, s>>>(order_s, hb[0], flop, B->cp, bpos, A->cp, rowsA, nnzcol, err);
  if (n_h2) k_sym_hash<2048, kHashNT><<<(unsigned)n_h2, kHashNT, 0, s>>>(order_s, hb[1], flop, B->cp, bpos, A->cp, rowsA, nnzcol, err);
  // tiny tier: its numeric kernel runs now into a staging buffer (a slot of
  // flop_j entries per column) and records the exact counts; after the
  // output scan the entries are copied into place.  One gather and sort per
  // column instead of two (symbolic + numeric), when the slots fit a
  // quarter of the free memory.
  if (n_h256) {
    CBX_TRY(dalloc(b_soff, n_h256 + 1, s, &stage_off));
    CBX_HIP(hipMemsetAsync(stage_off, 0, sizeof(int64_t), s));
    CBX_HIP(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, flop_s + hb[2], stage_off + 1, (int)n_h256, s));
    int64_t slots = 0;
    CBX_HIP(hipMemcpyAsync(&slots, stage_off + n_h256, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipStreamSynchronize(s));
    size_t freeb = 0, totb = 0;
    // a symbolic-only call (symbolic_out) stages nothing: it counts rows
    // without reading values, so it runs for operands of any value type
    if (symbolic_out == nullptr && hipMemGetInfo(&freeb, &totb) == hipSuccess &&
        (double)slots * (sizeof(int64_t) + sizeof(T)) <= 0.25 * (double)(freeb + cache_idle_bytes()) &&
        dalloc(b_sir, std::max<int64_t>(slots, 1), s, &stage_ir) == CBX_OK &&
        dalloc(b_sval, std::max<int64_t>(slots, 1), s, &stage_val) == CBX_OK)
      tiny_staged = true;
    (void)hipGetLastError();
    if (tiny_staged)
      launch_tiny<K, T, false>(cfg, order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA, valA, nnzcol, nullptr,
                               stage_ir, stage_val, err, stage_off, s, flt);
    else
      launch_tiny<K, T, true>(cfg, order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA, valA, nnzcol, nullptr, nullptr, nullptr, err, nullptr, s);
  }
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, nnzcol, cpC, (int)(nzcB + 1), s));
  if (n_dense) {
    CBX_HIP(hipMemsetAsync(bounds + 6, 0, sizeof(int64_t), s));
    k_sum_ordered<<<nblk(n_dense, 256), 256, 0, s>>>(order_s, n_dense, nnzcol, (unsigned long long*)(bounds + 6));
    CBX_HIP(hipMemcpyAsync(&nnz_main, bounds + 6, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  }
  CBX_HIP(hipMemcpyAsync(&nnzC, cpC + nzcB, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(hu, ucnt, sizeof(hu), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  tm.mark(2, s);
  if (symbolic_out) CBX_HIP(hipMemcpyAsync(symbolic_out, nnzcol, sizeof(int64_t) * nzcB, hipMemcpyDeviceToDevice, s));
  CBX_HIP(hipStreamSynchronize(s));
  if (herr) return set_error(CBX_ERR_HIP, "cbx_spgemm: symbolic pass failed (flag %d)", herr);
  if (symbolic_out) return CBX_OK;
  nnz_unfiltered = nnzC;
  if (flt.on) {
    unsigned long long dropped = 0;
    CBX_HIP(hipMemcpy(&dropped, flt.dropped, sizeof(dropped), hipMemcpyDeviceToHost));
    nnz_unfiltered += (int64_t)dropped;
    // one run per unit and per hash- or tiny-tier column at most
    const int64_t cap = (int64_t)hu[0] + hu[1] + hu[2] + nzcB + 16;
    CBX_RETURN_IF(cap >= (1ll << 31), CBX_ERR_LIMIT, "cbx_spgemm: too many output ranges for the fused threshold");
    CBX_TRY(dalloc(b_rsrc, cap, s, &flt.run_src));
    CBX_TRY(dalloc(b_rkept, cap, s, &flt.run_kept));
    CBX_TRY(dalloc(b_rcol, cap, s, &flt.run_col));
    CBX_TRY(dalloc(b_rcnt, 1, s, &flt.run_count));
    CBX_HIP(hipMemsetAsync(flt.run_count, 0, sizeof(int), s));
    flt.run_cap = (int)cap;
  }

  // --- output allocation (cached: repeated products reuse the arrays)
  auto alloc_out = [&]() -> cbx_status {
    if (!C->cp) CBX_TRY(dev_alloc((void**)&C->cp, sizeof(int64_t) * (nzcB + 1)));
    if (!C->jc) CBX_TRY(dev_alloc((void**)&C->jc, sizeof(int64_t) * std::max<int64_t>(nzcB, 1)));
    if (!C->ir) CBX_TRY(dev_alloc((void**)&C->ir, sizeof(int64_t) * std::max<int64_t>(nnzC, 1)));
    if (!C->val) CBX_TRY(dev_alloc(&C->val, sizeof(T) * std::max<int64_t>(nnzC, 1)));
    return CBX_OK;
  };
  cbx_status ast = alloc_out();
  if (ast != CBX_OK && rmap) {  // give the row maps back and use the two-sweep numeric pass
    CBX_HIP(hipStreamSynchronize(s));
    dev_free(rmap);
    rmap = nullptr;
    release_cache();
    ast = alloc_out();
  }
  CBX_TRY(ast);
  C->nnz = nnzC;
  tm.mark(3, s);

  // --- numeric
  CBX_RETURN_IF((int64_t)hu[0] * kLightNT >= (1ll << 32) || (int64_t)hu[1] * kHeavyNT >= (1ll << 32) ||
                    (int64_t)hu[2] * kLightNT >= (1ll << 32),
                CBX_ERR_LIMIT, "cbx_spgemm: too many units for a single launch");
  DevBuf b_ul_sorted, b_uh_sorted, b_us_sorted;
  if (use_units && cfg.unit_order == 1) {
    if (hu[0] > 1) CBX_TRY(sort_units_by_block(units_l, hu[0], S, s, b_ul_sorted, &units_l));
    if (hu[1] > 1) CBX_TRY(sort_units_by_block(units_h, hu[1], S, s, b_uh_sorted, &units_h));
    if (hu[2] > 1) CBX_TRY(sort_units_by_block(ul.units, hu[2], S, s, b_us_sorted, &ul.units));
  }
  if (use_units) {
    // large units: one sweep over the row map when it was kept (their
    // products exceed what the two-sweep kernel holds in registers); small
    // units keep the two-sweep kernel, whose second sweep mostly replays
    // registers
    if (hu[1] && rmap)
      k_num_units_rm<K, T, kHeavyNT, 4, heavy_k<K, T>(), kHeavyRB><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
          units_h, rmap, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, rm_cols, cpC, C->ir, (T*)C->val, err, flt,
          cfg.unit_xcd_group);
    else if (hu[1])
      k_num_units<K, T, kHeavyNT, 4, heavy_k<K, T>(), kHeavyRB, 4><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
          units_h, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, cpC, C->ir, (T*)C->val, err, flt, cfg.unit_xcd_group);
    if (hu[0])
      k_num_units<K, T, kLightNT, 4, kLightK, kLightRB, 2><<<(unsigned)hu[0], kLightNT, 0, s>>>(
          units_l, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, cpC, C->ir, (T*)C->val, err, flt, cfg.unit_xcd_group);
    if (hu[2])
      k_num_list<K, T, kLightNT, 4, kLightK><<<(unsigned)hu[2], kLightNT, 0, s>>>(
          ul.units, ul.lists, ul.off, ul.bcnt, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, cpC, C->ir, (T*)C->val,
          err, flt, cfg.unit_xcd_group);
  } else if (n_dense) {
    k_num_dense<K, T, kDenseNT><<<(unsigned)n_dense, kDenseNT, 0, s>>>(order_s, 0, B->cp, bpos, valB, A->cp, rowsA,
                                                                     valA, P, S, cpC, C->ir, (T*)C->val, err);
  }
  tm.mark(4, s);
  }  // !learned_done
  const bool sort = mode != CBX_SPGEMM_UNSORTED;
#define CBX_LAUNCH_HASH(HM, HNT, NCOL, FIRST)                                                                     \
  if (NCOL) {                                                                                                     \
    if (sort)                                                                                                     \
      k_num_hash<K, T, HM, HNT, true><<<(unsigned)(NCOL), HNT, 0, s>>>(                                           \
          order_s, FIRST, nnzcol, B->cp, bpos, valB, A->cp, rowsA, valA, cpC, C->ir, (T*)C->val, err, flt);      \
    else                                                                                                          \
      k_num_hash<K, T, HM, HNT, false><<<(unsigned)(NCOL), HNT, 0, s>>>(                                          \
          order_s, FIRST, nnzcol, B->cp, bpos, valB, A->cp, rowsA, valA, cpC, C->ir, (T*)C->val, err, flt);      \
  }
  // the 8192-slot table takes ~100 KB of LDS, so one workgroup per CU: give it 16 waves
  CBX_LAUNCH_HASH(8192, 1024, n_h8, hb[0]);
  CBX_LAUNCH_HASH(2048, kHashNT, n_h2, hb[1]);
  if (n_h256 && tiny_staged)
    k_tiny_place<T><<<wave_grid(n_h256), 256, 0, s>>>(order_s, hb[2], n_h256, stage_off, nnzcol, cpC, stage_ir,
                                                       stage_val, C->ir, (T*)C->val, flt, err);
  else if (n_h256)
    launch_tiny<K, T, false>(cfg, order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA, valA, nnzcol, cpC, C->ir,
                             (T*)C->val, err, nullptr, s, flt);
#undef CBX_LAUNCH_HASH
  CBX_HIP(hipGetLastError());

  if (flt.on) {  // the fused threshold's runs into the final arrays
    CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    CBX_HIP(hipStreamSynchronize(s));
    if (herr) return set_error(CBX_ERR_HIP, "cbx_spgemm: numeric pass inconsistency (flag %d)", herr);
    CBX_TRY(filter_compact<T>(flt, nzcB, nnzcol, cpC, tmp, tmp_bytes, C, s));
    nnzC = C->nnz;
  }
  // --- DCSC of C: keep the non-empty columns
  k_nonempty_flags<<<nblk(nzcB, 256), 256, 0, s>>>(nnzcol, nzcB, flags);
  CBX_HIP(hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, B->jc, flags, C->jc, nsel, (int)nzcB, s));
  CBX_HIP(hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, cpC, flags, C->cp, nsel + 1, (int)nzcB, s));
  k_set_last<<<1, 1, 0, s>>>(C->cp, nsel, cpC, nzcB);
  int64_t nzcC = 0;
  CBX_HIP(hipMemcpyAsync(&nzcC, nsel, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  tm.mark(5, s);
  CBX_HIP(hipStreamSynchronize(s));
  C->nzc = nzcC;
  if (herr) {
    cbx_dcsc_free(C);
    return set_error(CBX_ERR_HIP, "cbx_spgemm: numeric pass inconsistency (flag %d)", herr);
  }
  if (stats) {
    stats->flops = total_flop;
    stats->nnz = nnzC;
    stats->nnz_before_filter = learned_done ? nnzC : nnz_unfiltered;
    stats->nzc = nzcC;
    stats->ms_setup = tm.ms(0, 1);
    stats->ms_symbolic = tm.ms(1, 2);
    stats->ms_alloc = tm.ms(2, 3);
    stats->ms_numeric = tm.ms(3, 5);
    stats->ms_main_kernel = learned_done ? linfo.kernel_ms : tm.ms(3, 4);
    stats->ms_total = tm.ms(0, 5);
    stats->units_light = learned_done ? linfo.units : hu[0] + hu[2];
    stats->units_heavy = learned_done ? linfo.heavy_units : hu[1];
    stats->nnz_main = nnz_main;
    stats->overflow_cols = linfo.overflow_units;
    stats->predicted_nnz = (int64_t)linfo.pred_sum;
    stats->actual_nnz_predicted = learned_done ? linfo.nnzC - linfo.hash_total - linfo.exact_total : 0;
    stats->learned_fallback = learned && !learned_done;
  }
  return CBX_OK;
}

// A product whose small units took row lists and then found no room for its
// output runs again without the lists (its list units cannot run without
// them, and the lists are released when the first attempt returns).
template <int K, class T>
static cbx_status spgemm_typed(const cbx_dcsc* A, const cbx_dcsc* B, cbx_spgemm_mode mode, const SpgemmConfig& cfg,
                               cbx_dcsc* C, cbx_spgemm_stats* stats, hipStream_t s,
                               int64_t* symbolic_out = nullptr) {
  bool lists_used = false;
  const cbx_status st = spgemm_typed_once<K, T>(A, B, mode, cfg, C, stats, s, symbolic_out, &lists_used);
  if (st != CBX_ERR_NOMEM || !lists_used) return st;
  (void)hipStreamSynchronize(s);
  cbx_dcsc_free(C);
  release_cache();
  SpgemmConfig plain = cfg;
  plain.unit_lists = 0;
  return spgemm_typed_once<K, T>(A, B, mode, plain, C, stats, s, symbolic_out, nullptr);
}

cbx_status spgemm_features(const cbx_dcsc* A, const cbx_dcsc* B, float* feats, hipStream_t s) {
  const int64_t nzcA = A->nzc, nzcB = B->nzc, nnzA = A->nnz, nnzB = B->nnz;
  if (nzcB == 0) return CBX_OK;
  DevBuf b_colmap, b_rows, b_bpos, b_err;
  int32_t *colmap, *rowsA, *bpos;
  int* err;
  CBX_TRY(dalloc(b_colmap, A->n, s, &colmap));
  CBX_TRY(dalloc(b_rows, nnzA, s, &rowsA));
  CBX_TRY(dalloc(b_bpos, nnzB, s, &bpos));
  CBX_TRY(dalloc(b_err, 1, s, &err));
  CBX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  DevBuf b_vcnt;
  unsigned long long* vcnt;
  CBX_TRY(dalloc(b_vcnt, 6, s, &vcnt));
  CBX_TRY(launch_validate(A, 4, vcnt, err, s));
  CBX_TRY(launch_validate(B, 5, vcnt + 3, err, s));
  int herr = 0;
  CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  if (herr) return set_error(CBX_ERR_INVALID, "cbx_spgemm_features: operand is not a valid DCSC");
  k_fill_i32<<<nblk(A->n, 256), 256, 0, s>>>(colmap, A->n, -1);
  if (nzcA) k_colmap<<<nblk(nzcA, 256), 256, 0, s>>>(A->jc, nzcA, colmap);
  if (nnzA) k_rows32<<<nblk(nnzA, 256), 256, 0, s>>>(A->ir, nnzA, rowsA);
  if (nnzB) k_bpos<<<nblk(nnzB, 256), 256, 0, s>>>(B->ir, nnzB, colmap, bpos);
  k_features<<<wave_grid(nzcB), 256, 0, s>>>(B->cp, nzcB, bpos, A->cp, rowsA, A->m, feats);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

cbx_status spgemm_estimate_sampling(const cbx_dcsc* A, const cbx_dcsc* B, int rounds, float* est, double* total,
                                    hipStream_t s) {
  const int64_t nzcA = A->nzc, nzcB = B->nzc, nnzA = A->nnz, nnzB = B->nnz;
  if (total) *total = 0.0;
  if (nzcB == 0) return CBX_OK;
  DevBuf b_colmap, b_rows, b_bpos, b_err, b_vcnt, b_samp, b_est, b_sum;
  int32_t *colmap, *rowsA, *bpos;
  int* err;
  unsigned long long* vcnt;
  float *samp, *e;
  double* dsum;
  CBX_TRY(dalloc(b_colmap, A->n, s, &colmap));
  CBX_TRY(dalloc(b_rows, nnzA, s, &rowsA));
  CBX_TRY(dalloc(b_bpos, nnzB, s, &bpos));
  CBX_TRY(dalloc(b_err, 1, s, &err));
  CBX_TRY(dalloc(b_vcnt, 6, s, &vcnt));
  CBX_TRY(dalloc(b_samp, sampling_scratch_floats(A, rounds), s, &samp));
  CBX_TRY(dalloc(b_sum, 1, s, &dsum));
  e = est;
  if (!e) CBX_TRY(dalloc(b_est, nzcB, s, &e));
  CBX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  CBX_TRY(launch_validate(A, 4, vcnt, err, s));
  CBX_TRY(launch_validate(B, 5, vcnt + 3, err, s));
  int herr = 0;
  CBX_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  if (herr) return set_error(CBX_ERR_INVALID, "cbx_spgemm_estimate_sampling: operand is not a valid DCSC");
  k_fill_i32<<<nblk(A->n, 256), 256, 0, s>>>(colmap, A->n, -1);
  if (nzcA) k_colmap<<<nblk(nzcA, 256), 256, 0, s>>>(A->jc, nzcA, colmap);
  if (nnzA) k_rows32<<<nblk(nnzA, 256), 256, 0, s>>>(A->ir, nnzA, rowsA);
  if (nnzB) k_bpos<<<nblk(nnzB, 256), 256, 0, s>>>(B->ir, nnzB, colmap, bpos);
  CBX_TRY(sampling_estimate_launch(A, B, rounds, rowsA, bpos, samp, e, s));
  if (total) {
    CBX_HIP(hipMemsetAsync(dsum, 0, sizeof(double), s));
    k_sum_f32<<<256, 256, 0, s>>>(e, nzcB, dsum);
    CBX_HIP(hipMemcpyAsync(total, dsum, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

cbx_status spgemm_flops(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* flop_host, hipStream_t s) {
  const int64_t nzcA = A->nzc, nzcB = B->nzc, nnzB = B->nnz;
  if (nzcB == 0) return CBX_OK;
  DevBuf b_colmap, b_bpos, b_flop;
  int32_t *colmap, *bpos;
  int64_t* flop;
  CBX_TRY(dalloc(b_colmap, A->n, s, &colmap));
  CBX_TRY(dalloc(b_bpos, nnzB, s, &bpos));
  CBX_TRY(dalloc(b_flop, nzcB, s, &flop));
  k_fill_i32<<<nblk(A->n, 256), 256, 0, s>>>(colmap, A->n, -1);
  if (nzcA) k_colmap<<<nblk(nzcA, 256), 256, 0, s>>>(A->jc, nzcA, colmap);
  if (nnzB) k_bpos<<<nblk(nnzB, 256), 256, 0, s>>>(B->ir, nnzB, colmap, bpos);
  k_flop<<<wave_grid(nzcB), 256, 0, s>>>(B->cp, nzcB, bpos, A->cp, flop);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipMemcpyAsync(flop_host, flop, sizeof(int64_t) * nzcB, hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

// The symbolic pass alone for operands of any value type: with symbolic_out
// set, spgemm_typed runs only structure kernels (the unit, hash and tiny
// tiers' counting forms, no staged tiny numeric), so no value is read and the
// instantiation's value type is immaterial.
cbx_status spgemm_symbolic(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* nnz, hipStream_t s) {
  cbx_dcsc C;
  SpgemmConfig cfg = current_config();
  return spgemm_typed<kPlusTimes, int64_t>(A, B, CBX_SPGEMM_TWO_PHASE, cfg, &C, nullptr, s, nnz);
}

cbx_status spgemm_dispatch(const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring sr, cbx_spgemm_mode mode,
                           const SpgemmConfig& cfg, cbx_dcsc* C, cbx_spgemm_stats* stats, hipStream_t s) {
  switch (sr) {
    case CBX_SR_PLUS_TIMES:
      switch (A->dtype) {
        case CBX_I64: return spgemm_typed<kPlusTimes, int64_t>(A, B, mode, cfg, C, stats, s);
        case CBX_F64: return spgemm_typed<kPlusTimes, double>(A, B, mode, cfg, C, stats, s);
        case CBX_F32: return spgemm_typed<kPlusTimes, float>(A, B, mode, cfg, C, stats, s);
        case CBX_BOOL: return spgemm_typed<kPlusTimes, uint8_t>(A, B, mode, cfg, C, stats, s);
      }
      break;
    case CBX_SR_MIN_PLUS:
      switch (A->dtype) {
        case CBX_I64: return spgemm_typed<kMinPlus, int64_t>(A, B, mode, cfg, C, stats, s);
        case CBX_F64: return spgemm_typed<kMinPlus, double>(A, B, mode, cfg, C, stats, s);
        case CBX_F32: return spgemm_typed<kMinPlus, float>(A, B, mode, cfg, C, stats, s);
      }
      break;
    case CBX_SR_SELECT_MAX:
      switch (A->dtype) {
        case CBX_I64: return spgemm_typed<kSelectMax, int64_t>(A, B, mode, cfg, C, stats, s);
        case CBX_F64: return spgemm_typed<kSelectMax, double>(A, B, mode, cfg, C, stats, s);
        case CBX_F32: return spgemm_typed<kSelectMax, float>(A, B, mode, cfg, C, stats, s);
      }
      break;
  }
  return set_error(CBX_ERR_UNSUPPORTED, "cbx_spgemm: semiring %d with dtype %d is not provided", (int)sr,
                   (int)A->dtype);
}

}  // namespace cbx
