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

struct Unit {
  int32_t col;     // B stored-column position
  int32_t blk_lo;  // first block
  int32_t blk_hi;  // one past the last block
  int32_t cnt;     // outputs in the unit
  int32_t off;     // offset of the unit's first output within the column
  int32_t slot;    // the column's row-map slot (its workgroup in the symbolic launch)
};

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
                                            const int64_t* __restrict__ stage_off, int64_t slot) {
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
      if (key[u] != kNone && next != key[u]) {
        const int rank = before + __popcll(hm[u] & ((2ull << lane) - 1ull)) - 1;
        if (rank < have) {
          irC[out + rank] = key[u];
          valC[out + rank] = SR::out(x);
        }
      }
      carry = __shfl(x, kWave - 1, kWave);
      open = __shfl((int)(key[u] != kNone), kWave - 1, kWave) != 0;
      before += __popcll(hm[u]);
    }
  }
}

template <int K, class T, bool SYM>
__global__ __launch_bounds__(256) void k_tiny(const int32_t* __restrict__ order, int64_t first, int64_t ncols,
                                              const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                              const T* __restrict__ valB, const int64_t* __restrict__ cpA,
                                              const int32_t* __restrict__ rowsA, const T* __restrict__ valA,
                                              int64_t* __restrict__ nnzcol, const int64_t* __restrict__ cpC,
                                              int64_t* __restrict__ irC, T* __restrict__ valC, int* __restrict__ err,
                                              const int64_t* __restrict__ stage_off = nullptr) {
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
    if (filled > kTinyMax && lane == 0) atomicOr(err, 4);
    // sort only as many registers as the column fills (wave-uniform)
    if (filled <= kWave) tiny_finish<K, T, SYM, 1>(key, val, lane, col, nnzcol, cpC, irC, valC, err, stage_off, i);
    else if (filled <= 2 * kWave)
      tiny_finish<K, T, SYM, 2>(key, val, lane, col, nnzcol, cpC, irC, valC, err, stage_off, i);
    else tiny_finish<K, T, SYM, 4>(key, val, lane, col, nnzcol, cpC, irC, valC, err, stage_off, i);
  }
}

// Staged tiny columns into place: one wave per column copies its nnzcol[col]
// entries from its staging slot to cpC[col].
template <class T>
__global__ void k_tiny_place(const int32_t* __restrict__ order, int64_t first, int64_t ncols,
                             const int64_t* __restrict__ stage_off, const int64_t* __restrict__ nnzcol,
                             const int64_t* __restrict__ cpC, const int64_t* __restrict__ sir,
                             const T* __restrict__ sval, int64_t* __restrict__ irC, T* __restrict__ valC) {
  const int64_t i = global_wave();
  if (i >= ncols) return;
  const int col = order[first + i];
  const int64_t src = stage_off[i], dst = cpC[col], cnt = nnzcol[col];
  for (int64_t e = lane_id(); e < cnt; e += kWave) {
    irC[dst + e] = sir[src + e];
    valC[dst + e] = sval[src + e];
  }
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
                                                 int* __restrict__ err) {
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
  for (int i = threadIdx.x; i < cnt; i += NT) {
    irC[out + i] = keys[i];
    valC[out + i] = SR::out(acc[i]);
  }
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
template <int NT, int U, int IBLK>
__global__ __launch_bounds__(NT) void k_sym_units(const int32_t* __restrict__ order, int64_t first,
                                                  const int64_t* __restrict__ cpB, const int32_t* __restrict__ bpos,
                                                  const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                  const int32_t* __restrict__ P, int S, int kmax_l, int rblk_l,
                                                  int kmax_h, int rblk_h, int64_t heavy_thr,
                                                  int64_t* __restrict__ nnzcol, Unit* __restrict__ units_l,
                                                  int* __restrict__ n_l, int cap_l, Unit* __restrict__ units_h,
                                                  int* __restrict__ n_h, int cap_h, uint32_t* __restrict__ rmap,
                                                  int64_t rm_cols, int* __restrict__ err) {
  constexpr int BPP = 256;             // blocks per cutting pass (2^20 rows)
  constexpr int BMW = IBLK * kBRW;     // bitmap words of one sub-pass
  static_assert(BPP % IBLK == 0, "sub-passes tile the cutting pass");
  __shared__ uint32_t bits[BMW];
  __shared__ int bpre[BPP];           // block counts, then their inclusive prefix
  __shared__ Unit su[BPP];
  __shared__ SegStage<uint8_t, NT> st;
  __shared__ long long wsum[NT / kWave];
  __shared__ int s_nu, s_slot;
  const int col = order[first + blockIdx.x];
  const int64_t b0 = cpB[col], b1 = cpB[col + 1];
  const int npass = (S + BPP - 1) / BPP;
  const bool whole = S <= IBLK;  // one sub-pass covers every row: no tile look-ups
  int64_t colcnt = 0;
  int unit_off = 0;  // meaningful in thread 0 only
  for (int sp = 0; sp < npass; ++sp) {
    const int t_lo = sp * BPP, t_hi = min(S, t_lo + BPP), nblk = t_hi - t_lo;
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
        if (lane_id() == 0) bpre[i_lo - t_lo + b] = c;
      }
      __syncthreads();  // the counts are read, and the next sub-pass clears the bitmap
    }
    if (wave_id() == 0) {  // inclusive prefix over blocks
      int carry = 0;
      for (int b0i = 0; b0i < nblk; b0i += kWave) {
        const int b = b0i + lane_id();
        const int v = b < nblk ? bpre[b] : 0;
        const int inc = wave_incl_scan(v) + carry;
        if (b < nblk) bpre[b] = inc;
        carry = __shfl(inc, kWave - 1, kWave);
      }
    }
    __syncthreads();
    const int pass_total = bpre[nblk - 1];
    if (threadIdx.x == 0) {
      // greedy cut into units (a single block never exceeds kBR <= kmax)
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
        // a unit whose first blocks alone fill half the value window stops
        // at the window's rows, so the numeric pass handles it in one sweep
        const int dl = kmax / kBR;
        if (lo - start + 1 > dl && bpre[start + dl - 1] - before >= kmax / 2) lo = start + dl - 1;
        const int c = bpre[lo] - before;
        if (c > 0) {
          su[nu].col = col;
          su[nu].blk_lo = t_lo + start;
          su[nu].blk_hi = t_lo + lo + 1;
          su[nu].cnt = c;
          su[nu].off = unit_off;
          su[nu].slot = (int32_t)blockIdx.x;
          unit_off += c;
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
      for (int i = threadIdx.x; i < s_nu; i += NT)
        if (s_slot + i < cap) dst[s_slot + i] = su[i];
    }
    colcnt += pass_total;
    __syncthreads();
  }
  if (threadIdx.x == 0) nnzcol[col] = colcnt;
}

// Numeric pass of the unit tier: one workgroup per unit.  Sweep 1 sets the
// unit's presence bitmap, a block scan turns it into per-word ranks, sweep 2
// accumulates every product at its row's rank.  When the column's B entries
// fit one chunk, the first CR rounds of products (row, product value) stay in
// registers between the sweeps, so sweep 2 re-reads nothing for them.
template <int K, class T, int NT, int U, int KMAX, int RBLK, int CR>
__global__ __launch_bounds__(NT) void k_num_units(const Unit* __restrict__ units, const int64_t* __restrict__ cpB,
                                                  const int32_t* __restrict__ bpos, const T* __restrict__ valB,
                                                  const int64_t* __restrict__ cpA, const int32_t* __restrict__ rowsA,
                                                  const T* __restrict__ valA, const int32_t* __restrict__ P, int S,
                                                  const int64_t* __restrict__ cpC, int64_t* __restrict__ irC,
                                                  T* __restrict__ valC, int* __restrict__ err) {
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
  const Unit un = units[blockIdx.x];
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
  if (direct) {  // values already sit at their rows: one thread per row, stores stay coalesced
    for (int r = threadIdx.x; r < ninit; r += NT) {
      if (!((bits[r >> 5] >> (r & 31)) & 1u)) continue;
      const int q = rank_of(r);
      if (q < cnt) {
        irC[out + q] = (int64_t)row0 + r;
        valC[out + q] = SR::out(vals[r]);
      }
    }
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
  // emit: values in rank order (coalesced); then the rows, staged in rank
  // order in the value array's LDS (every Acc is at least 4 bytes) so that
  // they too leave in full, coalesced lines
  for (int i = threadIdx.x; i < cnt; i += NT) valC[out + i] = SR::out(vals[i]);
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
  for (int i = threadIdx.x; i < cnt; i += NT) irC[out + i] = (int64_t)row0 + rstage[i];
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
                                                     T* __restrict__ valC, int* __restrict__ err) {
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
  const Unit un = units[blockIdx.x];
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
  if (direct) {  // values sit at their rows: one thread per row, stores stay coalesced
    for (int r = threadIdx.x; r < ninit; r += NT) {
      if (!((bits[r >> 5] >> (r & 31)) & 1u)) continue;
      const int q = rank_of(r);
      if (q < cnt) {
        irC[out + q] = (int64_t)row0 + r;
        valC[out + q] = SR::out(vals[r]);
      }
    }
    return;
  }
  // values in rank order (coalesced); then the rows, staged in rank order in
  // the value array's LDS (every Acc is at least 4 bytes)
  for (int i = threadIdx.x; i < cnt; i += NT) valC[out + i] = SR::out(vals[i]);
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
  for (int i = threadIdx.x; i < cnt; i += NT) irC[out + i] = (int64_t)row0 + rstage[i];
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
  hipStream_t s = nullptr;
  ~DevBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
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
  size_t bytes = std::max<size_t>(count * sizeof(X), 256);
  hipError_t e = hipMallocAsync(&b.p, bytes, s);
  if (e != hipSuccess) {
    b.p = nullptr;
    return set_error(CBX_ERR_NOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
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
// light units span at most 16 blocks: ~53 KB of LDS, three workgroups per CU
constexpr int kLightNT = 512, kLightK = 4096, kLightRB = 16;    // <= 4096 outputs, <= 64K rows
constexpr int kHeavyNT = 1024, kHeavyK = 16384, kHeavyRB = 16;  // <= 16384 outputs, <= 64K rows
// learned units: one workgroup shape for all (512 threads, 4096-entry value
// window, 32 row blocks), cut to target * 4096 predicted outputs
constexpr int kLearnNT = 512, kLearnK = 4096, kLearnRB = 32, kLearnCR = 2;
constexpr int kLearnFlopCap = 131072;  // expected products per learned unit
constexpr int64_t kLearnChunk = 1 << 21;  // workgroups per learned launch
// Tiny tier launch: `waves` columns (one wave each) per workgroup, grid-stride
// beyond 256K columns in flight.
inline unsigned tiny_grid(int64_t ncols, int waves) {
  return (unsigned)std::min<int64_t>((ncols + waves - 1) / waves, (int64_t)262144 / waves);
}                            // learned units: expected products

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
  CachedBuf b_ul, b_uh;
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
  CBX_TRY(dalloc(b_hu, 2, s, &hucnt));
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
  if (nheavy) {
    cap_l = std::max<int64_t>(1, std::min<int64_t>(nheavy * (int64_t)S, 1ll << 30));
    cap_h = cap_l;
    CBX_TRY(calloc_cached(b_ul, cap_l, s, &units_l));
    CBX_TRY(calloc_cached(b_uh, cap_h, s, &units_h));
    CBX_HIP(hipMemsetAsync(hucnt, 0, 2 * sizeof(int), s));
    if (cfg.row_map_share > 0.0f) {
      const int64_t rm_words = nheavy * (int64_t)S * kBRW;
      size_t freeb = 0, totb = 0;
      if (hipMemGetInfo(&freeb, &totb) == hipSuccess &&
          4.0 * (double)rm_words <= (double)cfg.row_map_share * (double)(freeb + cache_idle_bytes()) &&
          dev_alloc((void**)&rm.p, sizeof(uint32_t) * rm_words) == CBX_OK)
        rm_cols = nheavy;
      (void)hipGetLastError();
    }
    k_sym_units<kSymUnitNT, 4, kSymUnitIBLK><<<(unsigned)nheavy, kSymUnitNT, 0, s>>>(
        hcols, 0, B->cp, bpos, A->cp, rowsA, P, S, kLightK, kLightRB, kHeavyK, kHeavyRB, cfg.unit_split_nnz, nnzcol,
        units_l, hucnt, (int)cap_l, units_h, hucnt + 1, (int)cap_h, rm.p, rm_cols, err);
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
  int hu[2] = {0, 0};
  CBX_HIP(hipMemcpy(hu, hucnt, sizeof(hu), hipMemcpyDeviceToHost));
  (void)hipEventRecord(kev[2], s);  // numeric of the heavy columns, after the offsets
  CBX_RETURN_IF(hu[0] > cap_l || hu[1] > cap_h, CBX_ERR_HIP, "cbx_spgemm: unit list overflow");
  CBX_RETURN_IF((int64_t)hu[0] * kLightNT >= (1ll << 32) || (int64_t)hu[1] * kHeavyNT >= (1ll << 32), CBX_ERR_LIMIT,
                "cbx_spgemm: too many units for a single launch");
  info.heavy_units = hu[0] + hu[1];
  if (hu[1] && rm.p)
    k_num_units_rm<K, T, kHeavyNT, 4, kHeavyK, kHeavyRB><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
        units_h, rm.p, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, rm_cols, cpC, C->ir,
        (T*)C->val, err);
  else if (hu[1])
    k_num_units<K, T, kHeavyNT, 4, kHeavyK, kHeavyRB, 4><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
        units_h, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, cpC, C->ir, (T*)C->val, err);
  if (hu[0])
    k_num_units<K, T, kLightNT, 4, kLightK, kLightRB, 2><<<(unsigned)hu[0], kLightNT, 0, s>>>(
        units_l, B->cp, bpos, (const T*)B->val, A->cp, rowsA, (const T*)A->val, P, S, cpC, C->ir, (T*)C->val, err);
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

template <int K, class T>
static cbx_status spgemm_typed(const cbx_dcsc* A, const cbx_dcsc* B, cbx_spgemm_mode mode, const SpgemmConfig& cfg,
                               cbx_dcsc* C, cbx_spgemm_stats* stats, hipStream_t s,
                               int64_t* symbolic_out = nullptr) {
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
  const int TR = kDenseTR;
  const int S16 = (int)((m + TR - 1) / TR);
  const int S = use_units ? S4 : S16;
  const int G = kSymRBW * 32 / TR;  // tiles per symbolic pass of the streaming tier

  DevBuf b_colmap, b_rows, b_bpos, b_flop, b_order, b_flop_s, b_order_s, b_P, b_nnz, b_cp, b_tmp, b_thr, b_bounds,
      b_flags, b_nsel, b_err, b_vcnt, b_ucnt, b_soff, b_sir, b_sval;
  CachedBuf b_ul, b_uh;
  int32_t *colmap, *rowsA, *bpos, *order, *order_s, *P;
  int *err, *ucnt;
  unsigned long long* vcnt;
  int64_t *stage_off = nullptr, *stage_ir = nullptr;  // staged tiny tier (see the symbolic pass)
  T* stage_val = nullptr;
  bool tiny_staged = false;
  int64_t *flop, *flop_s, *nnzcol, *cpC, *thr, *bounds, *nsel;
  uint8_t* flags;
  Unit *units_l = nullptr, *units_h = nullptr;
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
  CBX_TRY(dalloc(b_ucnt, 2, s, &ucnt));
  CBX_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  CBX_HIP(hipMemsetAsync(ucnt, 0, 2 * sizeof(int), s));

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
  int hu[2] = {0, 0}, herr = 0;
  if (learned) {
    if (n_h8) k_sym_hash<8192, kHashNT><<<(unsigned)n_h8, kHashNT, 0, s>>>(order_s, hb[0], flop, B->cp, bpos, A->cp, rowsA, nnzcol, err);
    if (n_h2) k_sym_hash<2048, kHashNT><<<(unsigned)n_h2, kHashNT, 0, s>>>(order_s, hb[1], flop, B->cp, bpos, A->cp, rowsA, nnzcol, err);
    if (n_h256) k_tiny<K, T, true><<<tiny_grid(n_h256, cfg.tiny_wg_waves), kWave * cfg.tiny_wg_waves, 0, s>>>(order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA, valA, nnzcol, nullptr, nullptr, nullptr, err);
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
    k_sym_units<kSymUnitNT, 4, kSymUnitIBLK><<<(unsigned)n_dense, kSymUnitNT, 0, s>>>(
        order_s, 0, B->cp, bpos, A->cp, rowsA, P, S, kLightK, kLightRB, kHeavyK, kHeavyRB, cfg.unit_split_nnz, nnzcol,
        units_l, ucnt, (int)cap_l, units_h, ucnt + 1, (int)cap_h, rmap, rm_cols, err);
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
    if (hipMemGetInfo(&freeb, &totb) == hipSuccess &&
        (double)slots * (sizeof(int64_t) + sizeof(T)) <= 0.25 * (double)(freeb + cache_idle_bytes()) &&
        dalloc(b_sir, std::max<int64_t>(slots, 1), s, &stage_ir) == CBX_OK &&
        dalloc(b_sval, std::max<int64_t>(slots, 1), s, &stage_val) == CBX_OK)
      tiny_staged = true;
    (void)hipGetLastError();
    if (tiny_staged)
      k_tiny<K, T, false><<<tiny_grid(n_h256, cfg.tiny_wg_waves), kWave * cfg.tiny_wg_waves, 0, s>>>(order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA,
                                                           valA, nnzcol, nullptr, stage_ir, stage_val, err,
                                                           stage_off);
    else
      k_tiny<K, T, true><<<tiny_grid(n_h256, cfg.tiny_wg_waves), kWave * cfg.tiny_wg_waves, 0, s>>>(order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA,
                                                          valA, nnzcol, nullptr, nullptr, nullptr, err);
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
  CBX_RETURN_IF((int64_t)hu[0] * kLightNT >= (1ll << 32) || (int64_t)hu[1] * kHeavyNT >= (1ll << 32), CBX_ERR_LIMIT,
                "cbx_spgemm: too many units for a single launch");
  if (use_units) {
    // large units: one sweep over the row map when it was kept (their
    // products exceed what the two-sweep kernel holds in registers); small
    // units keep the two-sweep kernel, whose second sweep mostly replays
    // registers
    if (hu[1] && rmap)
      k_num_units_rm<K, T, kHeavyNT, 4, kHeavyK, kHeavyRB><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
          units_h, rmap, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, rm_cols, cpC, C->ir, (T*)C->val, err);
    else if (hu[1])
      k_num_units<K, T, kHeavyNT, 4, kHeavyK, kHeavyRB, 4><<<(unsigned)hu[1], kHeavyNT, 0, s>>>(
          units_h, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, cpC, C->ir, (T*)C->val, err);
    if (hu[0])
      k_num_units<K, T, kLightNT, 4, kLightK, kLightRB, 2><<<(unsigned)hu[0], kLightNT, 0, s>>>(
          units_l, B->cp, bpos, valB, A->cp, rowsA, valA, P, S, cpC, C->ir, (T*)C->val, err);
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
          order_s, FIRST, nnzcol, B->cp, bpos, valB, A->cp, rowsA, valA, cpC, C->ir, (T*)C->val, err);           \
    else                                                                                                          \
      k_num_hash<K, T, HM, HNT, false><<<(unsigned)(NCOL), HNT, 0, s>>>(                                          \
          order_s, FIRST, nnzcol, B->cp, bpos, valB, A->cp, rowsA, valA, cpC, C->ir, (T*)C->val, err);           \
  }
  // the 8192-slot table takes ~100 KB of LDS, so one workgroup per CU: give it 16 waves
  CBX_LAUNCH_HASH(8192, 1024, n_h8, hb[0]);
  CBX_LAUNCH_HASH(2048, kHashNT, n_h2, hb[1]);
  if (n_h256 && tiny_staged)
    k_tiny_place<T><<<wave_grid(n_h256), 256, 0, s>>>(order_s, hb[2], n_h256, stage_off, nnzcol, cpC, stage_ir,
                                                       stage_val, C->ir, (T*)C->val);
  else if (n_h256)
    k_tiny<K, T, false><<<tiny_grid(n_h256, cfg.tiny_wg_waves), kWave * cfg.tiny_wg_waves, 0, s>>>(order_s, hb[2], n_h256, B->cp, bpos, valB, A->cp, rowsA,
                                                         valA, nnzcol, cpC, C->ir, (T*)C->val, err);
#undef CBX_LAUNCH_HASH
  CBX_HIP(hipGetLastError());

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
    stats->nzc = nzcC;
    stats->ms_setup = tm.ms(0, 1);
    stats->ms_symbolic = tm.ms(1, 2);
    stats->ms_alloc = tm.ms(2, 3);
    stats->ms_numeric = tm.ms(3, 5);
    stats->ms_main_kernel = learned_done ? linfo.kernel_ms : tm.ms(3, 4);
    stats->ms_total = tm.ms(0, 5);
    stats->units_light = learned_done ? linfo.units : hu[0];
    stats->units_heavy = learned_done ? linfo.heavy_units : hu[1];
    stats->nnz_main = nnz_main;
    stats->overflow_cols = linfo.overflow_units;
    stats->predicted_nnz = (int64_t)linfo.pred_sum;
    stats->actual_nnz_predicted = learned_done ? linfo.nnzC - linfo.hash_total - linfo.exact_total : 0;
    stats->learned_fallback = learned && !learned_done;
  }
  return CBX_OK;
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
