// Internal interface of the SpGEMM driver.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "cbx_internal.h"

namespace cbx {

// Tunables of the SpGEMM driver.
struct SpgemmConfig {
  // Columns with flop_j <= small_flop_max use the LDS hash tier, the rest the
  // row-tiled dense tier.  At most 4096 (the largest hash table holds 8192
  // slots and is kept at most half full).
  int64_t small_flop_max = 4096;
  // Hash tier table sizes: flop_j in (hash_mid_flop, small_flop_max] use the
  // 8192-slot table, (tiny_flop, hash_mid_flop] the 2048-slot table (each
  // table kept at most half full); columns with flop_j <= tiny_flop (<= 256)
  // take one wave each (the tiny tier).
  int64_t hash_mid_flop = 1024;
  int64_t tiny_flop = 256;
  // Learned mode: columns predicted to exceed heavy_nnz outputs are computed
  // exactly (symbolic, then two-phase units).
  int64_t heavy_nnz = 65536;
  // Unit tier: columns whose exact output exceeds unit_split_nnz use the
  // large (16384-output) units.
  int64_t unit_split_nnz = 131072;
  // Learned mode: device parameters of the nnz predictor (null = not loaded),
  // unit size target as a fraction of the value window, and the headroom of
  // the output capacity over the predicted total.
  const float* predictor = nullptr;
  // Learned mode without the network: with sampling_rounds > 0 the column
  // sizes come from the min-hash estimator of estimateNNZ_sampling
  // (sampling_estimate_launch) instead of the predictor.
  int sampling_rounds = 0;
  float learned_target = 0.75f;
  // Operands the library produced itself (a transpose, an earlier product)
  // are valid by construction and skip the device validation pass.
  bool trusted_a = false, trusted_b = false;
  float capacity_margin = 0.15f;
  // Two-phase mode: share of the free memory the large units' row maps may take (0: none).
  float row_map_share = 0.25f;
  // Workgroup shape of the tiny tier: columns (one wave each) per workgroup.
  int tiny_wg_waves = 4;
  // Exactly sized (two-phase) units in the symbolic pass's order (0) or
  // sorted by first row block (1).
  int unit_order = 1;
  // Tiny tier: register bitonic sort (0) or per-wave LDS hash table (1).
  int tiny_algo = 0;
  // Symbolic pass of the unit tier: per column (0) or in bands of 16 / 64
  // row blocks dispatched band by band.
  int sym_band = 0;
  // Fused threshold (set by the MCL expansion, two-phase mode only): the
  // product keeps only entries with !(v < filter_thr).
  bool filter = false;
  double filter_thr = 0.0;
  // cbx_mcl_expand may fuse its hard threshold into two-phase products.
  bool mcl_fuse = true;
  // Two-phase, per-column symbolic form: the symbolic pass writes the sorted
  // rows of the small-unit columns, whose units are then cut by output count
  // over any number of row blocks (k_num_list).
  int unit_lists = 1;
  // Numeric unit launches: runs of this many consecutive (row-block sorted)
  // units per XCD (0: the hardware's round-robin order).
  int unit_xcd_group = 0;
};

// The process-wide tunables (cbx_spgemm_set_config) with the installed
// predictor: the starting configuration of every product.
SpgemmConfig current_config();

// Per-column min-hash estimates of nnz(C(:,j)) (cbx_sampling.hip): `scratch`
// holds sampling_scratch_floats(A, rounds) floats; rowsA (int32 rows of A)
// and bpos (B's rows as A stored-column positions) come from the product's
// setup; est gets nzc(B) floats.  rounds in [2, 8].
size_t sampling_scratch_floats(const cbx_dcsc* A, int rounds);
cbx_status sampling_estimate_launch(const cbx_dcsc* A, const cbx_dcsc* B, int rounds, const int32_t* rowsA,
                                    const int32_t* bpos, float* scratch, float* est, hipStream_t s);
int sampling_rounds_setting();

// Launches the predictor's inference kernel (cbx_mlp.hip).
void mlp_infer_launch(const float* params, const float* X, int64_t n, float* out, hipStream_t s);
const float* predictor_params();

// Cached device allocations for arrays handed to callers (see cbx_runtime.hip).
cbx_status dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
void release_cache();
size_t cache_idle_bytes();  // bytes held by the cache and free for reuse

// Scratch of one call on stream s.  Blocks of kScratchCacheMin bytes or more
// come from the caching allocator above (plain hipMalloc on a miss, which is
// fast); smaller ones from the stream-ordered pool.  Growing the pool by
// gigabytes has taken up to 1.5 s per allocation, and hipFreeAsync of pool
// blocks of a few megabytes 2-4 ms of host time each (7 ms per R-MAT s20
// product at a 64 MB bound), so all but the smallest scratch stays out of
// it.  scratch_free waits for s before it caches a block.
constexpr size_t kScratchCacheMin = (size_t)1 << 20;
cbx_status scratch_alloc(void** p, size_t bytes, hipStream_t s);
void scratch_free(void* p, size_t bytes, hipStream_t s);

// The scratch blocks of one call, released together (at the latest when the
// set goes out of scope).  Requests of up to 64 KB are carved (256-byte
// aligned) from 1 MB chunks of the caching allocator: counters, verdicts and
// scan storage then cost no stream-pool call each (a pool free has taken
// 10-20 us of host time, several per call).
struct ScratchSet {
  static constexpr size_t kSmall = (size_t)64 << 10, kChunk = (size_t)1 << 20;
  hipStream_t s;
  std::vector<std::pair<void*, size_t>> blocks;
  uint8_t* chunk = nullptr;
  size_t chunk_used = 0;
  explicit ScratchSet(hipStream_t st) : s(st) {}
  ScratchSet(const ScratchSet&) = delete;
  ScratchSet& operator=(const ScratchSet&) = delete;
  bool take_bytes(void** p, size_t bytes) {
    bytes = bytes < 256 ? 256 : (bytes + 255) & ~(size_t)255;
    if (bytes <= kSmall) {
      if (!chunk || chunk_used + bytes > kChunk) {
        void* c = nullptr;
        if (scratch_alloc(&c, kChunk, s) != CBX_OK) return false;
        blocks.emplace_back(c, kChunk);
        chunk = (uint8_t*)c;
        chunk_used = 0;
      }
      *p = chunk + chunk_used;
      chunk_used += bytes;
      return true;
    }
    if (scratch_alloc(p, bytes, s) != CBX_OK) return false;
    blocks.emplace_back(*p, bytes);
    return true;
  }
  template <class X>
  bool take(X** p, size_t count) {
    void* q = nullptr;
    const bool ok = take_bytes(&q, count * sizeof(X));
    *p = (X*)q;
    return ok;
  }
  void release() {
    for (auto& b : blocks) scratch_free(b.first, b.second, s);
    blocks.clear();
    chunk = nullptr;
    chunk_used = 0;
  }
  ~ScratchSet() { release(); }
};

// Shared front-door helpers (cbx_runtime.hip).
size_t dtype_size(cbx_dtype t);
cbx_status init_device_once();
cbx_status check_dcsc(const cbx_dcsc* X, const char* name, cbx_location loc);

// flop_j of every stored column of B (estimateFLOP), copied to host memory.
cbx_status spgemm_flops(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* flop_host, hipStream_t s);
cbx_status spgemm_features(const cbx_dcsc* A, const cbx_dcsc* B, float* feats, hipStream_t s);
cbx_status spgemm_symbolic(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* nnz, hipStream_t s);
cbx_status spgemm_estimate_sampling(const cbx_dcsc* A, const cbx_dcsc* B, int rounds, float* est, double* total,
                                    hipStream_t s);

// Structural check of a device DCSC (the rules of cbx_dcsc_check) with
// caller-provided counters: the launch clears cnt[0..3) and runs the check on
// s; the result call waits for s and reads them.
void dcsc_check_launch(const cbx_dcsc* X, unsigned long long* cnt, hipStream_t s);
cbx_status dcsc_check_result(const unsigned long long* cnt, hipStream_t s);
// The per-column half alone (cnt[0..3) cleared first; cnt[0] violations,
// cnt[2] descents at column starts), for a pass that reads every entry
// anyway and does the per-entry half itself (rows in range, the descents
// ir[e+1] <= ir[e] of all entries, which must equal cnt[2]).
void dcsc_check_columns_launch(const cbx_dcsc* X, unsigned long long* cnt, hipStream_t s);

cbx_status spgemm_dispatch(const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring sr, cbx_spgemm_mode mode,
                           const SpgemmConfig& cfg, cbx_dcsc* C, cbx_spgemm_stats* stats, hipStream_t s);

}  // namespace cbx
