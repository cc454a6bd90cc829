// Runtime entry points of the C ABI: devices, matrix transfer, argument
// checks and the cbx_spgemm front door.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "cbx_device.h"
#include "cbx_spgemm.h"

namespace cbx {

size_t dtype_size(cbx_dtype t) {
  switch (t) {
    case CBX_I64: return 8;
    case CBX_F64: return 8;
    case CBX_F32: return 4;
    case CBX_BOOL: return 1;
  }
  return 0;
}

// Keep freed stream-ordered allocations in the device pool so repeated calls
// do not go back to the driver.
cbx_status init_device_once() {
  static std::once_flag flag;
  static cbx_status result = CBX_OK;
  std::call_once(flag, [] {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
      result = set_error(CBX_ERR_HIP, "no HIP device available");
      return;
    }
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t thr = ~0ull;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
  });
  return result;
}

// Structural checks shared by all entry points that take a DCSC.
cbx_status check_dcsc(const cbx_dcsc* X, const char* name, cbx_location loc) {
  CBX_RETURN_IF(!X, CBX_ERR_INVALID, "%s is null", name);
  CBX_RETURN_IF(X->location != loc, CBX_ERR_INVALID, "%s: expected %s memory", name,
                loc == CBX_DEVICE ? "device" : "host");
  CBX_RETURN_IF(X->m < 0 || X->n < 0 || X->nnz < 0 || X->nzc < 0, CBX_ERR_INVALID, "%s: negative size", name);
  CBX_RETURN_IF(X->m >= (1ll << 31) || X->n >= (1ll << 31), CBX_ERR_LIMIT,
                "%s: dimensions must be below 2^31 (got %lld x %lld)", name, (long long)X->m, (long long)X->n);
  CBX_RETURN_IF(X->nzc > X->n || X->nzc > X->nnz, CBX_ERR_INVALID, "%s: nzc=%lld inconsistent with n/nnz", name,
                (long long)X->nzc);
  CBX_RETURN_IF(dtype_size(X->dtype) == 0, CBX_ERR_INVALID, "%s: unknown dtype %d", name, (int)X->dtype);
  CBX_RETURN_IF(X->nnz > 0 && (!X->cp || !X->jc || !X->ir || !X->val), CBX_ERR_INVALID, "%s: null array", name);
  return CBX_OK;
}

// ------------------------------------------------------ caching allocator
// Device arrays handed to callers (matrix copies, SpGEMM outputs) come from
// a small caching allocator: cbx_dcsc_free returns blocks to a per-process
// free list and the next allocation of a similar size reuses them, so
// repeated products of the same shape do not pay hipMalloc/hipFree (which
// costs seconds for outputs of hundreds of GB).
namespace {
struct Cache {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;
  std::multimap<size_t, void*> idle;
  size_t idle_bytes = 0;
};
Cache& cache() {
  static Cache* c = new Cache();  // never destroyed: outlives static teardown order
  return *c;
}
}  // namespace

cbx_status dev_alloc(void** p, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  Cache& c = cache();
  {
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.idle.lower_bound(bytes);
    if (it != c.idle.end() && it->first <= bytes + bytes / 4 + (1u << 20)) {
      *p = it->second;
      c.live[*p] = it->first;
      c.idle_bytes -= it->first;
      c.idle.erase(it);
      return CBX_OK;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release_cache();
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_error(CBX_ERR_NOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> g(c.mu);
  c.live[*p] = bytes;
  return CBX_OK;
}

void dev_free(void* p) {
  if (!p) return;
  Cache& c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  auto it = c.live.find(p);
  if (it == c.live.end()) {
    (void)hipFree(p);
    return;
  }
  c.idle.emplace(it->second, p);
  c.idle_bytes += it->second;
  c.live.erase(it);
}

size_t cache_idle_bytes() {
  Cache& c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  return c.idle_bytes;
}

void release_cache() {
  Cache& c = cache();
  {
    std::lock_guard<std::mutex> g(c.mu);
    for (auto& kv : c.idle) (void)hipFree(kv.second);
    c.idle.clear();
    c.idle_bytes = 0;
  }
  // and the stream pool's reserve (scratch freed by finished calls), so that
  // a large hipMalloc can use it
  int dev = 0;
  hipMemPool_t pool;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
    (void)hipMemPoolTrimTo(pool, 0);
}

}  // namespace cbx

using namespace cbx;

extern "C" void cbx_release_cached_memory(void) { release_cache(); }

cbx_status cbx::scratch_alloc(void** p, size_t bytes, hipStream_t s) {
  bytes = std::max<size_t>(bytes, 256);
  if (bytes >= kScratchCacheMin) return dev_alloc(p, bytes);
  const hipError_t e = hipMallocAsync(p, bytes, s);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_error(CBX_ERR_NOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  return CBX_OK;
}

void cbx::scratch_free(void* p, size_t bytes, hipStream_t s) {
  if (!p) return;
  if (std::max<size_t>(bytes, 256) >= kScratchCacheMin) {
    (void)hipStreamSynchronize(s);  // the stream's kernels may still read it
    dev_free(p);
  } else {
    (void)hipFreeAsync(p, s);
  }
}

extern "C" int32_t cbx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" cbx_status cbx_device_name(char* buf, size_t len) {
  CBX_RETURN_IF(!buf || len == 0, CBX_ERR_INVALID, "cbx_device_name: empty buffer");
  buf[0] = 0;
  int dev = 0;
  hipDeviceProp_t prop;
  CBX_HIP(hipGetDevice(&dev));
  CBX_HIP(hipGetDeviceProperties(&prop, dev));
  if (prop.name[0])
    std::snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
  else  // some runtimes report no marketing name
    std::snprintf(buf, len, "%s", prop.gcnArchName);
  return CBX_OK;
}

extern "C" cbx_status cbx_device_synchronize(void) {
  CBX_HIP(hipDeviceSynchronize());
  return CBX_OK;
}

extern "C" void cbx_dcsc_free(cbx_dcsc* A) {
  if (!A) return;
  void* arrs[4] = {A->cp, A->jc, A->ir, A->val};
  for (void* p : arrs) {
    if (!p) continue;
    if (A->location == CBX_DEVICE) dev_free(p);
    else std::free(p);
  }
  std::memset(A, 0, sizeof(*A));
}

static cbx_status copy_dcsc(const cbx_dcsc* src, cbx_dcsc* dst, cbx_location to) {
  const size_t vs = dtype_size(src->dtype);
  std::memset(dst, 0, sizeof(*dst));
  dst->m = src->m;
  dst->n = src->n;
  dst->nnz = src->nnz;
  dst->nzc = src->nzc;
  dst->dtype = src->dtype;
  dst->location = to;
  const size_t bytes[4] = {sizeof(int64_t) * (size_t)(src->nzc + 1), sizeof(int64_t) * (size_t)src->nzc,
                           sizeof(int64_t) * (size_t)src->nnz, vs * (size_t)src->nnz};
  const void* s_arr[4] = {src->cp, src->jc, src->ir, src->val};
  void** d_arr[4] = {(void**)&dst->cp, (void**)&dst->jc, (void**)&dst->ir, &dst->val};
  const hipMemcpyKind kind = to == CBX_DEVICE ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  for (int i = 0; i < 4; ++i) {
    const size_t b = bytes[i] ? bytes[i] : 8;
    if (to == CBX_DEVICE) {
      if (dev_alloc(d_arr[i], b) != CBX_OK) {
        cbx_dcsc_free(dst);
        return CBX_ERR_NOMEM;
      }
    } else {
      *d_arr[i] = std::malloc(b);
      if (!*d_arr[i]) {
        cbx_dcsc_free(dst);
        return set_error(CBX_ERR_NOMEM, "host allocation of %zu bytes failed", b);
      }
    }
    if (bytes[i] && s_arr[i]) {
      hipError_t e = hipMemcpy(*d_arr[i], s_arr[i], bytes[i], kind);
      if (e != hipSuccess) {
        cbx_dcsc_free(dst);
        return set_error(CBX_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
      }
    } else if (i == 0 && to == CBX_DEVICE) {
      (void)hipMemset(*d_arr[i], 0, b);
    } else if (i == 0) {
      std::memset(*d_arr[i], 0, b);
    }
  }
  return CBX_OK;
}

extern "C" cbx_status cbx_malloc(void** ptr, size_t bytes) {
  CBX_RETURN_IF(!ptr, CBX_ERR_INVALID, "cbx_malloc: null argument");
  cbx_status st = init_device_once();
  if (st) return st;
  return dev_alloc(ptr, bytes);
}

extern "C" void cbx_free(void* ptr) { dev_free(ptr); }

extern "C" cbx_status cbx_memcpy(void* dst, const void* src, size_t bytes, int32_t kind) {
  CBX_RETURN_IF(kind < 0 || kind > 2, CBX_ERR_INVALID, "cbx_memcpy: kind must be 0, 1 or 2");
  if (bytes == 0) return CBX_OK;
  CBX_RETURN_IF(!dst || !src, CBX_ERR_INVALID, "cbx_memcpy: null pointer");
  const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
  CBX_HIP(hipMemcpy(dst, src, bytes, k[kind]));
  return CBX_OK;
}

extern "C" cbx_status cbx_dcsc_to_device(const cbx_dcsc* host_src, cbx_dcsc* dev_dst) {
  CBX_RETURN_IF(!dev_dst, CBX_ERR_INVALID, "cbx_dcsc_to_device: dst is null");
  cbx_status st = check_dcsc(host_src, "src", CBX_HOST);
  if (st) return st;
  st = init_device_once();
  if (st) return st;
  return copy_dcsc(host_src, dev_dst, CBX_DEVICE);
}

extern "C" cbx_status cbx_dcsc_to_host(const cbx_dcsc* dev_src, cbx_dcsc* host_dst) {
  CBX_RETURN_IF(!host_dst, CBX_ERR_INVALID, "cbx_dcsc_to_host: dst is null");
  cbx_status st = check_dcsc(dev_src, "src", CBX_DEVICE);
  if (st) return st;
  return copy_dcsc(dev_src, host_dst, CBX_HOST);
}

extern "C" cbx_status cbx_spgemm(const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring semiring, cbx_spgemm_mode mode,
                                 cbx_dcsc* C, cbx_spgemm_stats* stats, void* stream) {
  cbx_status st;
  if ((st = check_dcsc(A, "A", CBX_DEVICE))) return st;
  if ((st = check_dcsc(B, "B", CBX_DEVICE))) return st;
  CBX_RETURN_IF(!C, CBX_ERR_INVALID, "cbx_spgemm: C is null");
  CBX_RETURN_IF(A->n != B->m, CBX_ERR_INVALID, "cbx_spgemm: inner dimensions differ (A is %lld x %lld, B is %lld x %lld)",
                (long long)A->m, (long long)A->n, (long long)B->m, (long long)B->n);
  CBX_RETURN_IF(A->dtype != B->dtype, CBX_ERR_UNSUPPORTED, "cbx_spgemm: A and B must share a dtype");
  CBX_RETURN_IF(A->nnz >= (1ll << 31) || B->nnz >= (1ll << 31), CBX_ERR_LIMIT,
                "cbx_spgemm: operands must hold fewer than 2^31 nonzeros each");
  CBX_RETURN_IF(mode != CBX_SPGEMM_TWO_PHASE && mode != CBX_SPGEMM_UNSORTED && mode != CBX_SPGEMM_LEARNED,
                CBX_ERR_INVALID, "cbx_spgemm: unknown mode %d", (int)mode);
  if ((st = init_device_once())) return st;
  SpgemmConfig cfg = current_config();
  CBX_RETURN_IF(mode == CBX_SPGEMM_LEARNED && !cfg.predictor && cfg.sampling_rounds == 0, CBX_ERR_INVALID,
                "cbx_spgemm: learned mode needs a predictor (cbx_spgemm_set_predictor or "
                "cbx_spgemm_set_sampling_predictor)");
  return spgemm_dispatch(A, B, semiring, mode, cfg, C, stats, (hipStream_t)stream);
}

static cbx_status check_pair(const cbx_dcsc* A, const cbx_dcsc* B, const char* who) {
  cbx_status st;
  if ((st = check_dcsc(A, "A", CBX_DEVICE))) return st;
  if ((st = check_dcsc(B, "B", CBX_DEVICE))) return st;
  CBX_RETURN_IF(A->n != B->m, CBX_ERR_INVALID, "%s: inner dimensions differ", who);
  CBX_RETURN_IF(A->nnz >= (1ll << 31) || B->nnz >= (1ll << 31), CBX_ERR_LIMIT,
                "%s: operands must hold fewer than 2^31 nonzeros each", who);
  return init_device_once();
}

extern "C" cbx_status cbx_spgemm_symbolic(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* nnz, void* stream) {
  cbx_status st = check_pair(A, B, "cbx_spgemm_symbolic");
  if (st) return st;
  CBX_RETURN_IF(!nnz, CBX_ERR_INVALID, "cbx_spgemm_symbolic: null output");
  return spgemm_symbolic(A, B, nnz, (hipStream_t)stream);
}

extern "C" cbx_status cbx_spgemm_features(const cbx_dcsc* A, const cbx_dcsc* B, float* feats, void* stream) {
  cbx_status st = check_pair(A, B, "cbx_spgemm_features");
  if (st) return st;
  CBX_RETURN_IF(!feats, CBX_ERR_INVALID, "cbx_spgemm_features: null output");
  return spgemm_features(A, B, feats, (hipStream_t)stream);
}

extern "C" cbx_status cbx_spgemm_estimate_sampling(const cbx_dcsc* A, const cbx_dcsc* B, int32_t nrounds,
                                                   float* colest, double* total, void* stream) {
  cbx_status st = check_pair(A, B, "cbx_spgemm_estimate_sampling");
  if (st) return st;
  CBX_RETURN_IF(nrounds < 2 || nrounds > 8, CBX_ERR_INVALID, "cbx_spgemm_estimate_sampling: nrounds must be in [2, 8]");
  return spgemm_estimate_sampling(A, B, nrounds, colest, total, (hipStream_t)stream);
}

static float* g_predictor = nullptr;
static std::mutex g_predictor_mu;

const float* cbx::predictor_params() {
  std::lock_guard<std::mutex> g(g_predictor_mu);
  return g_predictor;
}

extern "C" cbx_status cbx_spgemm_set_predictor(const float* params) {
  std::lock_guard<std::mutex> g(g_predictor_mu);
  if (!params) {
    if (g_predictor) (void)hipFree(g_predictor);
    g_predictor = nullptr;
    return CBX_OK;
  }
  cbx_status st = init_device_once();
  if (st) return st;
  if (!g_predictor) CBX_HIP(hipMalloc(&g_predictor, sizeof(float) * CBX_MLP_PARAMS));
  CBX_HIP(hipMemcpy(g_predictor, params, sizeof(float) * CBX_MLP_PARAMS, hipMemcpyHostToDevice));
  return CBX_OK;
}

static int g_sampling_rounds = 0;

int cbx::sampling_rounds_setting() {
  std::lock_guard<std::mutex> g(g_predictor_mu);
  return g_sampling_rounds;
}

extern "C" cbx_status cbx_spgemm_set_sampling_predictor(int32_t nrounds) {
  CBX_RETURN_IF(nrounds != 0 && (nrounds < 2 || nrounds > 8), CBX_ERR_INVALID,
                "cbx_spgemm_set_sampling_predictor: nrounds must be 0 or in [2, 8]");
  std::lock_guard<std::mutex> g(g_predictor_mu);
  g_sampling_rounds = nrounds;
  return CBX_OK;
}

// ------------------------------------------------------------- tunables
static std::mutex g_config_mu;
static cbx_spgemm_config g_config = {4096, 1024, 256, 65536, 131072, 0.75f, 0.15f, 0.25f, 4, 1, 0, 0, 1, 1, 0};

SpgemmConfig cbx::current_config() {
  SpgemmConfig c;
  {
    std::lock_guard<std::mutex> g(g_config_mu);
    c.small_flop_max = g_config.hash_max_flop;
    c.hash_mid_flop = g_config.hash_mid_flop;
    c.tiny_flop = g_config.tiny_flop;
    c.heavy_nnz = g_config.heavy_nnz;
    c.unit_split_nnz = g_config.unit_split_nnz;
    c.learned_target = g_config.learned_target;
    c.capacity_margin = g_config.capacity_margin;
    c.row_map_share = g_config.row_map_share;
    c.tiny_wg_waves = g_config.tiny_wg_waves;
    c.unit_order = g_config.unit_order;
    c.tiny_algo = g_config.tiny_algo;
    c.sym_band = g_config.sym_band;
    c.mcl_fuse = g_config.mcl_fuse != 0;
    c.unit_lists = g_config.unit_lists;
    c.unit_xcd_group = g_config.unit_xcd_group;
  }
  c.predictor = predictor_params();
  c.sampling_rounds = sampling_rounds_setting();
  return c;
}

extern "C" cbx_status cbx_spgemm_get_config(cbx_spgemm_config* cfg) {
  CBX_RETURN_IF(!cfg, CBX_ERR_INVALID, "cbx_spgemm_get_config: null argument");
  std::lock_guard<std::mutex> g(g_config_mu);
  *cfg = g_config;
  return CBX_OK;
}

extern "C" cbx_status cbx_spgemm_set_config(const cbx_spgemm_config* cfg) {
  CBX_RETURN_IF(!cfg, CBX_ERR_INVALID, "cbx_spgemm_set_config: null argument");
  CBX_RETURN_IF(cfg->tiny_flop < 0 || cfg->tiny_flop > 256, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: tiny_flop must be in [0, 256]");
  CBX_RETURN_IF(cfg->hash_mid_flop < cfg->tiny_flop || cfg->hash_mid_flop > 1024, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: hash_mid_flop must be in [tiny_flop, 1024]");
  CBX_RETURN_IF(cfg->hash_max_flop < cfg->hash_mid_flop || cfg->hash_max_flop > 4096, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: hash_max_flop must be in [hash_mid_flop, 4096]");
  CBX_RETURN_IF(cfg->heavy_nnz < 4096, CBX_ERR_INVALID, "cbx_spgemm_set_config: heavy_nnz must be >= 4096");
  CBX_RETURN_IF(cfg->unit_split_nnz < 4096, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: unit_split_nnz must be >= 4096");
  CBX_RETURN_IF(!(cfg->learned_target > 0.05f && cfg->learned_target <= 4.0f), CBX_ERR_INVALID,
                "cbx_spgemm_set_config: learned_target must be in (0.05, 4]");
  CBX_RETURN_IF(!(cfg->capacity_margin >= 0.0f && cfg->capacity_margin <= 4.0f), CBX_ERR_INVALID,
                "cbx_spgemm_set_config: capacity_margin must be in [0, 4]");
  CBX_RETURN_IF(!(cfg->row_map_share >= 0.0f && cfg->row_map_share <= 1.0f), CBX_ERR_INVALID,
                "cbx_spgemm_set_config: row_map_share must be in [0, 1]");
  CBX_RETURN_IF(cfg->tiny_wg_waves != 1 && cfg->tiny_wg_waves != 2 && cfg->tiny_wg_waves != 4, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: tiny_wg_waves must be 1, 2 or 4");
  CBX_RETURN_IF(cfg->unit_order != 0 && cfg->unit_order != 1, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: unit_order must be 0 or 1");
  CBX_RETURN_IF(cfg->tiny_algo != 0 && cfg->tiny_algo != 1, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: tiny_algo must be 0 or 1");
  CBX_RETURN_IF(cfg->sym_band != 0 && cfg->sym_band != 16 && cfg->sym_band != 64, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: sym_band must be 0, 16 or 64");
  CBX_RETURN_IF(cfg->mcl_fuse != 0 && cfg->mcl_fuse != 1, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: mcl_fuse must be 0 or 1");
  CBX_RETURN_IF(cfg->unit_lists != 0 && cfg->unit_lists != 1, CBX_ERR_INVALID,
                "cbx_spgemm_set_config: unit_lists must be 0 or 1");
  {
    const int g = cfg->unit_xcd_group;
    CBX_RETURN_IF(g != 0 && g != 8 && g != 16 && g != 32 && g != 64 && g != 128, CBX_ERR_INVALID,
                  "cbx_spgemm_set_config: unit_xcd_group must be 0, 8, 16, 32, 64 or 128");
  }
  std::lock_guard<std::mutex> g(g_config_mu);
  g_config = *cfg;
  return CBX_OK;
}
