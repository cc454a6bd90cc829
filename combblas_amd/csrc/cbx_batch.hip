// Batches of independent local products, the transpose, and the batched
// Galerkin triple product for gfx950.
//
// Many small products (one per stencil in a multigrid hierarchy, one per
// process block) are run as ONE product of block-diagonal operands: the
// batch is packed into a single DCSC with row and column offsets (one launch
// per array kind for the whole batch), every tier of the SpGEMM then runs
// once over all columns of all products, and the result is split back into
// per-product matrices with one launch per array kind.  Column results do
// not depend on which other columns share a launch, so each product is
// exactly what a separate call would return.
//
// Transpose (SpDCCols::Transpose, include/CombBLAS/SpDCCols.cpp:871-1011)
// sorts (row, column) keys with a device radix sort and rebuilds DCSC.
//
// Galerkin (ReleaseTests/GalerkinNew.cpp:94-112): T = R^T is given, S = T
// transposed, AT = A*T, SAT = S*AT, all with PlusTimes<double>.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cbx_device.h"
#include "cbx_spgemm.h"

namespace cbx {
namespace {

constexpr int kMaxBatch = 4096;

#define CBX_TRY(x)                 \
  do {                             \
    cbx_status st_ = (x);          \
    if (st_ != CBX_OK) return st_; \
  } while (0)

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// One product's arrays inside a packed batch.
struct Seg {
  const int64_t* cp;
  const int64_t* jc;
  const int64_t* ir;
  const void* val;
  int64_t row_off, col_off;  // added to rows / column ids
  int64_t nz_off, c_off;     // first entry / first stored column in the packed matrix
  int64_t m, n, nnz, nzc;    // the member's own shape and sizes (for its bounds)
};

struct OutSeg {
  int64_t* cp;
  int64_t* jc;
  int64_t* ir;
  void* val;
  int64_t row_off, col_off;
  int64_t nz_off, c_off, nnz, nzc;
};

// Index of the segment holding position x: the last s with off[s] <= x.
__device__ __forceinline__ int find_seg(const int64_t* off, int count, int64_t x) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class V>
__global__ __launch_bounds__(256) void k_pack_entries(const Seg* __restrict__ segs, int count, int64_t total,
                                                      int64_t* __restrict__ ir, V* __restrict__ val) {
  __shared__ int64_t off[kMaxBatch];
  for (int s = threadIdx.x; s < count; s += blockDim.x) off[s] = segs[s].nz_off;
  __syncthreads();
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = find_seg(off, count, e);
    const int64_t i = e - off[s];
    ir[e] = segs[s].ir[i] + segs[s].row_off;
    val[e] = ((const V*)segs[s].val)[i];
  }
}

// Packs the column arrays and checks each member against its own bounds,
// which the packed matrix's validation cannot see (a row or column id that
// strays into the next member's block is in range there): column ids in
// [0, n_i), cp starting at 0 and closing at the member's nnz, and every
// non-empty column's first and last rows in [0, m_i) — all its rows, since
// the product's validation of the packed matrix checks that rows ascend
// within columns.  bad[0] counts violations.
__global__ __launch_bounds__(256) void k_pack_cols(const Seg* __restrict__ segs, int count, int64_t total,
                                                   int64_t total_nnz, int64_t* __restrict__ cp,
                                                   int64_t* __restrict__ jc, unsigned long long* __restrict__ bad) {
  __shared__ int64_t off[kMaxBatch];
  for (int s = threadIdx.x; s < count; s += blockDim.x) off[s] = segs[s].c_off;
  __syncthreads();
  long long nbad = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int s = find_seg(off, count, q);
    const int64_t i = q - off[s];
    const Seg& g = segs[s];
    const int64_t b = g.cp[i], e = g.cp[i + 1], j = g.jc[i];
    bool err = j < 0 || j >= g.n || (i == 0 && b != 0) || (i + 1 == g.nzc && e != g.nnz);
    if (!err && 0 <= b && b < e && e <= g.nnz) {  // a non-empty column inside the member's entries
      const int64_t r0 = g.ir[b], r1 = g.ir[e - 1];
      err = r0 < 0 || r0 >= g.m || r1 < 0 || r1 >= g.m;
    }
    nbad += err;
    cp[q] = b + g.nz_off;
    jc[q] = j + g.col_off;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cp[total] = total_nnz;
  block_add_counts<256>(nbad, 0, &bad[0], &bad[1]);
}

// qb[s] = first stored column of the packed matrix with column id >= bound[s].
__global__ void k_seg_bounds(const int64_t* __restrict__ jc, int64_t nzc, const int64_t* __restrict__ bound,
                             int count, const int64_t* __restrict__ cp, int64_t* __restrict__ qb,
                             int64_t* __restrict__ nzb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > count) return;
  const int64_t b = bound[s];
  int64_t lo = 0, hi = nzc;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (jc[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  qb[s] = lo;
  nzb[s] = cp[lo];
}

template <class V>
__global__ __launch_bounds__(256) void k_unpack_entries(const OutSeg* __restrict__ segs, int count, int64_t total,
                                                        const int64_t* __restrict__ ir, const V* __restrict__ val) {
  __shared__ int64_t off[kMaxBatch];
  for (int s = threadIdx.x; s < count; s += blockDim.x) off[s] = segs[s].nz_off;
  __syncthreads();
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = find_seg(off, count, e);
    const int64_t i = e - off[s];
    segs[s].ir[i] = ir[e] - segs[s].row_off;
    ((V*)segs[s].val)[i] = val[e];
  }
}

__global__ __launch_bounds__(256) void k_unpack_cols(const OutSeg* __restrict__ segs, int count, int64_t total,
                                                     const int64_t* __restrict__ cp, const int64_t* __restrict__ jc) {
  __shared__ int64_t off[kMaxBatch];
  for (int s = threadIdx.x; s < count; s += blockDim.x) off[s] = segs[s].c_off;
  __syncthreads();
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int s = find_seg(off, count, q);
    const int64_t i = q - off[s];
    segs[s].cp[i] = cp[q] - segs[s].nz_off;
    segs[s].jc[i] = jc[q] - segs[s].col_off;
  }
  // closing pointer of every product
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < count; s += (int64_t)gridDim.x * blockDim.x)
    segs[s].cp[segs[s].nzc] = segs[s].nnz;
}

// Transpose helpers: key = row << 32 | column, payload = entry index.
__global__ void k_entry_keys(const int64_t* __restrict__ cp, const int64_t* __restrict__ jc,
                             const int64_t* __restrict__ ir, int64_t nzc, uint64_t* __restrict__ key,
                             int32_t* __restrict__ idx) {
  const int64_t q = global_wave();
  if (q >= nzc) return;
  const uint64_t c = (uint64_t)jc[q];
  for (int64_t e = cp[q] + lane_id(); e < cp[q + 1]; e += kWave) {
    key[e] = ((uint64_t)ir[e] << 32) | c;
    idx[e] = (int32_t)e;
  }
}

__global__ void k_col_heads(const uint64_t* __restrict__ key, int64_t nnz, int32_t* __restrict__ head) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < nnz) head[e] = (e == 0 || (key[e] >> 32) != (key[e - 1] >> 32)) ? 1 : 0;
  if (e == nnz) head[e] = 0;
}

template <class V>
__global__ void k_transpose_write(const uint64_t* __restrict__ key, const int32_t* __restrict__ idx,
                                  const int32_t* __restrict__ head, const int32_t* __restrict__ hpos, int64_t nnz,
                                  const V* __restrict__ val, int64_t* __restrict__ ocp, int64_t* __restrict__ ojc,
                                  int64_t* __restrict__ oir, V* __restrict__ oval) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const uint64_t k = key[e];
  oir[e] = (int64_t)(k & 0xffffffffull);
  oval[e] = val[idx[e]];
  if (head[e]) {
    ojc[hpos[e]] = (int64_t)(k >> 32);
    ocp[hpos[e]] = e;
  }
}

cbx_status alloc_dev_dcsc(cbx_dcsc* X, int64_t m, int64_t n, int64_t nnz, int64_t nzc, cbx_dtype dt) {
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

struct Owned {
  cbx_dcsc x{};
  ~Owned() { cbx_dcsc_free(&x); }
};

struct DevMem {
  void* p = nullptr;
  ~DevMem() { dev_free(p); }
  template <class X>
  cbx_status get(X** out, size_t count) {
    CBX_TRY(dev_alloc(&p, sizeof(X) * std::max<size_t>(count, 1)));
    *out = (X*)p;
    return CBX_OK;
  }
};

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>(nblocks(n, 256), 1), 65536); }

template <class V>
cbx_status launch_pack_entries(const Seg* segs, int count, int64_t total, int64_t* ir, void* val, hipStream_t s) {
  if (total) k_pack_entries<V><<<grid_for(total), 256, 0, s>>>(segs, count, total, ir, (V*)val);
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}

template <class V>
cbx_status launch_unpack_entries(const OutSeg* segs, int count, int64_t total, const int64_t* ir, const void* val,
                                 hipStream_t s) {
  if (total) k_unpack_entries<V><<<grid_for(total), 256, 0, s>>>(segs, count, total, ir, (const V*)val);
  CBX_HIP(hipGetLastError());
  return CBX_OK;
}

}  // namespace

// Packs X[0..count) into one block-diagonal DCSC (rows and columns offset by
// the running sums of m and n), checking each member against its own bounds
// (k_pack_cols); the product validates the packed matrix as any operand.
static cbx_status pack(const cbx_dcsc* X, int64_t count, cbx_dcsc* out, hipStream_t s) {
  std::vector<Seg> segs(count);
  int64_t m = 0, n = 0, nnz = 0, nzc = 0;
  for (int64_t i = 0; i < count; ++i) {
    segs[i] = Seg{X[i].cp, X[i].jc, X[i].ir, X[i].val, m, n, nnz, nzc, X[i].m, X[i].n, X[i].nnz, X[i].nzc};
    m += X[i].m;
    n += X[i].n;
    nnz += X[i].nnz;
    nzc += X[i].nzc;
  }
  CBX_TRY(alloc_dev_dcsc(out, m, n, nnz, nzc, X[0].dtype));
  DevMem dsegs, dbad;
  Seg* d;
  unsigned long long* bad;
  CBX_TRY(dsegs.get(&d, count));
  CBX_TRY(dbad.get(&bad, 2));
  CBX_HIP(hipMemcpyAsync(d, segs.data(), sizeof(Seg) * count, hipMemcpyHostToDevice, s));
  CBX_HIP(hipMemsetAsync(bad, 0, 2 * sizeof(unsigned long long), s));
  k_pack_cols<<<grid_for(nzc), 256, 0, s>>>(d, (int)count, nzc, nnz, out->cp, out->jc, bad);
  CBX_HIP(hipGetLastError());
  switch (dtype_size(X[0].dtype)) {
    case 8: CBX_TRY(launch_pack_entries<uint64_t>(d, (int)count, nnz, out->ir, out->val, s)); break;
    case 4: CBX_TRY(launch_pack_entries<uint32_t>(d, (int)count, nnz, out->ir, out->val, s)); break;
    default: CBX_TRY(launch_pack_entries<uint8_t>(d, (int)count, nnz, out->ir, out->val, s)); break;
  }
  unsigned long long hbad = 0;
  CBX_HIP(hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));  // the segment table and counter are released at return
  if (hbad) {
    cbx_dcsc_free(out);
    return set_error(CBX_ERR_INVALID, "cbx_dcsc_check: a batch member is malformed (index outside its own shape)");
  }
  return CBX_OK;
}

// Splits a packed product into count matrices: product i owns the columns
// [col_bound[i], col_bound[i+1]) and rows offset by row_off[i].
static cbx_status unpack(const cbx_dcsc* P, int64_t count, const std::vector<int64_t>& col_bound,
                         const std::vector<int64_t>& row_off, const std::vector<int64_t>& ms,
                         const std::vector<int64_t>& ns, cbx_dcsc* out, hipStream_t s) {
  DevMem dbound, dqb, dnzb;
  int64_t *bound, *qb, *nzb;
  CBX_TRY(dbound.get(&bound, count + 1));
  CBX_TRY(dqb.get(&qb, count + 1));
  CBX_TRY(dnzb.get(&nzb, count + 1));
  CBX_HIP(hipMemcpyAsync(bound, col_bound.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, s));
  k_seg_bounds<<<nblocks(count + 1, 256), 256, 0, s>>>(P->jc, P->nzc, bound, (int)count, P->cp, qb, nzb);
  CBX_HIP(hipGetLastError());
  std::vector<int64_t> hqb(count + 1), hnzb(count + 1);
  CBX_HIP(hipMemcpyAsync(hqb.data(), qb, sizeof(int64_t) * (count + 1), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipMemcpyAsync(hnzb.data(), nzb, sizeof(int64_t) * (count + 1), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  std::vector<OutSeg> segs(count);
  for (int64_t i = 0; i < count; ++i) {
    const int64_t nzc = hqb[i + 1] - hqb[i], nnz = hnzb[i + 1] - hnzb[i];
    cbx_status st = alloc_dev_dcsc(&out[i], ms[i], ns[i], nnz, nzc, P->dtype);
    if (st) {
      for (int64_t k = 0; k < i; ++k) cbx_dcsc_free(&out[k]);
      return st;
    }
    segs[i] = OutSeg{out[i].cp, out[i].jc, out[i].ir, out[i].val, row_off[i], col_bound[i],
                     hnzb[i], hqb[i], nnz, nzc};
  }
  DevMem dsegs;
  OutSeg* d;
  CBX_TRY(dsegs.get(&d, count));
  CBX_HIP(hipMemcpyAsync(d, segs.data(), sizeof(OutSeg) * count, hipMemcpyHostToDevice, s));
  k_unpack_cols<<<grid_for(std::max<int64_t>(P->nzc, count)), 256, 0, s>>>(d, (int)count, P->nzc, P->cp, P->jc);
  CBX_HIP(hipGetLastError());
  switch (dtype_size(P->dtype)) {
    case 8: CBX_TRY(launch_unpack_entries<uint64_t>(d, (int)count, P->nnz, P->ir, P->val, s)); break;
    case 4: CBX_TRY(launch_unpack_entries<uint32_t>(d, (int)count, P->nnz, P->ir, P->val, s)); break;
    default: CBX_TRY(launch_unpack_entries<uint8_t>(d, (int)count, P->nnz, P->ir, P->val, s)); break;
  }
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

template <class V>
static cbx_status transpose_typed(const cbx_dcsc* X, cbx_dcsc* out, hipStream_t s) {
  const int64_t nnz = X->nnz;
  DevMem dk0, dk1, di0, di1, dh, dhp, dtmp;
  uint64_t *k0, *k1;
  int32_t *i0, *i1, *head, *hpos;
  CBX_TRY(dk0.get(&k0, nnz));
  CBX_TRY(dk1.get(&k1, nnz));
  CBX_TRY(di0.get(&i0, nnz));
  CBX_TRY(di1.get(&i1, nnz));
  CBX_TRY(dh.get(&head, nnz + 1));
  CBX_TRY(dhp.get(&hpos, nnz + 1));
  if (X->nzc) k_entry_keys<<<wave_grid(X->nzc), 256, 0, s>>>(X->cp, X->jc, X->ir, X->nzc, k0, i0);
  CBX_HIP(hipGetLastError());
  int rbits = 1;
  while (rbits < 31 && ((int64_t)1 << rbits) < X->m) ++rbits;
  size_t tb = 0, tb2 = 0;
  hipcub::DoubleBuffer<uint64_t> keys(k0, k1);
  hipcub::DoubleBuffer<int32_t> vals(i0, i1);
  CBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, vals, (int)nnz, 0, 32 + rbits, s));
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, head, hpos, (int)(nnz + 1), s));
  uint8_t* tmp;
  CBX_TRY(dtmp.get(&tmp, std::max(tb, tb2)));
  tb = std::max(tb, tb2);
  CBX_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, vals, (int)nnz, 0, 32 + rbits, s));
  k_col_heads<<<nblocks(nnz + 1, 256), 256, 0, s>>>(keys.Current(), nnz, head);
  CBX_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, head, hpos, (int)(nnz + 1), s));
  int32_t nzc = 0;
  CBX_HIP(hipMemcpyAsync(&nzc, hpos + nnz, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CBX_HIP(hipStreamSynchronize(s));
  CBX_TRY(alloc_dev_dcsc(out, X->n, X->m, nnz, nzc, X->dtype));
  if (nnz)
    k_transpose_write<V><<<nblocks(nnz, 256), 256, 0, s>>>(keys.Current(), vals.Current(), head, hpos, nnz,
                                                         (const V*)X->val, out->cp, out->jc, out->ir, (V*)out->val);
  CBX_HIP(hipGetLastError());
  CBX_HIP(hipMemcpyAsync(out->cp + nzc, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
  CBX_HIP(hipStreamSynchronize(s));
  return CBX_OK;
}

static cbx_status transpose(const cbx_dcsc* X, cbx_dcsc* out, hipStream_t s) {
  switch (dtype_size(X->dtype)) {
    case 8: return transpose_typed<uint64_t>(X, out, s);
    case 4: return transpose_typed<uint32_t>(X, out, s);
    default: return transpose_typed<uint8_t>(X, out, s);
  }
}

static void add_stats(cbx_spgemm_stats* acc, const cbx_spgemm_stats& s) {
  acc->flops += s.flops;
  acc->nnz += s.nnz;
  acc->nzc += s.nzc;
  acc->overflow_cols += s.overflow_cols;
  acc->predicted_nnz += s.predicted_nnz;
  acc->learned_fallback += s.learned_fallback;
  acc->units_light += s.units_light;
  acc->units_heavy += s.units_heavy;
  acc->nnz_main += s.nnz_main;
  acc->ms_setup += s.ms_setup;
  acc->ms_symbolic += s.ms_symbolic;
  acc->ms_alloc += s.ms_alloc;
  acc->ms_numeric += s.ms_numeric;
  acc->ms_main_kernel += s.ms_main_kernel;
}

static cbx_status check_batch(int64_t count, const cbx_dcsc* X, const char* name) {
  CBX_RETURN_IF(count < 1 || count > kMaxBatch, CBX_ERR_LIMIT, "batch size must be in [1, %d]", kMaxBatch);
  CBX_RETURN_IF(!X, CBX_ERR_INVALID, "%s is null", name);
  int64_t nnz = 0, m = 0, n = 0;
  for (int64_t i = 0; i < count; ++i) {
    CBX_TRY(check_dcsc(&X[i], name, CBX_DEVICE));
    CBX_RETURN_IF(X[i].dtype != X[0].dtype, CBX_ERR_UNSUPPORTED, "%s: batch members must share a dtype", name);
    nnz += X[i].nnz;
    m += X[i].m;
    n += X[i].n;
  }
  CBX_RETURN_IF(nnz >= (1ll << 31) || m >= (1ll << 31) || n >= (1ll << 31), CBX_ERR_LIMIT,
                "%s: the packed batch must hold fewer than 2^31 rows, columns and nonzeros", name);
  return CBX_OK;
}

static cbx_status spgemm_batched(int64_t count, const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring sr,
                                 cbx_spgemm_mode mode, const SpgemmConfig& cfg, cbx_dcsc* C,
                                 cbx_spgemm_stats* st, hipStream_t s) {
  Owned PA, PB, PC;
  CBX_TRY(pack(A, count, &PA.x, s));
  CBX_TRY(pack(B, count, &PB.x, s));
  cbx_spgemm_stats one;
  CBX_TRY(spgemm_dispatch(&PA.x, &PB.x, sr, mode, cfg, &PC.x, &one, s));
  add_stats(st, one);
  std::vector<int64_t> bound(count + 1, 0), roff(count, 0), ms(count), ns(count);
  for (int64_t i = 0; i < count; ++i) {
    bound[i + 1] = bound[i] + B[i].n;
    roff[i] = i ? roff[i - 1] + A[i - 1].m : 0;
    ms[i] = A[i].m;
    ns[i] = B[i].n;
  }
  return unpack(&PC.x, count, bound, roff, ms, ns, C, s);
}

static cbx_status galerkin_batched(int64_t count, const cbx_dcsc* A, const cbx_dcsc* T, const SpgemmConfig& cfg,
                                   cbx_dcsc* SAT, cbx_spgemm_stats* st, hipStream_t s) {
  Owned PA, PT, PS, PAT, PSAT;
  CBX_TRY(pack(A, count, &PA.x, s));
  CBX_TRY(pack(T, count, &PT.x, s));
  CBX_TRY(transpose(&PT.x, &PS.x, s));
  cbx_spgemm_stats one;
  CBX_TRY(spgemm_dispatch(&PA.x, &PT.x, CBX_SR_PLUS_TIMES, CBX_SPGEMM_TWO_PHASE, cfg, &PAT.x, &one, s));
  add_stats(st, one);
  SpgemmConfig cfg2 = cfg;  // S (a transpose) and A*T (a product) come from this library
  cfg2.trusted_a = cfg2.trusted_b = true;
  CBX_TRY(spgemm_dispatch(&PS.x, &PAT.x, CBX_SR_PLUS_TIMES, CBX_SPGEMM_TWO_PHASE, cfg2, &PSAT.x, &one, s));
  add_stats(st, one);
  std::vector<int64_t> bound(count + 1, 0), roff(count, 0), ms(count), ns(count);
  for (int64_t i = 0; i < count; ++i) {
    bound[i + 1] = bound[i] + T[i].n;
    roff[i] = bound[i];
    ms[i] = T[i].n;
    ns[i] = T[i].n;
  }
  return unpack(&PSAT.x, count, bound, roff, ms, ns, SAT, s);
}

}  // namespace cbx

using namespace cbx;

namespace {
struct Span {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit Span(hipStream_t st) : s(st) {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
  }
  double stop() {
    float f = 0;
    (void)hipEventRecord(b, s);
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&f, a, b);
    return f;
  }
  ~Span() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};
}  // namespace

extern "C" cbx_status cbx_dcsc_transpose(const cbx_dcsc* X, cbx_dcsc* out, void* stream) {
  cbx_status st;
  if ((st = check_dcsc(X, "X", CBX_DEVICE))) return st;
  CBX_RETURN_IF(!out, CBX_ERR_INVALID, "cbx_dcsc_transpose: out is null");
  CBX_RETURN_IF(X->nnz >= (1ll << 31), CBX_ERR_LIMIT, "cbx_dcsc_transpose: at most 2^31-1 nonzeros");
  if ((st = init_device_once())) return st;
  return transpose(X, out, (hipStream_t)stream);
}

extern "C" cbx_status cbx_spgemm_batched(int64_t count, const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring semiring,
                                         cbx_spgemm_mode mode, cbx_dcsc* C, cbx_spgemm_stats* stats, void* stream) {
  cbx_status st;
  if ((st = check_batch(count, A, "A")) || (st = check_batch(count, B, "B"))) return st;
  CBX_RETURN_IF(!C, CBX_ERR_INVALID, "cbx_spgemm_batched: C is null");
  for (int64_t i = 0; i < count; ++i)
    CBX_RETURN_IF(A[i].n != B[i].m, CBX_ERR_INVALID, "cbx_spgemm_batched: product %lld: inner dimensions differ",
                  (long long)i);
  CBX_RETURN_IF(A[0].dtype != B[0].dtype, CBX_ERR_UNSUPPORTED, "cbx_spgemm_batched: A and B must share a dtype");
  CBX_RETURN_IF(mode != CBX_SPGEMM_TWO_PHASE && mode != CBX_SPGEMM_UNSORTED && mode != CBX_SPGEMM_LEARNED,
                CBX_ERR_INVALID, "cbx_spgemm_batched: unknown mode %d", (int)mode);
  if ((st = init_device_once())) return st;
  SpgemmConfig cfg = current_config();
  CBX_RETURN_IF(mode == CBX_SPGEMM_LEARNED && !cfg.predictor && cfg.sampling_rounds == 0, CBX_ERR_INVALID,
                "cbx_spgemm_batched: learned mode needs a predictor (cbx_spgemm_set_predictor)");
  cbx_spgemm_stats local;
  cbx_spgemm_stats* sp = stats ? stats : &local;
  std::memset(sp, 0, sizeof(*sp));
  for (int64_t i = 0; i < count; ++i) std::memset(&C[i], 0, sizeof(cbx_dcsc));
  Span span((hipStream_t)stream);
  st = spgemm_batched(count, A, B, semiring, mode, cfg, C, sp, (hipStream_t)stream);
  sp->ms_total = span.stop();
  return st;
}

extern "C" cbx_status cbx_galerkin_batched(int64_t count, const cbx_dcsc* A, const cbx_dcsc* T, cbx_dcsc* SAT,
                                           cbx_spgemm_stats* stats, void* stream) {
  cbx_status st;
  if ((st = check_batch(count, A, "A")) || (st = check_batch(count, T, "T"))) return st;
  CBX_RETURN_IF(!SAT, CBX_ERR_INVALID, "cbx_galerkin_batched: SAT is null");
  CBX_RETURN_IF(A[0].dtype != CBX_F64 && A[0].dtype != CBX_F32, CBX_ERR_UNSUPPORTED,
                "cbx_galerkin_batched: values must be float64 or float32");
  CBX_RETURN_IF(A[0].dtype != T[0].dtype, CBX_ERR_UNSUPPORTED, "cbx_galerkin_batched: A and T must share a dtype");
  for (int64_t i = 0; i < count; ++i)
    CBX_RETURN_IF(A[i].m != A[i].n || A[i].n != T[i].m, CBX_ERR_INVALID,
                  "cbx_galerkin_batched: product %lld: A must be square with as many rows as T", (long long)i);
  if ((st = init_device_once())) return st;
  SpgemmConfig cfg = current_config();
  cbx_spgemm_stats local;
  cbx_spgemm_stats* sp = stats ? stats : &local;
  std::memset(sp, 0, sizeof(*sp));
  for (int64_t i = 0; i < count; ++i) std::memset(&SAT[i], 0, sizeof(cbx_dcsc));
  Span span((hipStream_t)stream);
  st = galerkin_batched(count, A, T, cfg, SAT, sp, (hipStream_t)stream);
  sp->ms_total = span.stop();
  return st;
}
