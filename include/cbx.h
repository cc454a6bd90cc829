/* cbx.h -- C ABI of the MI355X (gfx950) node-local sparse kernels.
 *
 * This library replaces the node-local kernels that CombBLAS calls once per
 * SUMMA stage.  Every entry point below names the CombBLAS interface it
 * stands in for (paths relative to the CombBLAS source tree).  Matrices cross
 * the boundary as DCSC (doubly compressed sparse columns), exactly the
 * arrays of combblas::Dcsc<int64_t,NT> (include/CombBLAS/dcsc.h:125-131):
 *   cp[nzc+1]  column pointers of the non-empty columns
 *   jc[nzc]    indices of the non-empty columns (ascending)
 *   ir[nnz]    row indices, ascending and unique within each column
 *   val[nnz]   values, element type given by `dtype`
 * Indices are int64 on the boundary; internally row indices are 32-bit, so
 * every dimension must be below 2^31.
 *
 * Errors are returned as cbx_status codes (never exceptions); the text of the
 * last error on the calling thread is available from cbx_last_error().
 * All device work of one call runs on the stream passed in (NULL = the null
 * stream); calls on different streams do not share scratch memory.
 */
#ifndef CBX_H
#define CBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t cbx_status;
enum {
  CBX_OK = 0,
  CBX_ERR_INVALID = 1,     /* bad argument, shape mismatch, unsorted input */
  CBX_ERR_HIP = 2,         /* a HIP runtime call failed (no device, fault) */
  CBX_ERR_NOMEM = 3,       /* host or device allocation failed */
  CBX_ERR_UNSUPPORTED = 4, /* semiring/dtype/mode combination not provided */
  CBX_ERR_LIMIT = 5        /* a size exceeds a documented limit */
};

/* Semirings of include/CombBLAS/Semirings.h.  The value type is the dtype of
 * the operands (both operands and the result share it, NT1 = NT2 = NTO).
 *   PLUS_TIMES  PlusTimesSRing<T,T>   (Semirings.h:212-232); with CBX_BOOL it is
 *               bool + bool -> OR, bool * bool -> AND (promote_trait<bool,bool>).
 *   MIN_PLUS    MinPlusSRing<T,T>     (Semirings.h:235-255): add = min,
 *               multiply = inf_plus, saturating at numeric_limits<T>::max()
 *               (Semirings.h:40-47).
 *   SELECT_MAX  SelectMaxSRing<T,T>   (Semirings.h:165-190): add = max,
 *               multiply = a*b. */
typedef int32_t cbx_semiring;
enum { CBX_SR_PLUS_TIMES = 0, CBX_SR_MIN_PLUS = 1, CBX_SR_SELECT_MAX = 2 };

typedef int32_t cbx_dtype;
enum { CBX_I64 = 0, CBX_F64 = 1, CBX_F32 = 2, CBX_BOOL = 3 /* one byte, 0 or 1 */ };

typedef int32_t cbx_location;
enum { CBX_HOST = 0, CBX_DEVICE = 1 };

/* A DCSC matrix (see the top of this file).  `location` says whether the
 * arrays live in host memory or in device memory. */
typedef struct cbx_dcsc {
  int64_t m, n;     /* rows, columns */
  int64_t nnz, nzc; /* nonzeros, non-empty columns */
  int64_t* cp;
  int64_t* jc;
  int64_t* ir;
  void* val;
  cbx_dtype dtype;
  cbx_location location;
} cbx_dcsc;

/* ---------------------------------------------------------------- runtime */

/* Text of the last error raised on the calling thread ("" if none). */
const char* cbx_last_error(void);

/* Number of visible devices (0 when there is no GPU); never fails. */
int32_t cbx_device_count(void);

/* Block until all work on every stream of the current device has finished. */
cbx_status cbx_device_synchronize(void);

/* Free the arrays of a matrix that this library allocated (host arrays with
 * free(), device arrays with hipFree), then zero the struct.  NULL-safe. */
void cbx_dcsc_free(cbx_dcsc* A);

/* Copy a matrix between host and device.  `dst` receives newly allocated
 * arrays owned by the caller (release with cbx_dcsc_free). */
cbx_status cbx_dcsc_to_device(const cbx_dcsc* host_src, cbx_dcsc* dev_dst);
cbx_status cbx_dcsc_to_host(const cbx_dcsc* dev_src, cbx_dcsc* host_dst);

/* Structural check of a device DCSC: cp consistent, jc strictly increasing
 * and below n, rows strictly increasing within each column and below m.
 * CBX_OK or CBX_ERR_INVALID.  (The preconditions of Dcsc, dcsc.h:46-136.) */
cbx_status cbx_dcsc_check(const cbx_dcsc* X);

/* Copy only the column structure of a device DCSC to host arrays
 * (cp_host[nzc+1], jc_host[nzc]); either pointer may be NULL. */
cbx_status cbx_dcsc_structure_to_host(const cbx_dcsc* X, int64_t* cp_host, int64_t* jc_host);

/* Per stored column j of a device DCSC: sum_i mix64(i ^ salt) * bits(X(i,j))
 * in wrapping uint64 arithmetic, written to host_out[0..nzc).  Linear in X,
 * so a product's checksums follow from its operands in O(nnz) on the host. */
cbx_status cbx_dcsc_column_checksums(const cbx_dcsc* X, uint64_t salt, uint64_t* host_out);

/* ------------------------------------------------------------- generators */

enum {
  CBX_GEN_SYMMETRIZE = 1u,    /* add (c,r) for every generated (r,c) */
  CBX_GEN_NO_SELF_LOOPS = 2u, /* drop edges with r == c */
  CBX_GEN_PATTERN = 4u        /* duplicates collapse to one entry of value 1 */
};

/* R-MAT matrix of 2^scale vertices and edgefactor*2^scale generated edges,
 * quadrant probabilities (a, b, c, 1-a-b-c), no noise, no scrambling.  Each
 * edge carries an int64 value uniform in [1,9]; duplicate edges are summed
 * unless CBX_GEN_PATTERN.  Deterministic for a given seed, independent of the
 * thread count.  Output: host DCSC, dtype CBX_I64.  The stand-in for the
 * Graph500 generator path of DistEdgeList.cpp:223 (GenGraph500Data). */
cbx_status cbx_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                        uint64_t seed, uint32_t flags, cbx_dcsc* out);

/* ----------------------------------------------------------------- SpGEMM */

/* Modes of the local product. */
typedef int32_t cbx_spgemm_mode;
enum {
  /* LocalHybridSpGEMM (mtSpGEMM.h:213-460): two-phase (symbolic count, then
   * numeric), output rows sorted within every column. */
  CBX_SPGEMM_TWO_PHASE = 0,
  /* LocalSpGEMMHash(..., sort=false) (mtSpGEMM.h:463-656): same entries, rows
   * within a column in no particular order. */
  CBX_SPGEMM_UNSORTED = 1,
  /* Learned single pass: a per-column nnz predictor replaces the symbolic
   * pass; mispredicted columns are recomputed exactly.  Output identical to
   * CBX_SPGEMM_TWO_PHASE. */
  CBX_SPGEMM_LEARNED = 2
};

/* Per-call counters and phase times (milliseconds, from device events on the
 * call's stream). */
typedef struct cbx_spgemm_stats {
  int64_t flops;          /* multiply-adds, sum over columns of estimateFLOP */
  int64_t nnz;            /* nnz of the product */
  int64_t nzc;            /* non-empty output columns */
  int64_t overflow_cols;  /* learned mode: columns re-run on the exact path */
  int64_t predicted_nnz;  /* learned mode: sum of the predicted column sizes */
  int64_t units_light;    /* row-range units of the numeric pass (small/large) */
  int64_t units_heavy;
  int64_t nnz_main;       /* outputs written by the row-range-unit kernels */
  double ms_setup;        /* validation, A column map, flop count, binning */
  double ms_symbolic;     /* symbolic pass (two-phase) or prediction (learned) */
  double ms_alloc;        /* output allocation */
  double ms_numeric;      /* numeric pass including exact re-runs and DCSC assembly */
  double ms_main_kernel;  /* the row-range-unit numeric kernels alone */
  double ms_total;
} cbx_spgemm_stats;

/* Return the device memory that freed matrices left in the library's cache
 * (cbx_dcsc_free keeps device blocks for reuse by later calls). */
void cbx_release_cached_memory(void);

/* C = A*B on `semiring` (mtSpGEMM.h:213-217 LocalHybridSpGEMM<SR,NTO>(A, B,
 * clearA=false, clearB=false) and mtSpGEMM.h:463-467 LocalSpGEMMHash).
 * A (m x k) and B (k x n) are device DCSC of the same dtype.  On success `C`
 * receives a device DCSC (m x n) whose arrays the library allocated; the
 * caller releases them with cbx_dcsc_free.  Explicit zeros produced by the
 * semiring are kept, like the reference (an entry exists wherever some
 * A(i,k)*B(k,j) term exists).  `stats` may be NULL.  Synchronous with
 * respect to the host (it returns when C is complete). */
cbx_status cbx_spgemm(const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring semiring,
                      cbx_spgemm_mode mode, cbx_dcsc* C, cbx_spgemm_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CBX_H */
