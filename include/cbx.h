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

/* Name and architecture of the current device, e.g. "AMD Instinct MI355X
 * (gfx950:sramecc+:xnack-)", written NUL-terminated into buf[len]. */
cbx_status cbx_device_name(char* buf, size_t len);

/* Block until all work on every stream of the current device has finished. */
cbx_status cbx_device_synchronize(void);

/* Free the arrays of a matrix that this library allocated (host arrays with
 * free(), device arrays with hipFree), then zero the struct.  NULL-safe. */
void cbx_dcsc_free(cbx_dcsc* A);

/* Raw device buffers from the library's caching allocator (for vectors passed
 * to cbx_spmspv / cbx_bfs_run), and copies: kind 0 host->device, 1
 * device->host, 2 device->device. */
cbx_status cbx_malloc(void** ptr, size_t bytes);
void cbx_free(void* ptr);
cbx_status cbx_memcpy(void* dst, const void* src, size_t bytes, int32_t kind);

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

/* Column-stochastic power-law matrix, n x n, dtype CBX_F32 (the input of an
 * MCL expansion step, Applications/MCL.cpp:573, which reads its graph from a
 * file instead).  Degrees d_i = min(dmax, floor(xmin * u^(-1/(alpha-1)))) with
 * xmin chosen so that the mean degree is `mean` (every vertex has degree >= 1);
 * edges from the directed configuration model (column j has d_j entries, row i
 * appears d_i times over all columns, stubs paired in seeded random order);
 * log-normal(0,1) weights, duplicate pairs summed, columns normalised to sum
 * 1.  Deterministic for a given seed, independent of the thread count. */
cbx_status cbx_gen_powerlaw_stochastic(int64_t n, double mean, double alpha, int64_t dmax, uint64_t seed,
                                       cbx_dcsc* out);

/* 27-point Poisson stencil on a g^3 grid (26 on the diagonal, -1 for every
 * neighbour inside the grid), dtype CBX_F64; vertex (x,y,z) is x+g*(y+g*z).
 * The operand of ReleaseTests/GalerkinNew.cpp (read there from
 * grid3d_k5.txt). */
cbx_status cbx_gen_stencil27(int32_t g, cbx_dcsc* out);

/* 2x2x2 aggregation of a g^3 grid onto a ceil(g/2)^3 grid (odd edges give
 * smaller aggregates), entries 1.0, dtype CBX_F64: transposed = 0 gives R
 * (coarse x fine), 1 gives T = R^T (fine x coarse; restrict_T_grid3d_k5.txt
 * in GalerkinNew.cpp). */
cbx_status cbx_gen_aggregation(int32_t g, int32_t transposed, cbx_dcsc* out);

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
  /* Learned single pass: the installed nnz predictor (cbx_spgemm_set_predictor)
   * replaces the symbolic pass of the large columns; each row-range unit's
   * exact size comes from its own presence bitmap and its output offset from
   * a decoupled look-back, units larger than predicted are re-run exactly,
   * and an exceeded output capacity falls back to the two-phase path.
   * Output identical to CBX_SPGEMM_TWO_PHASE. */
  CBX_SPGEMM_LEARNED = 2
};

/* Process-wide tunables of the local product (the bin thresholds and table
 * sizes the reference hard-codes: the heap/hash switch cr < 2 at
 * mtSpGEMM.h:311 and minHashTableSize / hashScale at :366-367).  Picked per
 * workload by combblas_amd/autotune.py; defaults {4096, 1024, 256, 65536,
 * 131072, 0.75, 0.15, 0.25, 4, 1, 0, 0, 1, 1, 0}. */
typedef struct cbx_spgemm_config {
  int64_t hash_max_flop;   /* columns with flop_j <= this use the LDS hash tier (<= 4096) */
  int64_t hash_mid_flop;   /* ... with flop_j > this use the 8192-slot table, else the 2048-slot one (<= 1024) */
  int64_t tiny_flop;       /* columns with flop_j <= this take one wave each, sorted in registers (<= 256) */
  int64_t heavy_nnz;       /* learned mode: columns predicted to hold more outputs are computed exactly (symbolic
                              pass, then two-phase units) after the single pass (>= 4096) */
  int64_t unit_split_nnz;  /* unit tier: columns with more (exact) outputs use the 16384-output units, the
                              others the 4096-output ones (>= 4096) */
  float learned_target;    /* learned mode: predicted outputs per unit as a fraction of the 4096-output value
                              window; above 1 units are cut larger than their window and re-swept in windows
                              (exact, slower) */
  float capacity_margin;   /* learned mode: output capacity headroom over the predicted total */
  float row_map_share;     /* two-phase mode: at most this share of the free device memory holds the symbolic
                              pass's row maps of the columns that can form 16384-output units, which those
                              units then read instead of rebuilding them (0 disables; in [0, 1]) */
  int32_t tiny_wg_waves;   /* workgroup shape of the tiny tier: columns (one wave each) per workgroup, 1, 2 or 4 */
  int32_t unit_order;      /* order of the exactly sized row-range units in the numeric pass: 0 as the symbolic
                              pass cut them (column by column), 1 by first row block (the units in flight then
                              gather from a narrow band of A's rows, which the L2 caches hold) */
  int32_t tiny_algo;       /* tiny tier: 0 sorts a column's products in registers (bitonic network across lanes),
                              1 merges them in a per-wave LDS hash table and ranks the distinct rows */
  int32_t sym_band;        /* symbolic pass of the row-range units: 0 one workgroup per column over all its rows;
                              16 or 64 one workgroup per (column, band of that many 4096-row blocks), dispatched
                              band by band so the gathers in flight stay in one band of A's rows */
  int32_t mcl_fuse;        /* cbx_mcl_expand with two-phase sizing and recover_num 0: 1 fuses the hard threshold
                              into the product's emission (only entries that can survive the epilogue are
                              written; same result), 0 writes the whole product first */
  int32_t unit_lists;      /* two-phase mode with sym_band 0: 1 lets the symbolic pass write the sorted rows of the
                              columns that form 4096-output units (2 bytes per output, when they fit a fifth of
                              the free memory), so those columns are cut by output count alone, over any number
                              of row blocks, and each such unit gathers its products once; 0 cuts them at 16
                              blocks and counts each unit's rows in a first sweep */
  int32_t unit_xcd_group;  /* numeric pass of the row-range units: 0 leaves the hardware's round-robin deal of
                              workgroups to the 8 XCDs; 8, 16, 32, 64 or 128 give each XCD runs of that many
                              consecutive units (sorted by row block with unit_order 1), so an XCD's L2 caches
                              one band of A's rows at a time */
} cbx_spgemm_config;

cbx_status cbx_spgemm_get_config(cbx_spgemm_config* cfg);
/* Validates and installs `cfg` for every later product in the process. */
cbx_status cbx_spgemm_set_config(const cbx_spgemm_config* cfg);

/* Per-call counters and phase times (milliseconds, from device events on the
 * call's stream). */
typedef struct cbx_spgemm_stats {
  int64_t flops;          /* multiply-adds, sum over columns of estimateFLOP */
  int64_t nnz;            /* nnz of the product */
  int64_t nzc;            /* non-empty output columns */
  int64_t overflow_cols;  /* learned mode: units whose exact size exceeded the value window and were
                             re-run exactly in windows of ranks (caught by a device counter) */
  int64_t predicted_nnz;  /* learned mode: sum of the predicted sizes of the columns the single pass
                             sized by prediction */
  int64_t actual_nnz_predicted; /* learned mode: their exact sizes (the product's outputs in those columns) */
  int64_t learned_fallback; /* learned mode: 1 if the predicted output capacity was exceeded and the
                               product was recomputed by the exact two-phase path */
  int64_t units_light;    /* row-range units of the numeric pass (small/large) */
  int64_t units_heavy;
  int64_t nnz_main;       /* outputs written by the row-range-unit kernels */
  double ms_setup;        /* validation, A column map, flop count, binning */
  double ms_symbolic;     /* symbolic pass (two-phase) or prediction (learned) */
  double ms_alloc;        /* output allocation */
  double ms_numeric;      /* numeric pass including exact re-runs and DCSC assembly */
  double ms_main_kernel;  /* the row-range-unit numeric kernels alone */
  double ms_total;
  int64_t nnz_before_filter; /* entries before a fused threshold (cbx_mcl_expand's products); = nnz otherwise */
} cbx_spgemm_stats;

/* Installs the nnz predictor used by CBX_SPGEMM_LEARNED (host array of
 * CBX_MLP_PARAMS floats, copied to the device), or removes it (NULL). */
cbx_status cbx_spgemm_set_predictor(const float* params);

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

/* -------------------------------------------------- learned nnz predictor */

/* Symbolic column counts: nnz of C(:,j) for every stored column j of B
 * (estimateNNZ_Hash, mtSpGEMM.h:806-933), written to device int64[nzc_B].
 * These are the predictor's training labels. */
cbx_status cbx_spgemm_symbolic(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* nnz, void* stream);

/* Cheap per-column features of C = A*B for every stored column j of B,
 * device float[nzc_B][6]:
 *   log2(1+nnz(B(:,j))), log2(1+flop_j), log2(1+max_k nnz(A(:,k))),
 *   log2(1+flop_j/nnz(B(:,j))), row span of the contributing A columns / m,
 *   log2(1 + span*(1-exp(-flop_j/span)))   (k ranges over B(:,j)). */
cbx_status cbx_spgemm_features(const cbx_dcsc* A, const cbx_dcsc* B, float* feats, void* stream);

/* estimateNNZ_sampling (mtSpGEMM.h:936-1054) on the device, kept per column:
 * colest[j] (device float[nzc_B], may be NULL) = the min-hash estimate of
 * nnz(C(:,j)) from nrounds exponential samples per row of A (the reference's
 * sample stream: minstd_rand0 from seed 1, -log(1-u)), and *total (host, may
 * be NULL) their sum, the reference's return value before its cast to int64.
 * Matches the reference up to the last bits of -log.  nrounds in [2, 8]
 * (the reference's default is 5). */
cbx_status cbx_spgemm_estimate_sampling(const cbx_dcsc* A, const cbx_dcsc* B, int32_t nrounds, float* colest,
                                        double* total, void* stream);

/* Learned mode without the network: with nrounds in [2, 8] the single pass
 * sizes its units from the per-column estimates of
 * cbx_spgemm_estimate_sampling instead of the installed MLP (the method
 * without learning, for comparison); 0 returns to the MLP. */
cbx_status cbx_spgemm_set_sampling_predictor(int32_t nrounds);

/* The predictor: a 6 -> 64 -> 64 -> 1 ReLU MLP in FP32 predicting
 * log2(1 + nnz_j) from standardised features.  Parameters (device float
 * array of CBX_MLP_PARAMS): W1[64][6] b1[64] W2[64][64] b2[64] W3[64] b3
 * mu[6] sd[6] (feature mean and standard deviation used for scaling). */
#define CBX_MLP_FEATURES 6
#define CBX_MLP_HIDDEN 64
#define CBX_MLP_PARAMS (64 * 6 + 64 + 64 * 64 + 64 + 64 + 1 + 12)

typedef struct cbx_mlp_train_opts {
  int32_t steps;  /* Adam steps */
  int32_t batch;  /* samples per step, drawn with replacement */
  float lr;       /* Adam learning rate */
  int32_t pad;
  uint64_t seed;
} cbx_mlp_train_opts;

/* out[i] = predicted log2(1 + nnz) for features X[i][6] (device arrays). */
cbx_status cbx_mlp_forward(const float* params, const float* X, int64_t n, float* out, void* stream);
/* Mean-squared-error gradient over all n samples (device arrays; grad gets
 * CBX_MLP_PARAMS floats, the last 12 zero) and the loss (host). */
cbx_status cbx_mlp_gradient(const float* params, const float* X, const float* y, int32_t n, float* grad, float* loss);
/* Adam training on the device with the library's own kernels; loss_curve
 * (host, optional) receives each step's mini-batch loss. */
cbx_status cbx_mlp_train(float* params, const float* X, const float* y, int64_t n, const cbx_mlp_train_opts* opts,
                         float* loss_curve);

/* ----------------------------------------------------------------- SpMSpV */

/* y = A (x) x on `semiring` with a sparse vector x (SpImpl::SpMXSpV,
 * SpImpl.cpp:55-158 and SpImpl.cpp:167-213, via generic_gespmv,
 * Friends.h:448-482).  A: device DCSC (m x n), dtype CBX_BOOL or `xdtype`.
 * x: nx device indices (ascending, unique, below n) with device values of
 * `xdtype` (CBX_I64, CBX_F64 or CBX_F32).  y: device arrays of capacity m;
 * *ny (host) receives the number of entries, indices ascending.  An entry
 * exists wherever some A(i,k) x_k term exists.  A boolean A contributes x_k
 * itself (the reference's boolean specialisation never calls multiply), so
 * SELECT_MAX on a boolean A is SelectMaxSRing<bool,T> (Semirings.h:193-210). */
cbx_status cbx_spmspv(const cbx_dcsc* A, const int32_t* xind, const void* xval, int32_t nx, cbx_semiring semiring,
                      cbx_dtype xdtype, int32_t* yind, void* yval, int32_t* ny, void* stream);

/* cbx_spmspv without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as cbx_spmspv; the structure check of A and
 * every scratch array (column map, dense accumulator, bitmap, scan storage)
 * live in one static device array of CBX_SPMSPV_NOALLOC_SCRATCH_BYTES per GPU,
 * kept until the process ends.  It needs about 4*n + sizeof(value)*m + m/4
 * bytes plus a few KiB (12.4 MiB for m = n = 2^20 with 8-byte values); a
 * larger A is refused with CBX_ERR_LIMIT before any device work.  All calls
 * on one GPU share that array: calls that overlap in time on one GPU (other
 * host threads, other streams) race and must be serialised by the caller.
 * Synchronous with respect to the host, like cbx_spmspv. */
#define CBX_SPMSPV_NOALLOC_SCRATCH_BYTES (16u << 20)
cbx_status cbx_spmspv_noalloc(const cbx_dcsc* A, const int32_t* xind, const void* xval, int32_t nx,
                              cbx_semiring semiring, cbx_dtype xdtype, int32_t* yind, void* yval, int32_t* ny,
                              void* stream);

/* -------------------------------------------------------------------- BFS */

/* y = A*x with a dense x: dcsc_gespmv<SR>(A, x, y) (Friends.h:63-77), y[i] =
 * add(y[i], multiply(A(i,j), x[j])) for every stored entry, accumulated into
 * y in place (initialise it, e.g. to the semiring's identity).  x: device
 * array of A->n values, y: device array of A->m values, both of xdtype; A
 * shares that dtype or is boolean (then x[j] itself is added).  Integer
 * semirings are exact; floating-point sums are formed in an unspecified
 * order. */
cbx_status cbx_spmv_dense(const cbx_dcsc* A, const void* x, void* y, cbx_semiring semiring, cbx_dtype xdtype,
                          void* stream);

/* Top-down BFS of TopDownBFS.cpp:442-448 on one GPU:
 *   fringe.setNumToInd(); fringe = SpMV<SelectMaxSRing<bool,int64>>(A, fringe);
 *   fringe = EWiseMult(fringe, parents, true, -1); parents.Set(fringe);
 * until the fringe is empty.  parents[v] is the largest-index frontier
 * neighbour through which v was first reached; the root's own entry is set
 * when a later level reaches it back (as in the reference), vertices never
 * reached stay -1.  A plan holds the adjacency matrix in the layout the
 * kernels read plus one captured hipGraph per frontier parity; it is built
 * once and run for many roots, like the reference's prepared matrix. */
typedef struct cbx_bfs_plan_s* cbx_bfs_plan;

typedef struct cbx_bfs_stats {
  int32_t levels;   /* SpMV iterations (TopDownBFS's "Number iterations") */
  int32_t bottom_up_levels; /* of which run bottom-up (see cbx_bfs_plan_set_direction) */
  int64_t reached;  /* vertices that received a parent */
  double ms;        /* device time of the whole BFS */
} cbx_bfs_stats;

/* A: device DCSC, square (values ignored); normally the symmetric pattern of
 * an undirected graph.  The plan compares A's pattern with its transpose's: a
 * directed (asymmetric) pattern runs every level top-down, and
 * cbx_bfs_plan_set_direction refuses CBX_BFS_BOTTOM_UP for it. */
cbx_status cbx_bfs_plan_create(const cbx_dcsc* A, cbx_bfs_plan* plan);

/* How each level is computed.  Top-down is TopDownBFS's SpMV: every frontier
 * vertex offers itself to its unvisited neighbours (atomic max).  Bottom-up
 * has every unvisited vertex scan its neighbours from the largest index down
 * and stop at the first frontier vertex: the same largest frontier neighbour,
 * so parents and levels are identical in all three settings.  The default,
 * CBX_BFS_DIRECTION_OPTIMIZING, chooses per level: bottom-up while the
 * frontier's edges exceed 1/256 of the unexplored edges.  No counterpart in
 * the reference, which runs top-down only (TopDownBFS.cpp:442-448). */
typedef enum cbx_bfs_direction {
  CBX_BFS_DIRECTION_OPTIMIZING = 0,
  CBX_BFS_TOP_DOWN = 1,
  CBX_BFS_BOTTOM_UP = 2
} cbx_bfs_direction;
cbx_status cbx_bfs_plan_set_direction(cbx_bfs_plan plan, cbx_bfs_direction direction);
/* parents: device int64[n], overwritten. */
cbx_status cbx_bfs_run(cbx_bfs_plan plan, int64_t root, int64_t* parents, cbx_bfs_stats* stats);
void cbx_bfs_plan_destroy(cbx_bfs_plan plan);

/* ------------------------------------------------------------ MCL expansion */

typedef struct cbx_mcl_stats {
  int64_t flops;           /* multiply-adds of the product */
  int64_t nnz_product;     /* entries of the product before the epilogue */
  int64_t nnz;             /* entries kept */
  int64_t phases;          /* column phases the product ran in */
  int64_t recovered_cols;  /* columns whose cut came from recovery */
  int64_t selected_cols;   /* columns whose cut came from selection */
  int64_t reselected_cols; /* selected columns then recovered (recovery after selection) */
  int64_t overflow_cols;   /* learned mode, summed over phases (see cbx_spgemm_stats) */
  int64_t learned_fallback;
  double ms_product, ms_prune, ms_total;
} cbx_mcl_stats;

/* MCLPruneRecoverySelect(A, hardThreshold, selectNum, recoverNum, recoverPct,
 * kselectVersion) (ParFriends.h:185-345) on one process's block, for a
 * device DCSC of float32 or float64 values.  Per stored column: entries >
 * hard_threshold are counted and summed; a column with fewer than recover_num
 * of them, more entries in all, and a sum below recover_pct is cut at its
 * recover_num-th largest entry (Kselect, SpParMat.cpp:1413-1690: the
 * smallest entry if it has fewer); otherwise a column with more than
 * select_num of them is cut at its select_num-th largest entry, and then at
 * its recover_num-th largest if fewer than recover_num entries survive that
 * cut and their sum is below recover_pct; other columns are cut at
 * hard_threshold.  Entries with !(v < cut) are kept (PruneColumn with
 * std::less: ties at the cut are kept, and so are entries equal to the
 * threshold); emptied columns are dropped.  Column statistics are summed in
 * FP64 (the reference sums sequentially in the value type).  `out` receives
 * a new device DCSC (release with cbx_dcsc_free); `C` is not modified. */
cbx_status cbx_mcl_prune(const cbx_dcsc* C, double hard_threshold, int64_t select_num, int64_t recover_num,
                         double recover_pct, cbx_dcsc* out, cbx_mcl_stats* stats, void* stream);

/* One MCL expansion A*B followed by cbx_mcl_prune, with B's stored columns
 * split into `phases` flop-balanced column phases (MemEfficientSpGEMM,
 * ParFriends.h:452-780: ColSplit, per-phase product and prune,
 * ColConcatenate) so the unpruned product never has to fit at once;
 * phases <= 0 picks the count from free device memory.  mode:
 * CBX_SPGEMM_TWO_PHASE or CBX_SPGEMM_LEARNED (plus-times semiring). */
cbx_status cbx_mcl_expand(const cbx_dcsc* A, const cbx_dcsc* B, double hard_threshold, int64_t select_num,
                          int64_t recover_num, double recover_pct, int64_t phases, cbx_spgemm_mode mode,
                          cbx_dcsc* out, cbx_mcl_stats* stats, void* stream);

/* ----------------------------------------------- transpose, batches, Galerkin */

/* SpDCCols::Transpose (SpDCCols.cpp:871-1011): out = X^T as a new device DCSC
 * (rows sorted within columns). */
cbx_status cbx_dcsc_transpose(const cbx_dcsc* X, cbx_dcsc* out, void* stream);

/* Stage merge: MultiwayMerge<SR>(parts, m, n) (MultiwayMerge.h:428-541, and
 * MultiwayMergeHash with sorted=true, :553) followed by SpDCCols(SpTuples)
 * (SpDCCols.cpp:109-184).  out = the sum under the semiring's add of count
 * device DCSC parts of one shape and dtype (an array of count structs), as a
 * new device DCSC; entries with equal (row, col) are combined, none is
 * dropped.  Integer, min-plus and select-max results are exact; floating-
 * point plus-times sums of duplicates are formed in the product's
 * accumulation order (the reference's heap order may round differently).
 * The parts are validated as cbx_dcsc_check does, each against its own
 * m x n (CBX_ERR_INVALID). */
cbx_status cbx_dcsc_merge(int32_t count, const cbx_dcsc* parts, cbx_semiring semiring, cbx_dcsc* out, void* stream);

/* Edge list to matrix on the device: the Graph500 constructor SpTuples(maxnnz,
 * m, n, edges, removeloops) (SpTuples.cpp:70-110: values 1, duplicates summed,
 * self loops dropped) and SpTuples::RemoveDuplicates(BinOp) (SpTuples.cpp:271)
 * followed by SpDCCols(SpTuples).  rows, cols: device int64 arrays of nnz
 * entries in any order; vals: device array of nnz values of dtype, or NULL
 * for all ones.  Entries with equal (row, col) are combined with the
 * semiring's add; with remove_loops != 0 entries with row == col are
 * dropped.  Indices outside m x n are an error.  At most 2^31 - 1 entries. */
cbx_status cbx_dcsc_from_coo(int64_t m, int64_t n, const int64_t* rows, const int64_t* cols, const void* vals,
                             int64_t nnz, cbx_dtype dtype, cbx_semiring semiring, int32_t remove_loops, cbx_dcsc* out,
                             void* stream);

/* count independent products C[i] = A[i]*B[i] (arrays of count device
 * descriptors, one dtype) computed together: the batch is packed into one
 * block-diagonal product, so every SpGEMM kernel is launched once for the
 * whole batch, then split into count new device matrices.  Each C[i] equals
 * what cbx_spgemm(&A[i], &B[i], ...) returns.  1 <= count <= 4096; the packed
 * batch must hold fewer than 2^31 rows, columns and nonzeros.  Every member
 * is validated against its own shape (CBX_ERR_INVALID for a malformed one). */
cbx_status cbx_spgemm_batched(int64_t count, const cbx_dcsc* A, const cbx_dcsc* B, cbx_semiring semiring,
                              cbx_spgemm_mode mode, cbx_dcsc* C, cbx_spgemm_stats* stats, void* stream);

/* Batched Galerkin triple products (ReleaseTests/GalerkinNew.cpp:94-112):
 * for each i, S = T[i]^T, AT = A[i]*T[i], SAT[i] = S*AT with PlusTimes, where
 * A[i] is square and T[i] = R^T is the fine x coarse prolongation.  One
 * transpose and two batched products for the whole batch; members are
 * validated as in cbx_spgemm_batched. */
cbx_status cbx_galerkin_batched(int64_t count, const cbx_dcsc* A, const cbx_dcsc* T, cbx_dcsc* SAT,
                                cbx_spgemm_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CBX_H */
