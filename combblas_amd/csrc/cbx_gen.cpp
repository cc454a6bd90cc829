// Deterministic synthetic-matrix generators (host side, OpenMP).
//
// R-MAT follows the quadrant recursion of the Graph500 / CombBLAS generator
// (graph500-1.2/generator, used through DistEdgeList.cpp:223 GenGraph500Data)
// but without noise or vertex scrambling: every edge draws `scale` quadrants
// with probabilities (a, b, c, 1-a-b-c).  Randomness is counter based
// (splitmix64 of seed and edge id), so the matrix does not depend on the
// number of threads.  Output is DCSC with int64 indices, rows sorted and
// unique within each column, duplicates combined as the flags ask.
#include "cbx_internal.h"

#include <parallel/algorithm>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

namespace cbx {

static inline uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct EdgeRng {
  uint64_t s;
  EdgeRng(uint64_t seed, uint64_t edge) : s(splitmix(seed ^ splitmix(edge * 0xD1B54A32D192ED03ULL))) {}
  inline uint64_t next() { s += 0x9E3779B97F4A7C15ULL; return splitmix(s); }
  inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

struct Triple { uint64_t key; int64_t val; };  // key = col << 32 | row

// Sort triples by (col,row) and fold duplicates, then emit host DCSC.
static cbx_status triples_to_dcsc(std::vector<Triple>& t, int64_t m, int64_t n, bool pattern,
                                  cbx_dcsc* out) {
  __gnu_parallel::sort(t.begin(), t.end(),
                       [](const Triple& x, const Triple& y) { return x.key < y.key; });
  // fold duplicates in place
  size_t w = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    if (w > 0 && t[w - 1].key == t[i].key) {
      if (!pattern) t[w - 1].val += t[i].val;
    } else {
      t[w++] = t[i];
    }
  }
  t.resize(w);
  int64_t nnz = (int64_t)w;
  int64_t nzc = 0;
  for (size_t i = 0; i < w; ++i)
    if (i == 0 || (t[i].key >> 32) != (t[i - 1].key >> 32)) ++nzc;
  std::memset(out, 0, sizeof(*out));
  out->m = m; out->n = n; out->nnz = nnz; out->nzc = nzc;
  out->dtype = CBX_I64; out->location = CBX_HOST;
  out->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nzc + 1));
  out->jc = (int64_t*)std::malloc(sizeof(int64_t) * (nzc > 0 ? nzc : 1));
  out->ir = (int64_t*)std::malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  out->val = std::malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  if (!out->cp || !out->jc || !out->ir || !out->val) {
    cbx_dcsc_free(out);
    return set_error(CBX_ERR_NOMEM, "generator: host allocation failed");
  }
  int64_t* val = (int64_t*)out->val;
  int64_t c = 0;
  for (int64_t i = 0; i < nnz; ++i) {
    int64_t col = (int64_t)(t[i].key >> 32);
    if (i == 0 || col != (int64_t)(t[i - 1].key >> 32)) {
      out->jc[c] = col;
      out->cp[c] = i;
      ++c;
    }
    out->ir[i] = (int64_t)(t[i].key & 0xffffffffULL);
    val[i] = pattern ? 1 : t[i].val;
  }
  out->cp[nzc] = nnz;
  return CBX_OK;
}

}  // namespace cbx

using namespace cbx;

extern "C" cbx_status cbx_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                                   uint64_t seed, uint32_t flags, cbx_dcsc* out) {
  if (!out) return set_error(CBX_ERR_INVALID, "cbx_gen_rmat: out is null");
  if (scale < 1 || scale > 30 || edgefactor < 1)
    return set_error(CBX_ERR_INVALID, "cbx_gen_rmat: scale must be in [1,30], edgefactor >= 1");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return set_error(CBX_ERR_INVALID, "cbx_gen_rmat: quadrant probabilities out of range");
  const int64_t n = (int64_t)1 << scale;
  const int64_t nedges = n * (int64_t)edgefactor;
  const bool sym = flags & CBX_GEN_SYMMETRIZE;
  const bool noloops = flags & CBX_GEN_NO_SELF_LOOPS;
  const bool pattern = flags & CBX_GEN_PATTERN;
  const double ab = a + b, abc = a + b + c;
  std::vector<Triple> t;
  try {
    t.resize(sym ? 2 * nedges : nedges);
  } catch (...) {
    return set_error(CBX_ERR_NOMEM, "cbx_gen_rmat: host allocation failed");
  }
  std::vector<uint8_t> keep(nedges, 1);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nedges; ++e) {
    EdgeRng rng(seed, (uint64_t)e);
    uint64_t row = 0, col = 0;
    for (int l = 0; l < scale; ++l) {
      double u = rng.uniform();
      uint64_t rb, cb;
      if (u < a) { rb = 0; cb = 0; }
      else if (u < ab) { rb = 0; cb = 1; }
      else if (u < abc) { rb = 1; cb = 0; }
      else { rb = 1; cb = 1; }
      row = (row << 1) | rb;
      col = (col << 1) | cb;
    }
    int64_t v = 1 + (int64_t)(((rng.next() >> 32) * 9ULL) >> 32);  // uniform in [1,9]
    if (noloops && row == col) keep[e] = 0;
    t[e] = Triple{(col << 32) | row, v};
    if (sym) t[nedges + e] = Triple{(row << 32) | col, v};
  }
  if (noloops) {
    size_t w = 0;
    const size_t total = t.size();
    for (size_t i = 0; i < total; ++i) {
      int64_t e = (int64_t)(i < (size_t)nedges ? i : i - nedges);
      if (keep[e]) t[w++] = t[i];
    }
    t.resize(w);
  }
  return triples_to_dcsc(t, n, n, pattern, out);
}

// ----------------------------------------------------------------------------
// Power-law, column-stochastic matrix (the input of an MCL expansion step,
// Applications/MCL.cpp:573).  Vertex i gets degree
//   d_i = min(dmax, floor(xmin * u_i^(-1/(alpha-1)))),   u_i uniform in (0,1],
// with xmin found by bisection so that the sample mean of d equals `mean`.
// Edges come from the directed configuration model: column j holds d_j
// out-stubs, the row stubs are the same multiset (vertex i appears d_i times)
// in a seeded random order, and the e-th column stub is paired with the e-th
// row stub.  Every edge carries a log-normal(0,1) weight; duplicate (row,col)
// pairs add their weights, and each column is divided by its sum (in double)
// before rounding to float, so columns sum to 1 up to float rounding.
namespace cbx {

struct WTriple { uint64_t key; double val; };

static double degree_of(double xmin, double alpha, double u, int64_t dmax) {
  double d = std::floor(xmin * std::pow(u, -1.0 / (alpha - 1.0)));
  return d > (double)dmax ? (double)dmax : d;
}

}  // namespace cbx

extern "C" cbx_status cbx_gen_powerlaw_stochastic(int64_t n, double mean, double alpha, int64_t dmax,
                                                  uint64_t seed, cbx_dcsc* out) {
  if (!out) return set_error(CBX_ERR_INVALID, "cbx_gen_powerlaw_stochastic: out is null");
  if (n < 1 || n > ((int64_t)1 << 31) - 1)
    return set_error(CBX_ERR_INVALID, "cbx_gen_powerlaw_stochastic: n must be in [1, 2^31)");
  if (!(alpha > 1.0) || !(mean >= 1.0) || dmax < 1 || (double)dmax < mean)
    return set_error(CBX_ERR_INVALID, "cbx_gen_powerlaw_stochastic: need alpha > 1, mean >= 1, dmax >= mean");
  std::vector<double> u(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    EdgeRng r(seed ^ 0x5DE6A11ULL, (uint64_t)i);
    u[i] = 1.0 - r.uniform();  // (0,1]
  }
  auto sample_mean = [&](double xmin) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += degree_of(xmin, alpha, u[i], dmax);
    return s / (double)n;
  };
  double lo = 1.0, hi = mean + 1.0;
  if (sample_mean(lo) > mean) hi = lo;  // even xmin = 1 overshoots: use it
  else {
    while (sample_mean(hi) < mean && hi < 1e12) hi *= 2;
    for (int it = 0; it < 80; ++it) {
      double mid = 0.5 * (lo + hi);
      (sample_mean(mid) < mean ? lo : hi) = mid;
    }
  }
  const double xmin = hi;
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    int64_t d = (int64_t)degree_of(xmin, alpha, u[i], dmax);
    off[i + 1] = off[i] + (d < 1 ? 1 : d);
  }
  const int64_t nst = off[n];
  std::vector<WTriple> t;
  std::vector<std::pair<uint64_t, uint32_t>> rows;
  try {
    t.resize(nst);
    rows.resize(nst);
  } catch (...) {
    return set_error(CBX_ERR_NOMEM, "cbx_gen_powerlaw_stochastic: host allocation failed");
  }
  // row stubs in a seeded random order: sort (random key, vertex)
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t s = off[i]; s < off[i + 1]; ++s)
      rows[s] = {splitmix(seed ^ splitmix((uint64_t)s + 0x1234567ULL)), (uint32_t)i};
  __gnu_parallel::sort(rows.begin(), rows.end());
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t s = off[j]; s < off[j + 1]; ++s) {
      EdgeRng r(seed ^ 0x10A0A1ULL, (uint64_t)s);
      // Box-Muller: one standard normal from two uniforms
      double u1 = 1.0 - r.uniform(), u2 = r.uniform();
      double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      t[s] = WTriple{((uint64_t)j << 32) | rows[s].second, std::exp(z)};
    }
  std::vector<std::pair<uint64_t, uint32_t>>().swap(rows);
  __gnu_parallel::sort(t.begin(), t.end(), [](const WTriple& x, const WTriple& y) { return x.key < y.key; });
  size_t w = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    if (w > 0 && t[w - 1].key == t[i].key) t[w - 1].val += t[i].val;
    else t[w++] = t[i];
  }
  t.resize(w);
  const int64_t nnz = (int64_t)w;
  int64_t nzc = 0;
  for (size_t i = 0; i < w; ++i)
    if (i == 0 || (t[i].key >> 32) != (t[i - 1].key >> 32)) ++nzc;
  std::memset(out, 0, sizeof(*out));
  out->m = n; out->n = n; out->nnz = nnz; out->nzc = nzc;
  out->dtype = CBX_F32; out->location = CBX_HOST;
  out->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nzc + 1));
  out->jc = (int64_t*)std::malloc(sizeof(int64_t) * (nzc > 0 ? nzc : 1));
  out->ir = (int64_t*)std::malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  out->val = std::malloc(sizeof(float) * (nnz > 0 ? nnz : 1));
  if (!out->cp || !out->jc || !out->ir || !out->val) {
    cbx_dcsc_free(out);
    return set_error(CBX_ERR_NOMEM, "cbx_gen_powerlaw_stochastic: host allocation failed");
  }
  int64_t c = 0;
  for (int64_t i = 0; i < nnz; ++i) {
    if (i == 0 || (t[i].key >> 32) != (t[i - 1].key >> 32)) {
      out->jc[c] = (int64_t)(t[i].key >> 32);
      out->cp[c] = i;
      ++c;
    }
    out->ir[i] = (int64_t)(t[i].key & 0xffffffffULL);
  }
  out->cp[nzc] = nnz;
  float* val = (float*)out->val;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t q = 0; q < nzc; ++q) {
    double s = 0;
    for (int64_t p = out->cp[q]; p < out->cp[q + 1]; ++p) s += t[p].val;
    for (int64_t p = out->cp[q]; p < out->cp[q + 1]; ++p) val[p] = (float)(t[p].val / s);
  }
  return CBX_OK;
}

// ----------------------------------------------------------------------------
// Galerkin operands (ReleaseTests/GalerkinNew.cpp reads them from files:
// grid3d_k5.txt and restrict_T_grid3d_k5.txt).  Vertex (x,y,z) of a g^3 grid
// is row/column x + g*(y + g*z).
namespace cbx {

static cbx_status alloc_host_dcsc(cbx_dcsc* out, int64_t m, int64_t n, int64_t nnz, int64_t nzc, cbx_dtype dt,
                                  size_t vs) {
  std::memset(out, 0, sizeof(*out));
  out->m = m; out->n = n; out->nnz = nnz; out->nzc = nzc;
  out->dtype = dt; out->location = CBX_HOST;
  out->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nzc + 1));
  out->jc = (int64_t*)std::malloc(sizeof(int64_t) * (nzc > 0 ? nzc : 1));
  out->ir = (int64_t*)std::malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
  out->val = std::malloc(vs * (nnz > 0 ? nnz : 1));
  if (!out->cp || !out->jc || !out->ir || !out->val) {
    cbx_dcsc_free(out);
    return set_error(CBX_ERR_NOMEM, "generator: host allocation failed");
  }
  return CBX_OK;
}

}  // namespace cbx

// 27-point Poisson stencil: 26 on the diagonal, -1 for each of the up to 26
// neighbours inside the grid (Dirichlet boundary).  Symmetric, FP64.
extern "C" cbx_status cbx_gen_stencil27(int32_t g, cbx_dcsc* out) {
  if (!out) return set_error(CBX_ERR_INVALID, "cbx_gen_stencil27: out is null");
  if (g < 1 || g > 1024) return set_error(CBX_ERR_INVALID, "cbx_gen_stencil27: edge must be in [1, 1024]");
  const int64_t n = (int64_t)g * g * g;
  auto deg = [g](int64_t v) {
    const int x = (int)(v % g), y = (int)((v / g) % g), z = (int)(v / ((int64_t)g * g));
    const int a = 1 + (x > 0) + (x < g - 1), b = 1 + (y > 0) + (y < g - 1), c = 1 + (z > 0) + (z < g - 1);
    return (int64_t)(a * b * c);
  };
  std::vector<int64_t> cp(n + 1, 0);
  for (int64_t v = 0; v < n; ++v) cp[v + 1] = cp[v] + deg(v);
  cbx_status st = alloc_host_dcsc(out, n, n, cp[n], n, CBX_F64, sizeof(double));
  if (st) return st;
  std::memcpy(out->cp, cp.data(), sizeof(int64_t) * (n + 1));
  double* val = (double*)out->val;
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; ++v) {
    out->jc[v] = v;
    const int x = (int)(v % g), y = (int)((v / g) % g), z = (int)(v / ((int64_t)g * g));
    int64_t w = cp[v];
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {  // rows ascend: z slowest, x fastest
          const int X = x + dx, Y = y + dy, Z = z + dz;
          if (X < 0 || Y < 0 || Z < 0 || X >= g || Y >= g || Z >= g) continue;
          out->ir[w] = X + (int64_t)g * (Y + (int64_t)g * Z);
          val[w] = (dx | dy | dz) ? -1.0 : 26.0;
          ++w;
        }
  }
  return CBX_OK;
}

// 2x2x2 aggregation: coarse vertex (x/2, y/2, z/2) of a ceil(g/2)^3 grid.
// transposed = 0: R, coarse x fine, one 1.0 per column.  transposed = 1:
// T = R^T, fine x coarse, the up to 8 fine vertices of each aggregate in a
// column (the restrict_T input of GalerkinNew.cpp).
extern "C" cbx_status cbx_gen_aggregation(int32_t g, int32_t transposed, cbx_dcsc* out) {
  if (!out) return set_error(CBX_ERR_INVALID, "cbx_gen_aggregation: out is null");
  if (g < 1 || g > 1024) return set_error(CBX_ERR_INVALID, "cbx_gen_aggregation: edge must be in [1, 1024]");
  const int gc = (g + 1) / 2;
  const int64_t n = (int64_t)g * g * g, nc = (int64_t)gc * gc * gc;
  if (!transposed) {
    cbx_status st = alloc_host_dcsc(out, nc, n, n, n, CBX_F64, sizeof(double));
    if (st) return st;
    double* val = (double*)out->val;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) {
      const int x = (int)(v % g), y = (int)((v / g) % g), z = (int)(v / ((int64_t)g * g));
      out->cp[v] = v;
      out->jc[v] = v;
      out->ir[v] = x / 2 + (int64_t)gc * (y / 2 + (int64_t)gc * (z / 2));
      val[v] = 1.0;
    }
    out->cp[n] = n;
    return CBX_OK;
  }
  auto span = [g](int c) { return (2 * c + 1 < g) ? 2 : 1; };
  std::vector<int64_t> cp(nc + 1, 0);
  for (int64_t J = 0; J < nc; ++J) {
    const int X = (int)(J % gc), Y = (int)((J / gc) % gc), Z = (int)(J / ((int64_t)gc * gc));
    cp[J + 1] = cp[J] + span(X) * span(Y) * span(Z);
  }
  cbx_status st = alloc_host_dcsc(out, n, nc, n, nc, CBX_F64, sizeof(double));
  if (st) return st;
  std::memcpy(out->cp, cp.data(), sizeof(int64_t) * (nc + 1));
  double* val = (double*)out->val;
#pragma omp parallel for schedule(static)
  for (int64_t J = 0; J < nc; ++J) {
    out->jc[J] = J;
    const int X = (int)(J % gc), Y = (int)((J / gc) % gc), Z = (int)(J / ((int64_t)gc * gc));
    int64_t w = cp[J];
    for (int dz = 0; dz < span(Z); ++dz)
      for (int dy = 0; dy < span(Y); ++dy)
        for (int dx = 0; dx < span(X); ++dx) {
          out->ir[w] = (2 * X + dx) + (int64_t)g * ((2 * Y + dy) + (int64_t)g * (2 * Z + dz));
          val[w] = 1.0;
          ++w;
        }
  }
  return CBX_OK;
}
