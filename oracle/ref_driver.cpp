// ORACLE -- test infrastructure only.
//
// A thin C entry layer over the genuine CombBLAS headers, compiled from the
// reference tree where it lies (see oracle/Makefile) into
// oracle/_ref/libcbref.so.  It exists to pin the restatement in
// oracle/cbx_oracle.cpp and to write golden fixtures; it is never part of
// the product and never runs on a machine without the reference tree at
// build time.
#include <mpi.h>
#include <omp.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "CombBLAS/CombBLAS.h"
#include "../include/cbx.h"

// Globals the CombBLAS headers declare extern (CombBLAS.h:76-102).
int cblas_splits = 1;
double cblas_alltoalltime, cblas_allgathertime, cblas_localspmvtime, cblas_mergeconttime, cblas_transvectime;
double mcl_Abcasttime, mcl_Bbcasttime, mcl_localspgemmtime, mcl_multiwaymergetime, mcl_kselecttime,
    mcl_prunecolumntime, mcl_symbolictime;
double mcl3d_conversiontime, mcl3d_symbolictime, mcl3d_Abcasttime, mcl3d_Bbcasttime, mcl3d_SUMMAtime,
    mcl3d_localspgemmtime, mcl3d_SUMMAmergetime, mcl3d_reductiontime, mcl3d_3dmergetime, mcl3d_kselecttime;

using namespace combblas;

namespace {

void ensure_mpi() {
  int inited = 0;
  MPI_Initialized(&inited);
  if (!inited) {
    int provided;
    MPI_Init_thread(nullptr, nullptr, MPI_THREAD_SERIALIZED, &provided);
  }
}

template <class NT, class ST>
SpDCCols<int64_t, NT>* make_mat(const cbx_dcsc& X) {
  auto* M = new SpDCCols<int64_t, NT>(X.nnz, X.m, X.n, X.nzc);
  if (X.nnz == 0) return M;
  Dcsc<int64_t, NT>* d = M->GetDCSC();
  std::memcpy(d->cp, X.cp, sizeof(int64_t) * (X.nzc + 1));
  std::memcpy(d->jc, X.jc, sizeof(int64_t) * X.nzc);
  std::memcpy(d->ir, X.ir, sizeof(int64_t) * X.nnz);
  const ST* v = (const ST*)X.val;
  for (int64_t i = 0; i < X.nnz; ++i) d->numx[i] = (NT)v[i];
  return M;
}

template <class NT, class ST>
void export_mat(const SpTuples<int64_t, NT>& T, int64_t m, int64_t n, cbx_dtype dt, cbx_dcsc* C) {
  std::memset(C, 0, sizeof(*C));
  C->m = m;
  C->n = n;
  C->dtype = dt;
  C->location = CBX_HOST;
  const int64_t nnz = T.getnnz();
  int64_t nzc = 0;
  for (int64_t i = 0; i < nnz; ++i)
    if (i == 0 || T.colindex(i) != T.colindex(i - 1)) ++nzc;
  C->nnz = nnz;
  C->nzc = nzc;
  C->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nzc + 1));
  C->jc = (int64_t*)std::malloc(sizeof(int64_t) * (nzc ? nzc : 1));
  C->ir = (int64_t*)std::malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  ST* v = (ST*)std::malloc(sizeof(ST) * (nnz ? nnz : 1));
  C->val = v;
  int64_t c = 0;
  for (int64_t i = 0; i < nnz; ++i) {
    if (i == 0 || T.colindex(i) != T.colindex(i - 1)) {
      C->jc[c] = T.colindex(i);
      C->cp[c] = i;
      ++c;
    }
    C->ir[i] = T.rowindex(i);
    v[i] = (ST)T.numvalue(i);
  }
  C->cp[nzc] = nnz;
}

template <class SR, class NT, class ST>
int run(const cbx_dcsc* A, const cbx_dcsc* B, int mode, cbx_dcsc* C) {
  ensure_mpi();
  auto* MA = make_mat<NT, ST>(*A);
  auto* MB = make_mat<NT, ST>(*B);
  SpTuples<int64_t, NT>* T;
  if (mode == 0)
    T = LocalHybridSpGEMM<SR, NT>(*MA, *MB, false, false);
  else
    T = LocalSpGEMMHash<SR, NT>(*MA, *MB, false, false, mode == 1);
  export_mat<NT, ST>(*T, A->m, B->n, A->dtype, C);
  delete T;
  delete MA;
  delete MB;
  return 0;
}

template <class NT>
void export_dcsc(SpDCCols<int64_t, NT>& M, cbx_dtype dt, cbx_dcsc* C) {
  std::memset(C, 0, sizeof(*C));
  C->m = M.getnrow();
  C->n = M.getncol();
  C->dtype = dt;
  C->location = CBX_HOST;
  const int64_t nnz = M.getnnz();
  Dcsc<int64_t, NT>* d = nnz > 0 ? M.GetDCSC() : nullptr;
  const int64_t nzc = d ? d->nzc : 0;
  C->nnz = nnz;
  C->nzc = nzc;
  C->cp = (int64_t*)std::malloc(sizeof(int64_t) * (nzc + 1));
  C->jc = (int64_t*)std::malloc(sizeof(int64_t) * (nzc ? nzc : 1));
  C->ir = (int64_t*)std::malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  NT* v = (NT*)std::malloc(sizeof(NT) * (nnz ? nnz : 1));
  C->val = v;
  C->cp[0] = 0;
  if (!d) return;
  std::memcpy(C->cp, d->cp, sizeof(int64_t) * (nzc + 1));
  std::memcpy(C->jc, d->jc, sizeof(int64_t) * nzc);
  std::memcpy(C->ir, d->ir, sizeof(int64_t) * nnz);
  std::memcpy(v, d->numx, sizeof(NT) * nnz);
}

// MultiwayMerge<SR> (MultiwayMerge.h:428) of the parts' column-sorted tuples.
template <class SR, class NT, class ST>
int merge(int32_t count, const cbx_dcsc* parts, cbx_dcsc* C) {
  ensure_mpi();
  std::vector<SpDCCols<int64_t, NT>*> mats;
  std::vector<SpTuples<int64_t, NT>*> tups;
  for (int32_t i = 0; i < count; ++i) {
    mats.push_back(make_mat<NT, ST>(parts[i]));
    tups.push_back(new SpTuples<int64_t, NT>(*mats.back()));
  }
  const int64_t m = count ? parts[0].m : 0, n = count ? parts[0].n : 0;
  SpTuples<int64_t, NT>* T = MultiwayMerge<SR>(tups, m, n, false);
  export_mat<NT, ST>(*T, m, n, count ? parts[0].dtype : CBX_I64, C);
  if (!(count == 1 && T == tups[0])) delete T;
  for (auto* t : tups) delete t;
  for (auto* x : mats) delete x;
  return 0;
}

// MCLPruneRecoverySelect (ParFriends.h:185) on a 1x1 process grid.
template <class NT>
int mcl_prune(const cbx_dcsc* A, double thr, int64_t select, int64_t recover, double pct, cbx_dcsc* C) {
  ensure_mpi();
  typedef SpDCCols<int64_t, NT> DER;
  std::shared_ptr<CommGrid> grid(new CommGrid(MPI_COMM_WORLD, 1, 1));
  SpParMat<int64_t, NT, DER> M(make_mat<NT, NT>(*A), grid);
  MCLPruneRecoverySelect(M, (NT)thr, (int64_t)select, (int64_t)recover, (NT)pct, 1);
  export_dcsc<NT>(M.seq(), A->dtype, C);
  return 0;
}

}  // namespace

// generic_gespmv<SR> on an A of the vector's value type (heap form).
template <class SR, class NT>
static int32_t ref_spmspv_typed(const cbx_dcsc* A, const int32_t* xind, const NT* xval, int32_t nx, int32_t* yind,
                                NT* yval) {
  ensure_mpi();
  auto* MA = make_mat<NT, NT>(*A);
  std::vector<int32_t> indy;
  std::vector<NT> numy;
  PreAllocatedSPA<NT> spa;
  generic_gespmv<SR>(*MA, xind, xval, nx, indy, numy, spa);
  std::copy(indy.begin(), indy.end(), yind);
  std::copy(numy.begin(), numy.end(), yval);
  delete MA;
  return (int32_t)indy.size();
}

extern "C" {

int32_t ref_mcl_prune(const cbx_dcsc* A, double thr, int64_t select, int64_t recover, double pct, cbx_dcsc* C) {
  if (A->dtype == CBX_F32) return mcl_prune<float>(A, thr, select, recover, pct, C);
  if (A->dtype == CBX_F64) return mcl_prune<double>(A, thr, select, recover, pct, C);
  return CBX_ERR_UNSUPPORTED;
}

// mode 0: LocalHybridSpGEMM; 1: LocalSpGEMMHash(sort=true); 2: sort=false.
int ref_spgemm(const cbx_dcsc* A, const cbx_dcsc* B, int32_t semiring, int32_t mode, cbx_dcsc* C) {
  switch (semiring) {
    case CBX_SR_PLUS_TIMES:
      switch (A->dtype) {
        case CBX_I64: return run<PlusTimesSRing<int64_t, int64_t>, int64_t, int64_t>(A, B, mode, C);
        case CBX_F64: return run<PlusTimesSRing<double, double>, double, double>(A, B, mode, C);
        case CBX_F32: return run<PlusTimesSRing<float, float>, float, float>(A, B, mode, C);
        case CBX_BOOL: return run<PlusTimesSRing<bool, bool>, bool, uint8_t>(A, B, mode, C);
      }
      break;
    case CBX_SR_MIN_PLUS:
      switch (A->dtype) {
        case CBX_I64: return run<MinPlusSRing<int64_t, int64_t>, int64_t, int64_t>(A, B, mode, C);
        case CBX_F64: return run<MinPlusSRing<double, double>, double, double>(A, B, mode, C);
        case CBX_F32: return run<MinPlusSRing<float, float>, float, float>(A, B, mode, C);
      }
      break;
    case CBX_SR_SELECT_MAX:
      switch (A->dtype) {
        case CBX_I64: return run<SelectMaxSRing<int64_t, int64_t>, int64_t, int64_t>(A, B, mode, C);
        case CBX_F64: return run<SelectMaxSRing<double, double>, double, double>(A, B, mode, C);
        case CBX_F32: return run<SelectMaxSRing<float, float>, float, float>(A, B, mode, C);
      }
      break;
  }
  return -1;
}

// SpTuples' Graph500 constructor (SpTuples.cpp:70): values 1, duplicates
// summed, self loops dropped when removeloops.
int ref_graph500_tuples(int64_t m, int64_t n, const int64_t* edges, int64_t nedges, int32_t removeloops,
                        cbx_dcsc* C) {
  std::vector<int64_t> e(edges, edges + 2 * nedges);
  SpTuples<int64_t, int64_t> T(nedges, m, n, e, removeloops != 0);
  export_mat<int64_t, int64_t>(T, m, n, CBX_I64, C);
  return 0;
}

int ref_multiway_merge(int32_t count, const cbx_dcsc* parts, int32_t semiring, cbx_dcsc* C) {
  if (count < 1) return CBX_ERR_INVALID;
  switch (semiring) {
    case CBX_SR_PLUS_TIMES:
      switch (parts[0].dtype) {
        case CBX_I64: return merge<PlusTimesSRing<int64_t, int64_t>, int64_t, int64_t>(count, parts, C);
        case CBX_F64: return merge<PlusTimesSRing<double, double>, double, double>(count, parts, C);
        case CBX_F32: return merge<PlusTimesSRing<float, float>, float, float>(count, parts, C);
        case CBX_BOOL: return merge<PlusTimesSRing<bool, bool>, bool, uint8_t>(count, parts, C);
      }
      break;
    case CBX_SR_MIN_PLUS:
      switch (parts[0].dtype) {
        case CBX_I64: return merge<MinPlusSRing<int64_t, int64_t>, int64_t, int64_t>(count, parts, C);
        case CBX_F64: return merge<MinPlusSRing<double, double>, double, double>(count, parts, C);
        case CBX_F32: return merge<MinPlusSRing<float, float>, float, float>(count, parts, C);
      }
      break;
    case CBX_SR_SELECT_MAX:
      switch (parts[0].dtype) {
        case CBX_I64: return merge<SelectMaxSRing<int64_t, int64_t>, int64_t, int64_t>(count, parts, C);
        case CBX_F64: return merge<SelectMaxSRing<double, double>, double, double>(count, parts, C);
        case CBX_F32: return merge<SelectMaxSRing<float, float>, float, float>(count, parts, C);
      }
      break;
  }
  return -1;
}

// estimateFLOP + estimateNNZ_Hash of the reference.
int ref_symbolic(const cbx_dcsc* A, const cbx_dcsc* B, int64_t* flop, int64_t* nnz) {
  ensure_mpi();
  auto* MA = make_mat<int64_t, int64_t>(*A);
  auto* MB = make_mat<int64_t, int64_t>(*B);
  int64_t* f = estimateFLOP(*MA, *MB);
  int64_t* z = estimateNNZ_Hash(*MA, *MB, f);
  std::memcpy(flop, f, sizeof(int64_t) * B->nzc);
  std::memcpy(nnz, z, sizeof(int64_t) * B->nzc);
  delete[] f;
  delete[] z;
  delete MA;
  delete MB;
  return 0;
}

// estimateNNZ_sampling (mtSpGEMM.h:936-1054), the genuine template: its
// global estimate of nnz(A*B) with nrounds samples per row.
int64_t ref_estimate_sampling(const cbx_dcsc* A, const cbx_dcsc* B, int32_t nrounds) {
  ensure_mpi();
  auto* MA = make_mat<int64_t, int64_t>(*A);
  auto* MB = make_mat<int64_t, int64_t>(*B);
  const int64_t est = estimateNNZ_sampling(*MA, *MB, nrounds);
  delete MA;
  delete MB;
  return est;
}

// Times combblas::LocalHybridSpGEMM<PlusTimesSRing<int64,int64>,int64>(A, B)
// alone (operands converted to SpDCCols beforehand; the returned SpTuples is
// deleted after the clock stops).
int ref_time_hybrid_i64(const cbx_dcsc* A, const cbx_dcsc* B, double* secs, int64_t* nnz) {
  ensure_mpi();
  auto* MA = make_mat<int64_t, int64_t>(*A);
  auto* MB = make_mat<int64_t, int64_t>(*B);
  const double t0 = omp_get_wtime();
  SpTuples<int64_t, int64_t>* T = LocalHybridSpGEMM<PlusTimesSRing<int64_t, int64_t>, int64_t>(*MA, *MB, false, false);
  const double t1 = omp_get_wtime();
  *secs = t1 - t0;
  *nnz = T->getnnz();
  delete T;
  delete MA;
  delete MB;
  return 0;
}

// LocalHybridSpGEMM<PlusTimesSRing<int64,int64>> alone (secs_product) and
// together with the SpDCCols(SpTuples) conversion the SUMMA drivers apply
// to its output (SpDCCols.cpp:109-184; secs_dcsc covers both).
int ref_time_hybrid_dcsc_i64(const cbx_dcsc* A, const cbx_dcsc* B, double* secs_product, double* secs_dcsc,
                             int64_t* nnz) {
  ensure_mpi();
  auto* MA = make_mat<int64_t, int64_t>(*A);
  auto* MB = make_mat<int64_t, int64_t>(*B);
  const double t0 = omp_get_wtime();
  SpTuples<int64_t, int64_t>* T = LocalHybridSpGEMM<PlusTimesSRing<int64_t, int64_t>, int64_t>(*MA, *MB, false, false);
  const double t1 = omp_get_wtime();
  auto* D = new SpDCCols<int64_t, int64_t>(*T, false);
  const double t2 = omp_get_wtime();
  *secs_product = t1 - t0;
  *secs_dcsc = t2 - t0;
  *nnz = D->getnnz();
  delete D;
  delete T;
  delete MA;
  delete MB;
  return 0;
}

// generic_gespmv<SelectMaxSRing<bool,int64_t>> (Friends.h:448-462) on a
// boolean A: the reference's SpMXSpV with its sequence heap.
int32_t ref_spmspv_selectmax_bool(const cbx_dcsc* A, const int32_t* xind, const int64_t* xval, int32_t nx,
                                  int32_t* yind, int64_t* yval) {
  ensure_mpi();
  auto* MA = make_mat<bool, uint8_t>(*A);
  std::vector<int32_t> indy;
  std::vector<int64_t> numy;
  PreAllocatedSPA<int64_t> spa;
  generic_gespmv<SelectMaxSRing<bool, int64_t>>(*MA, xind, xval, nx, indy, numy, spa);
  std::copy(indy.begin(), indy.end(), yind);
  std::copy(numy.begin(), numy.end(), yval);
  delete MA;
  return (int32_t)indy.size();
}

// generic_gespmv<PlusTimesSRing<int64,int64>> on an int64 A (heap form).
int32_t ref_spmspv_plustimes_i64(const cbx_dcsc* A, const int32_t* xind, const int64_t* xval, int32_t nx,
                                 int32_t* yind, int64_t* yval) {
  ensure_mpi();
  auto* MA = make_mat<int64_t, int64_t>(*A);
  std::vector<int32_t> indy;
  std::vector<int64_t> numy;
  PreAllocatedSPA<int64_t> spa;
  generic_gespmv<PlusTimesSRing<int64_t, int64_t>>(*MA, xind, xval, nx, indy, numy, spa);
  std::copy(indy.begin(), indy.end(), yind);
  std::copy(numy.begin(), numy.end(), yval);
  delete MA;
  return (int32_t)indy.size();
}

// generic_gespmv<MinPlusSRing<int64,int64>> (inf_plus multiply, min add) and
// generic_gespmv<PlusTimesSRing<double,double>> on A of the same value type.
int32_t ref_spmspv_minplus_i64(const cbx_dcsc* A, const int32_t* xind, const int64_t* xval, int32_t nx,
                               int32_t* yind, int64_t* yval) {
  return ref_spmspv_typed<MinPlusSRing<int64_t, int64_t>, int64_t>(A, xind, xval, nx, yind, yval);
}

int32_t ref_spmspv_plustimes_f64(const cbx_dcsc* A, const int32_t* xind, const double* xval, int32_t nx,
                                 int32_t* yind, double* yval) {
  return ref_spmspv_typed<PlusTimesSRing<double, double>, double>(A, xind, xval, nx, yind, yval);
}

// The TopDownBFS level loop (TopDownBFS.cpp:442-448) on one process with the
// reference's local SpMSpV; the vector glue restates FullyDistSpVec's
// setNumToInd (FullyDistSpVec.h:250-259), EWiseMult with exclude
// (ParFriends.h:2700-2760) and FullyDistVec::Set (FullyDistVec.cpp:1036).
int32_t ref_bfs(const cbx_dcsc* A, int64_t root, int64_t* parents) {
  ensure_mpi();
  auto* MA = make_mat<bool, uint8_t>(*A);
  const int64_t n = A->n;
  std::fill(parents, parents + n, (int64_t)-1);
  std::vector<int32_t> ind(1, (int32_t)root), indy;
  std::vector<int64_t> num, numy;
  PreAllocatedSPA<int64_t> spa;
  int32_t it = 0;
  while (!ind.empty()) {
    num.assign(ind.begin(), ind.end());
    indy.clear();
    numy.clear();
    generic_gespmv<SelectMaxSRing<bool, int64_t>>(*MA, ind.data(), num.data(), (int32_t)ind.size(), indy, numy, spa);
    ind.clear();
    num.clear();
    for (size_t i = 0; i < indy.size(); ++i)
      if (parents[indy[i]] == -1) {
        ind.push_back(indy[i]);
        num.push_back(numy[i]);
      }
    for (size_t i = 0; i < ind.size(); ++i) parents[ind[i]] = num[i];
    ++it;
  }
  delete MA;
  return it;
}

// The TopDownBFS level loop with the threaded local SpMSpV of a
// multithreaded run: the matrix row-split into `splits` pieces
// (SpParMat::ActivateThreading -> SpDCCols::RowSplit, TopDownBFS.cpp:178)
// and each level's product through generic_gespmv_threaded with one
// recipient (Friends.h:181-320; LocalSpMV, ParFriends.h:1965), default SPA
// as in SpMV(Aeff, fringe, optbuf) (TopDownBFS.cpp:445).  For `nroots`
// roots: parents of the last root, the summed seconds of the level loops
// (the row split is done once, outside the clock) and the summed levels.
int32_t ref_bfs_threaded(const cbx_dcsc* A, const int64_t* roots, int32_t nroots, int32_t splits, int64_t* parents,
                         double* secs) {
  ensure_mpi();
  auto* MA = make_mat<bool, uint8_t>(*A);
  MA->RowSplit(splits);
  const int64_t n = A->n;
  int32_t total_levels = 0;
  double t = 0.0;
  std::vector<int32_t> ind, nind;
  std::vector<int64_t> num, nnum;
  for (int32_t q = 0; q < nroots; ++q) {
    const double t0 = omp_get_wtime();
    std::fill(parents, parents + n, (int64_t)-1);
    ind.assign(1, (int32_t)roots[q]);
    while (!ind.empty()) {
      num.assign(ind.begin(), ind.end());  // setNumToInd
      int32_t* sind = nullptr;
      int64_t* snum = nullptr;
      int* sdis = nullptr;
      PreAllocatedSPA<int64_t> spa;
      const int ny = generic_gespmv_threaded<SelectMaxSRing<bool, int64_t>>(*MA, ind.data(), num.data(),
                                                                             (int32_t)ind.size(), sind, snum, sdis,
                                                                             1, spa);
      nind.clear();
      nnum.clear();
      for (int i = 0; i < ny; ++i)  // EWiseMult(fringe, parents, exclude=true, -1)
        if (parents[sind[i]] == -1) {
          nind.push_back(sind[i]);
          nnum.push_back(snum[i]);
        }
      for (size_t i = 0; i < nind.size(); ++i) parents[nind[i]] = nnum[i];  // parents.Set(fringe)
      delete[] sind;
      delete[] snum;
      delete[] sdis;
      ind.swap(nind);
      ++total_levels;
    }
    t += omp_get_wtime() - t0;
  }
  if (secs) *secs = t;
  delete MA;
  return total_levels;
}

// dcsc_gespmv<SR>(A, x, y) (Friends.h:63-77), the genuine template.
int32_t ref_spmv_dense(const cbx_dcsc* A, const void* x, void* y, int32_t semiring) {
  ensure_mpi();
  if (A->dtype == CBX_BOOL) {  // SelectMax over a boolean matrix, int64 vectors (the BFS form)
    auto* M = make_mat<bool, uint8_t>(*A);
    dcsc_gespmv<SelectMaxSRing<bool, int64_t>>(*M, (const int64_t*)x, (int64_t*)y);
    delete M;
    return 0;
  }
  if (A->dtype == CBX_I64) {
    auto* M = make_mat<int64_t, int64_t>(*A);
    if (semiring == CBX_SR_PLUS_TIMES)
      dcsc_gespmv<PlusTimesSRing<int64_t, int64_t>>(*M, (const int64_t*)x, (int64_t*)y);
    else if (semiring == CBX_SR_MIN_PLUS)
      dcsc_gespmv<MinPlusSRing<int64_t, int64_t>>(*M, (const int64_t*)x, (int64_t*)y);
    else
      dcsc_gespmv<SelectMaxSRing<int64_t, int64_t>>(*M, (const int64_t*)x, (int64_t*)y);
    delete M;
    return 0;
  }
  if (A->dtype == CBX_F64 && semiring == CBX_SR_PLUS_TIMES) {
    auto* M = make_mat<double, double>(*A);
    dcsc_gespmv<PlusTimesSRing<double, double>>(*M, (const double*)x, (double*)y);
    delete M;
    return 0;
  }
  return CBX_ERR_UNSUPPORTED;
}

void ref_free(cbx_dcsc* C) {
  std::free(C->cp);
  std::free(C->jc);
  std::free(C->ir);
  std::free(C->val);
  std::memset(C, 0, sizeof(*C));
}

}  // extern "C"
