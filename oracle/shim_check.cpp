// ORACLE -- test infrastructure only.
//
// Integration check of include/cbx_combblas.hpp: genuine CombBLAS objects
// (SpDCCols built from an R-MAT matrix) go through the GPU drop-ins and
// through the reference's own LocalHybridSpGEMM / LocalSpGEMMHash,
// SpMXSpV and generic_gespmv_threaded (on a row-split boolean matrix, the
// BFS form) on the CPU; the results must be identical (integer semirings,
// boolean).  Built by
// oracle/Makefile (target `ref`) into oracle/_ref/shim_check against the
// reference headers where they lie; exits 77 when no GPU is visible.
#include <mpi.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <vector>

#include "CombBLAS/CombBLAS.h"
#include "../include/cbx_combblas.hpp"

int cblas_splits = 1;
double cblas_alltoalltime, cblas_allgathertime, cblas_localspmvtime, cblas_mergeconttime, cblas_transvectime;
double mcl_Abcasttime, mcl_Bbcasttime, mcl_localspgemmtime, mcl_multiwaymergetime, mcl_kselecttime,
    mcl_prunecolumntime, mcl_symbolictime;
double mcl3d_conversiontime, mcl3d_symbolictime, mcl3d_Abcasttime, mcl3d_Bbcasttime, mcl3d_SUMMAtime,
    mcl3d_localspgemmtime, mcl3d_SUMMAmergetime, mcl3d_reductiontime, mcl3d_3dmergetime, mcl3d_kselecttime;

using namespace combblas;

template <class NT>
SpDCCols<int64_t, NT>* from_cbx(const cbx_dcsc& X) {
  auto* M = new SpDCCols<int64_t, NT>(X.nnz, X.m, X.n, X.nzc);
  Dcsc<int64_t, NT>* d = M->GetDCSC();
  std::memcpy(d->cp, X.cp, sizeof(int64_t) * (X.nzc + 1));
  std::memcpy(d->jc, X.jc, sizeof(int64_t) * X.nzc);
  std::memcpy(d->ir, X.ir, sizeof(int64_t) * X.nnz);
  const int64_t* v = (const int64_t*)X.val;
  for (int64_t i = 0; i < X.nnz; ++i) d->numx[i] = (NT)(v[i] % 7 - 3);  // signed values
  return M;
}

template <class IT, class NT>
std::vector<std::tuple<IT, IT, NT>> sorted_tuples(const SpTuples<IT, NT>& T) {
  std::vector<std::tuple<IT, IT, NT>> v;
  for (int64_t i = 0; i < T.getnnz(); ++i) v.emplace_back(T.colindex(i), T.rowindex(i), T.numvalue(i));
  std::sort(v.begin(), v.end());
  return v;
}

template <class SR, class NT>
bool check(const cbx_dcsc& G, const char* name) {
  auto* A = from_cbx<NT>(G);
  SpTuples<int64_t, NT>* ref = LocalHybridSpGEMM<SR, NT>(*A, *A, false, false);
  SpTuples<int64_t, NT>* gpu = LocalHybridSpGEMM_HIP<SR, NT>(*A, *A, false, false);
  bool ok = ref->getnnz() == gpu->getnnz();
  for (int64_t i = 0; ok && i < ref->getnnz(); ++i)  // same order: column-major, rows ascending
    ok = ref->rowindex(i) == gpu->rowindex(i) && ref->colindex(i) == gpu->colindex(i) &&
         ref->numvalue(i) == gpu->numvalue(i);
  SpTuples<int64_t, NT>* gh = LocalSpGEMMHash_HIP<SR, NT>(*A, *A, false, false, false);
  ok = ok && sorted_tuples(*gh) == sorted_tuples(*ref);
  SpDCCols<int64_t, NT>* gd = LocalHybridSpGEMM_HIP_DCSC<SR, NT>(*A, *A, false, false);
  SpDCCols<int64_t, NT> rd(*ref, false);
  ok = ok && (*gd == rd);
  std::printf("%-28s nnz=%lld %s\n", name, (long long)ref->getnnz(), ok ? "identical" : "DIFFERENT");
  delete ref;
  delete gpu;
  delete gh;
  delete gd;
  delete A;
  return ok;
}

// A sparse x over every 7th column of A, integer values in [-5, 5].
template <class VT>
void make_x(int64_t n, std::vector<int32_t>& xi, std::vector<VT>& xv) {
  for (int32_t j = 0; j < n; j += 7) {
    xi.push_back(j);
    xv.push_back((VT)(j % 11 - 5));
  }
}

template <class SR, class NUM, class VT>
bool check_spmxspv(const cbx_dcsc& G, const char* name) {
  auto* A = from_cbx<NUM>(G);
  std::vector<int32_t> xi;
  std::vector<VT> xv;
  make_x<VT>(G.n, xi, xv);
  std::vector<int32_t> ri, gi;
  std::vector<VT> rv, gv;
  PreAllocatedSPA<VT> spa;
  SpMXSpV<SR>(*A->GetDCSC(), (int32_t)G.m, xi.data(), xv.data(), (int32_t)xi.size(), ri, rv, spa);
  SpMXSpV_HIP<SR>(*A->GetDCSC(), (int32_t)G.m, xi.data(), xv.data(), (int32_t)xi.size(), gi, gv, spa);
  const bool ok = !ri.empty() && ri == gi && rv == gv;
  std::printf("%-28s ny=%zu %s\n", name, ri.size(), ok ? "identical" : "DIFFERENT");
  delete A;
  return ok;
}

template <class SR, class VT>
bool check_gespmv_threaded(const cbx_dcsc& G, int splits, int p_c, const char* name) {
  auto* A = from_cbx<bool>(G);
  A->RowSplit(splits);
  std::vector<int32_t> xi;
  std::vector<VT> xv;
  make_x<VT>(G.n, xi, xv);
  PreAllocatedSPA<VT> spa;
  int32_t *ri = nullptr, *gi = nullptr;
  VT *rv = nullptr, *gv = nullptr;
  int *rd = nullptr, *gd = nullptr;
  const int rn = generic_gespmv_threaded<SR>(*A, xi.data(), xv.data(), (int32_t)xi.size(), ri, rv, rd, p_c, spa);
  const int gn = generic_gespmv_threaded_HIP<SR>(*A, xi.data(), xv.data(), (int32_t)xi.size(), gi, gv, gd, p_c, spa);
  bool ok = rn > 0 && rn == gn && std::equal(rd, rd + p_c, gd);
  for (int i = 0; ok && i < rn; ++i) ok = ri[i] == gi[i] && rv[i] == gv[i];
  std::printf("%-28s splits=%d p_c=%d n=%d %s\n", name, splits, p_c, rn, ok ? "identical" : "DIFFERENT");
  delete[] ri;
  delete[] rv;
  delete[] rd;
  delete[] gi;
  delete[] gv;
  delete[] gd;
  delete A;
  return ok;
}

int main() {
  if (cbx_device_count() < 1) {
    std::printf("no GPU visible\n");
    return 77;
  }
  MPI_Init(nullptr, nullptr);
  cbx_dcsc G;
  if (cbx_gen_rmat(12, 8, 0.57, 0.19, 0.19, 0xCB1A5 ^ 12, 0, &G) != CBX_OK) return 2;
  bool ok = true;
  ok &= check<PlusTimesSRing<int64_t, int64_t>, int64_t>(G, "PlusTimes<int64>");
  ok &= check<MinPlusSRing<int64_t, int64_t>, int64_t>(G, "MinPlus<int64>");
  ok &= check<SelectMaxSRing<int64_t, int64_t>, int64_t>(G, "SelectMax<int64>");
  ok &= check<PlusTimesSRing<bool, bool>, bool>(G, "PlusTimes<bool> (OR/AND)");
  ok &= check_spmxspv<SelectMaxSRing<bool, int64_t>, bool, int64_t>(G, "SpMXSpV SelectMax<bool,int64>");
  ok &= check_spmxspv<PlusTimesSRing<int64_t, int64_t>, int64_t, int64_t>(G, "SpMXSpV PlusTimes<int64>");
  ok &= check_spmxspv<MinPlusSRing<int64_t, int64_t>, int64_t, int64_t>(G, "SpMXSpV MinPlus<int64>");
  ok &= check_gespmv_threaded<SelectMaxSRing<bool, int64_t>, int64_t>(G, 32, 4, "gespmv_threaded SelectMax");
  ok &= check_gespmv_threaded<SelectMaxSRing<bool, int64_t>, int64_t>(G, 7, 3, "gespmv_threaded SelectMax");
  cbx_dcsc_free(&G);
  MPI_Finalize();
  std::printf(ok ? "shim ok\n" : "shim MISMATCH\n");
  return ok ? 0 : 1;
}
