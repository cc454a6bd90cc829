// cbx_combblas.hpp -- drop-in GPU counterparts of CombBLAS's node-local
// SpGEMM and SpMSpV calls, header-only, over the C ABI in cbx.h.
//
// Include after "CombBLAS/CombBLAS.h" and link libcbx.so.  The functions
// keep the reference templates' parameters and return types:
//
//   LocalHybridSpGEMM_HIP<SR,NTO>(A, B, clearA, clearB, aux)
//       replaces LocalHybridSpGEMM<SR,NTO> (include/CombBLAS/mtSpGEMM.h:213-217),
//       returning a new SpTuples<IT,NTO>, column-major, rows sorted.
//   LocalSpGEMMHash_HIP<SR,NTO>(A, B, clearA, clearB, sort)
//       replaces LocalSpGEMMHash<SR,NTO> (mtSpGEMM.h:463-467); sort=false
//       leaves rows of a column in no particular order.
//   LocalHybridSpGEMM_HIP_DCSC<SR,NTO>(A, B, clearA, clearB)
//       the same product returned as a new SpDCCols<IT,NTO> (what the SUMMA
//       drivers build from the tuples, ParFriends.h:1548-1553), without the
//       tuple round trip.
//   SpMXSpV_HIP<SR>(Adcsc, mA, indx, numx, veclen, indy, numy, SPA)
//       replaces SpMXSpV<SR,IT,NUM,IVT,OVT> (include/CombBLAS/SpImpl.h:49-56):
//       y = A x for a sparse x, appended to indy/numy with rows ascending.
//   generic_gespmv_threaded_HIP<SR>(A, indx, numx, nnzx, sendindbuf,
//                                   sendnumbuf, sdispls, p_c, SPA)
//       replaces generic_gespmv_threaded<SR> (Friends.h:181-320): y = A x,
//       cut by recipient (row r goes to min(r / (m / p_c), p_c - 1) with the
//       recipient's local index), in new[]'d sendindbuf / sendnumbuf /
//       sdispls that the caller deletes, returning the entry count.  A may be
//       row-split (SpDCCols::RowSplit) or not; the result does not depend on
//       the splits.
//
// Supported semirings (SpGEMM): PlusTimesSRing<T,T>, MinPlusSRing<T,T>,
// SelectMaxSRing<T,T> with T in {int64_t, double, float} and
// PlusTimesSRing<bool,bool>; NT1 = NT2 = NTO = T.  (SpMSpV): the same three
// over NUM = IVT = OVT = T, and over a boolean matrix NUM = bool with
// IVT = OVT = T -- SelectMaxSRing<bool,int64_t> is the BFS semiring
// (Semirings.h:193-210); a boolean matrix contributes x_k itself.  Any other
// combination is a compile-time error (there is no silent CPU fallback).  Failures throw
// std::runtime_error with cbx_last_error()'s text (the reference itself has
// no error path; a failed GPU call must not be mistaken for a result).
//
// `aux` is accepted for signature compatibility and ignored (the device
// builds its own column map).  Operands are copied to the device per call;
// callers that multiply the same block repeatedly should hold device copies
// and call cbx_spgemm directly.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "cbx.h"

namespace cbx_combblas {

template <class T>
struct dtype_of;
template <>
struct dtype_of<int64_t> { static constexpr cbx_dtype value = CBX_I64; };
template <>
struct dtype_of<double> { static constexpr cbx_dtype value = CBX_F64; };
template <>
struct dtype_of<float> { static constexpr cbx_dtype value = CBX_F32; };
template <>
struct dtype_of<bool> { static constexpr cbx_dtype value = CBX_BOOL; };

// Maps a CombBLAS semiring type to the ABI's enum (Semirings.h:165-255).
template <class SR>
struct semiring_of { static constexpr bool supported = false; };
template <class T>
struct semiring_of<combblas::PlusTimesSRing<T, T>> {
  static constexpr bool supported = true;
  static constexpr cbx_semiring value = CBX_SR_PLUS_TIMES;
  using type = T;
};
template <class T>
struct semiring_of<combblas::MinPlusSRing<T, T>> {
  static constexpr bool supported = !std::is_same<T, bool>::value;
  static constexpr cbx_semiring value = CBX_SR_MIN_PLUS;
  using type = T;
};
template <class T>
struct semiring_of<combblas::SelectMaxSRing<T, T>> {
  static constexpr bool supported = !std::is_same<T, bool>::value;
  static constexpr cbx_semiring value = CBX_SR_SELECT_MAX;
  using type = T;
};

inline void check(cbx_status st, const char* what) {
  if (st != CBX_OK) throw std::runtime_error(std::string(what) + ": " + cbx_last_error());
}

// Host view of a SpDCCols as an int64 cbx_dcsc; copies only when the index
// type is not int64_t.
template <class IT, class NT>
struct HostView {
  cbx_dcsc d{};
  std::vector<int64_t> cp, jc, ir;
  explicit HostView(const combblas::SpDCCols<IT, NT>& X) {
    d.m = X.getnrow();
    d.n = X.getncol();
    d.nnz = X.getnnz();
    d.dtype = dtype_of<NT>::value;
    d.location = CBX_HOST;
    static int64_t zero = 0;
    combblas::Dcsc<IT, NT>* dc = X.GetDCSC();
    if (d.nnz == 0 || dc == nullptr) {
      d.nnz = 0;
      d.cp = &zero;
      return;
    }
    d.nzc = dc->nzc;
    d.val = dc->numx;
    if (std::is_same<IT, int64_t>::value) {
      d.cp = reinterpret_cast<int64_t*>(dc->cp);
      d.jc = reinterpret_cast<int64_t*>(dc->jc);
      d.ir = reinterpret_cast<int64_t*>(dc->ir);
    } else {
      cp.assign(dc->cp, dc->cp + d.nzc + 1);
      jc.assign(dc->jc, dc->jc + d.nzc);
      ir.assign(dc->ir, dc->ir + d.nnz);
      d.cp = cp.data();
      d.jc = jc.data();
      d.ir = ir.data();
    }
  }
};

struct DevGuard {
  cbx_dcsc d{};
  ~DevGuard() { cbx_dcsc_free(&d); }
};

// Runs the product on the device and leaves the result in host DCSC form.
template <class SR, class NTO, class IT, class NT1, class NT2>
void product_to_host(const combblas::SpDCCols<IT, NT1>& A, const combblas::SpDCCols<IT, NT2>& B,
                     cbx_spgemm_mode mode, DevGuard& host_out) {
  static_assert(semiring_of<SR>::supported, "cbx: semiring not provided on the GPU");
  static_assert(std::is_same<NT1, NT2>::value && std::is_same<NT1, NTO>::value &&
                    std::is_same<typename semiring_of<SR>::type, NTO>::value,
                "cbx: operands and result must share the semiring's value type");
  HostView<IT, NT1> ha(A);
  HostView<IT, NT2> hb(B);
  DevGuard da, db, dc;
  check(cbx_dcsc_to_device(&ha.d, &da.d), "cbx_dcsc_to_device(A)");
  check(cbx_dcsc_to_device(&hb.d, &db.d), "cbx_dcsc_to_device(B)");
  check(cbx_spgemm(&da.d, &db.d, semiring_of<SR>::value, mode, &dc.d, nullptr, nullptr), "cbx_spgemm");
  check(cbx_dcsc_to_host(&dc.d, &host_out.d), "cbx_dcsc_to_host");
}

template <class IT, class NTO>
combblas::SpTuples<IT, NTO>* to_tuples(const cbx_dcsc& c) {
  if (c.nnz == 0) return new combblas::SpTuples<IT, NTO>(0, (IT)c.m, (IT)c.n);
  auto* t = new std::tuple<IT, IT, NTO>[c.nnz];
  const NTO* v = static_cast<const NTO*>(c.val);
  for (int64_t col = 0; col < c.nzc; ++col)
    for (int64_t e = c.cp[col]; e < c.cp[col + 1]; ++e) t[e] = std::make_tuple((IT)c.ir[e], (IT)c.jc[col], v[e]);
  // new[] storage, isOpNew = false (SpTuples.cpp:150-160 frees with delete[])
  return new combblas::SpTuples<IT, NTO>(c.nnz, (IT)c.m, (IT)c.n, t, true, false);
}

// SpMSpV: the ABI semiring and vector dtype of SR over a NUM matrix.
template <class SR, class NUM, class IVT, class OVT>
struct spmspv_of { static constexpr bool supported = false; };
template <template <class, class> class SRT, class T>
struct spmspv_sr {
  static constexpr bool known = std::is_same<SRT<T, T>, combblas::PlusTimesSRing<T, T>>::value ||
                                std::is_same<SRT<T, T>, combblas::MinPlusSRing<T, T>>::value ||
                                std::is_same<SRT<T, T>, combblas::SelectMaxSRing<T, T>>::value;
  static constexpr cbx_semiring value = std::is_same<SRT<T, T>, combblas::PlusTimesSRing<T, T>>::value
                                            ? CBX_SR_PLUS_TIMES
                                            : (std::is_same<SRT<T, T>, combblas::MinPlusSRing<T, T>>::value
                                                   ? CBX_SR_MIN_PLUS
                                                   : CBX_SR_SELECT_MAX);
};
template <template <class, class> class SRT, class NUM, class T>
struct spmspv_of<SRT<NUM, T>, NUM, T, T> {
  static constexpr bool vec_ok = std::is_same<T, int64_t>::value || std::is_same<T, double>::value ||
                                 std::is_same<T, float>::value;
  static constexpr bool supported =
      vec_ok && spmspv_sr<SRT, T>::known && (std::is_same<NUM, T>::value || std::is_same<NUM, bool>::value);
  static constexpr cbx_semiring value = spmspv_sr<SRT, T>::value;
  static constexpr cbx_dtype vdtype = dtype_of<T>::value;
};

// A host Dcsc (with mA rows and ncols columns) as an int64 cbx_dcsc.
template <class IT, class NUM>
struct DcscView {
  cbx_dcsc d{};
  std::vector<int64_t> cp, jc, ir;
  std::vector<uint8_t> bval;
  DcscView(const combblas::Dcsc<IT, NUM>& X, int64_t m, int64_t ncols) {
    d.m = m;
    d.n = ncols;
    d.nnz = X.nz;
    d.nzc = X.nzc;
    d.dtype = dtype_of<NUM>::value;
    d.location = CBX_HOST;
    cp.assign(X.cp, X.cp + X.nzc + 1);
    jc.assign(X.jc, X.jc + X.nzc);
    ir.assign(X.ir, X.ir + X.nz);
    d.cp = cp.data();
    d.jc = jc.data();
    d.ir = ir.data();
    if (std::is_same<NUM, bool>::value) {
      bval.resize((size_t)X.nz);
      for (int64_t i = 0; i < X.nz; ++i) bval[i] = X.numx[i] ? 1 : 0;
      d.val = bval.data();
    } else {
      d.val = (void*)X.numx;
    }
  }
};

// The whole matrix of a (possibly row-split) SpDCCols as one host DCSC: the
// splits hold consecutive row ranges of perpiece = m / splits rows with local
// row indices (SpDCCols::RowSplit), so appending each column's entries split
// by split keeps its rows ascending.
template <class IT, class NUM>
struct WholeView {
  using Stored = typename std::conditional<std::is_same<NUM, bool>::value, uint8_t, NUM>::type;
  cbx_dcsc d{};
  std::vector<int64_t> cp, jc, ir;
  std::vector<Stored> val;
  explicit WholeView(const combblas::SpDCCols<IT, NUM>& X) {
    d.m = X.getnrow();
    d.n = X.getncol();
    d.dtype = dtype_of<NUM>::value;
    d.location = CBX_HOST;
    const int splits = X.getnsplit();
    std::vector<const combblas::Dcsc<IT, NUM>*> parts;
    if (splits > 0)
      for (int i = 0; i < splits; ++i) parts.push_back(X.GetInternal(i));
    else
      parts.push_back(X.GetDCSC());
    const int64_t perpiece = splits > 0 ? d.m / splits : 0;
    std::vector<int64_t> cnt((size_t)d.n + 1, 0);
    for (auto* P : parts)
      if (P)
        for (int64_t q = 0; q < P->nzc; ++q) cnt[P->jc[q] + 1] += P->cp[q + 1] - P->cp[q];
    for (int64_t j = 0; j < d.n; ++j) cnt[j + 1] += cnt[j];
    const int64_t nnz = cnt[d.n];
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    ir.resize((size_t)nnz);
    val.resize((size_t)nnz);
    for (size_t k = 0; k < parts.size(); ++k) {
      auto* P = parts[k];
      if (!P) continue;
      const int64_t off = (int64_t)k * perpiece;
      for (int64_t q = 0; q < P->nzc; ++q)
        for (int64_t e = P->cp[q]; e < P->cp[q + 1]; ++e) {
          const int64_t at = fill[P->jc[q]]++;
          ir[at] = P->ir[e] + off;
          val[at] = (Stored)P->numx[e];
        }
    }
    cp.push_back(0);
    for (int64_t j = 0; j < d.n; ++j)
      if (cnt[j + 1] > cnt[j]) {
        jc.push_back(j);
        cp.push_back(cnt[j + 1]);
      }
    d.nnz = nnz;
    d.nzc = (int64_t)jc.size();
    d.cp = cp.data();
    d.jc = jc.data();
    d.ir = ir.data();
    d.val = val.data();
  }
};

// y = A x on the device for a host matrix view and host vector; returns the
// entries in host vectors, rows ascending.
template <class SR, class NUM, class IVT, class OVT>
void spmspv_to_host(const cbx_dcsc& hA, const int32_t* indx, const IVT* numx, int32_t nx, std::vector<int32_t>& yi,
                    std::vector<OVT>& yv) {
  using S = spmspv_of<SR, NUM, IVT, OVT>;
  static_assert(S::supported, "cbx: SpMSpV semiring / value types not provided on the GPU");
  yi.clear();
  yv.clear();
  if (hA.nnz == 0 || nx == 0) return;
  DevGuard dA;
  check(cbx_dcsc_to_device(&hA, &dA.d), "cbx_dcsc_to_device(A)");
  struct Buf {
    void* p = nullptr;
    ~Buf() { cbx_free(p); }
  } xi, xv, oi, ov;
  const size_t m = (size_t)std::max<int64_t>(hA.m, 1);
  check(cbx_malloc(&xi.p, sizeof(int32_t) * (size_t)nx), "cbx_malloc");
  check(cbx_malloc(&xv.p, sizeof(IVT) * (size_t)nx), "cbx_malloc");
  check(cbx_malloc(&oi.p, sizeof(int32_t) * m), "cbx_malloc");
  check(cbx_malloc(&ov.p, sizeof(OVT) * m), "cbx_malloc");
  check(cbx_memcpy(xi.p, indx, sizeof(int32_t) * (size_t)nx, 0), "cbx_memcpy");
  check(cbx_memcpy(xv.p, numx, sizeof(IVT) * (size_t)nx, 0), "cbx_memcpy");
  int32_t ny = 0;
  check(cbx_spmspv(&dA.d, (const int32_t*)xi.p, xv.p, nx, S::value, S::vdtype, (int32_t*)oi.p, ov.p, &ny, nullptr),
        "cbx_spmspv");
  yi.resize((size_t)ny);
  yv.resize((size_t)ny);
  if (ny) {
    check(cbx_memcpy(yi.data(), oi.p, sizeof(int32_t) * (size_t)ny, 1), "cbx_memcpy");
    check(cbx_memcpy(yv.data(), ov.p, sizeof(OVT) * (size_t)ny, 1), "cbx_memcpy");
  }
}

}  // namespace cbx_combblas

namespace combblas {

template <typename SR, typename NTO, typename IT, typename NT1, typename NT2>
SpTuples<IT, NTO>* LocalHybridSpGEMM_HIP(const SpDCCols<IT, NT1>& A, const SpDCCols<IT, NT2>& B, bool clearA,
                                         bool clearB, IT* aux = nullptr) {
  (void)aux;
  SpTuples<IT, NTO>* out;
  if (A.isZero() || B.isZero()) {  // mtSpGEMM.h:224-227
    out = new SpTuples<IT, NTO>(0, A.getnrow(), B.getncol());
  } else {
    cbx_combblas::DevGuard h;
    cbx_combblas::product_to_host<SR, NTO>(A, B, CBX_SPGEMM_TWO_PHASE, h);
    out = cbx_combblas::to_tuples<IT, NTO>(h.d);
  }
  if (clearA) delete const_cast<SpDCCols<IT, NT1>*>(&A);
  if (clearB) delete const_cast<SpDCCols<IT, NT2>*>(&B);
  return out;
}

template <typename SR, typename NTO, typename IT, typename NT1, typename NT2>
SpTuples<IT, NTO>* LocalSpGEMMHash_HIP(const SpDCCols<IT, NT1>& A, const SpDCCols<IT, NT2>& B, bool clearA,
                                       bool clearB, bool sort = true) {
  SpTuples<IT, NTO>* out;
  if (A.isZero() || B.isZero()) {
    out = new SpTuples<IT, NTO>(0, A.getnrow(), B.getncol());
  } else {
    cbx_combblas::DevGuard h;
    cbx_combblas::product_to_host<SR, NTO>(A, B, sort ? CBX_SPGEMM_TWO_PHASE : CBX_SPGEMM_UNSORTED, h);
    out = cbx_combblas::to_tuples<IT, NTO>(h.d);
  }
  if (clearA) delete const_cast<SpDCCols<IT, NT1>*>(&A);
  if (clearB) delete const_cast<SpDCCols<IT, NT2>*>(&B);
  return out;
}

// The product as SpDCCols, built with the sized constructor so the Dcsc owns
// new[] arrays (~Dcsc deletes them unconditionally, dcsc.cpp:1426-1439).
template <typename SR, typename NTO, typename IT, typename NT1, typename NT2>
SpDCCols<IT, NTO>* LocalHybridSpGEMM_HIP_DCSC(const SpDCCols<IT, NT1>& A, const SpDCCols<IT, NT2>& B, bool clearA,
                                              bool clearB) {
  SpDCCols<IT, NTO>* out;
  if (A.isZero() || B.isZero()) {
    out = new SpDCCols<IT, NTO>(0, A.getnrow(), B.getncol(), 0);
  } else {
    cbx_combblas::DevGuard h;
    cbx_combblas::product_to_host<SR, NTO>(A, B, CBX_SPGEMM_TWO_PHASE, h);
    const cbx_dcsc& c = h.d;
    out = new SpDCCols<IT, NTO>((IT)c.nnz, (IT)c.m, (IT)c.n, (IT)c.nzc);
    if (c.nnz > 0) {
      Dcsc<IT, NTO>* d = out->GetDCSC();
      for (int64_t i = 0; i <= c.nzc; ++i) d->cp[i] = (IT)c.cp[i];
      for (int64_t i = 0; i < c.nzc; ++i) d->jc[i] = (IT)c.jc[i];
      for (int64_t i = 0; i < c.nnz; ++i) d->ir[i] = (IT)c.ir[i];
      std::memcpy(d->numx, c.val, sizeof(NTO) * (size_t)c.nnz);
    }
  }
  if (clearA) delete const_cast<SpDCCols<IT, NT1>*>(&A);
  if (clearB) delete const_cast<SpDCCols<IT, NT2>*>(&B);
  return out;
}

// y = A x for a sparse x, the counterpart of SpMXSpV overload #1
// (SpImpl.h:49-56).  Like the reference it appends to indy / numy, combining
// the first new entry with indy.back() under SR::add when their rows agree.
// The matrix's column count is taken as one past the largest of its stored
// columns and x's indices.  SPA is accepted for signature compatibility.
template <class SR, class IT, class NUM, class IVT, class OVT>
void SpMXSpV_HIP(const Dcsc<IT, NUM>& Adcsc, int32_t mA, const int32_t* indx, const IVT* numx, int32_t veclen,
                 std::vector<int32_t>& indy, std::vector<OVT>& numy, PreAllocatedSPA<OVT>& SPA) {
  (void)SPA;
  if (veclen <= 0 || Adcsc.nz == 0) return;
  const int64_t ncols = std::max<int64_t>(Adcsc.nzc ? (int64_t)Adcsc.jc[Adcsc.nzc - 1] + 1 : 0,
                                          (int64_t)indx[veclen - 1] + 1);
  cbx_combblas::DcscView<IT, NUM> hv(Adcsc, mA, ncols);
  std::vector<int32_t> yi;
  std::vector<OVT> yv;
  cbx_combblas::spmspv_to_host<SR, NUM, IVT, OVT>(hv.d, indx, numx, veclen, yi, yv);
  size_t first = 0;
  if (!yi.empty() && !indy.empty() && indy.back() == yi[0]) {
    numy.back() = SR::add(numy.back(), yv[0]);
    first = 1;
  }
  indy.insert(indy.end(), yi.begin() + first, yi.end());
  numy.insert(numy.end(), yv.begin() + first, yv.end());
}

// generic_gespmv_threaded (Friends.h:181-320) on the device: y = A x, then
// the entries cut by recipient into new[]'d send buffers.  Returns the entry
// count; with an empty A or x, sendindbuf / sendnumbuf are null and sdispls
// all zero, as in the reference.
template <typename SR, typename IU, typename NUM, typename DER, typename IVT, typename OVT>
int generic_gespmv_threaded_HIP(const SpMat<IU, NUM, DER>& A, const int32_t* indx, const IVT* numx, int32_t nnzx,
                                int32_t*& sendindbuf, OVT*& sendnumbuf, int*& sdispls, int p_c,
                                PreAllocatedSPA<OVT>& SPA) {
  (void)SPA;
  sdispls = new int[p_c]();
  if (A.getnnz() == 0 || nnzx == 0) {
    sendindbuf = nullptr;
    sendnumbuf = nullptr;
    return 0;
  }
  const DER& M = static_cast<const DER&>(A);
  cbx_combblas::WholeView<IU, NUM> hv(M);
  std::vector<int32_t> yi;
  std::vector<OVT> yv;
  cbx_combblas::spmspv_to_host<SR, NUM, IVT, OVT>(hv.d, indx, numx, nnzx, yi, yv);
  const int32_t nlocrows = static_cast<int32_t>(A.getnrow());
  const int32_t perproc = nlocrows / p_c;
  if (perproc <= 0) {
    delete[] sdispls;
    sdispls = nullptr;
    throw std::invalid_argument("generic_gespmv_threaded_HIP: fewer rows than recipients");
  }
  const int32_t last_rec = p_c - 1;
  const int total = (int)yi.size();
  sendindbuf = new int32_t[total];
  sendnumbuf = new OVT[total];
  int rec = 0;  // sdispls[k] = entries bound for recipients below k
  for (int i = 0; i < total; ++i) {
    const int32_t cur = std::min(yi[i] / perproc, last_rec);
    while (rec < cur) sdispls[++rec] = i;
    sendindbuf[i] = yi[i] - perproc * cur;
    sendnumbuf[i] = yv[i];
  }
  while (rec < last_rec) sdispls[++rec] = total;
  return total;
}

}  // namespace combblas
