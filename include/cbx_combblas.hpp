// cbx_combblas.hpp -- drop-in GPU counterparts of CombBLAS's node-local
// SpGEMM calls, header-only, over the C ABI in cbx.h.
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
//
// Supported semirings: PlusTimesSRing<T,T>, MinPlusSRing<T,T>,
// SelectMaxSRing<T,T> with T in {int64_t, double, float} and
// PlusTimesSRing<bool,bool>; NT1 = NT2 = NTO = T.  Any other combination is a
// compile-time error (there is no silent CPU fallback).  Failures throw
// std::runtime_error with cbx_last_error()'s text (the reference itself has
// no error path; a failed GPU call must not be mistaken for a result).
//
// `aux` is accepted for signature compatibility and ignored (the device
// builds its own column map).  Operands are copied to the device per call;
// callers that multiply the same block repeatedly should hold device copies
// and call cbx_spgemm directly.
#pragma once

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

}  // namespace combblas
