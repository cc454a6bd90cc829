// Device-side semiring functors for gfx950.
//
// Each functor restates one CombBLAS semiring (include/CombBLAS/Semirings.h)
// for a single value type T shared by both operands and the result:
//   PlusTimes  Semirings.h:212-232  add a+b, multiply a*b; for bool the
//              promote_trait<bool,bool> arithmetic gives OR / AND.
//   MinPlus    Semirings.h:235-255  add min, multiply inf_plus (:40-47),
//              which saturates at numeric_limits<T>::max().
//   SelectMax  Semirings.h:165-190  add max, multiply a*b.
//
// Accumulation happens in LDS with atomics.  `Acc` is the LDS accumulator
// type and `ident()` a true identity of `add` for it, so an accumulator slot
// that starts at ident() and receives the products x1..xk holds
// x1 (+) ... (+) xk exactly as the reference's first-insert-then-add does:
//   PlusTimes integers 0; floating point -0.0 (so a lone -0.0 stays -0.0);
//   MinPlus   max();
//   SelectMax lowest() (the reference's id() of -1 is not an identity of max
//             for products below -1, so it is not used here);
//   bool      0 in a 32-bit word accumulated with OR.
// Whether an entry exists is tracked separately (bitmap or hash key), so
// explicit zeros survive like in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <limits>

namespace cbx {

enum SemiringKind : int { kPlusTimes = 0, kMinPlus = 1, kSelectMax = 2 };

template <int K, class T>
struct Sem;

// -------------------------------------------------------------- helpers
template <class A>
__device__ __forceinline__ void lds_add(A* p, A v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class A>
__device__ __forceinline__ void lds_min(A* p, A v) {
  __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class A>
__device__ __forceinline__ void lds_max(A* p, A v) {
  __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Floating min/max through a compare-and-swap loop on the bit pattern, with
// std::min / std::max semantics: the stored value is replaced only when the
// new value is strictly smaller (larger).
template <class F, class U>
__device__ __forceinline__ void lds_cas_minmax(F* p, F v, bool is_min) {
  U* up = reinterpret_cast<U*>(p);
  U old = __hip_atomic_load(up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (true) {
    F cur = __builtin_bit_cast(F, old);
    bool better = is_min ? (v < cur) : (cur < v);
    if (!better) return;
    U want = __builtin_bit_cast(U, v);
    if (__hip_atomic_compare_exchange_strong(up, &old, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return;
  }
}

// -------------------------------------------------------------- PlusTimes
template <>
struct Sem<kPlusTimes, int64_t> {
  using S = int64_t;
  using Acc = unsigned long long;  // two's-complement wrap like the reference
  static __device__ __forceinline__ Acc ident() { return 0ull; }
  static __device__ __forceinline__ Acc mul(S a, S b) { return (Acc)a * (Acc)b; }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_add(p, v); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a + b; }
  static __device__ __forceinline__ S out(Acc a) { return (S)a; }
};
template <>
struct Sem<kPlusTimes, double> {
  using S = double;
  using Acc = double;
  static __device__ __forceinline__ Acc ident() { return -0.0; }
  static __device__ __forceinline__ Acc mul(S a, S b) { return a * b; }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_add(p, v); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a + b; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};
template <>
struct Sem<kPlusTimes, float> {
  using S = float;
  using Acc = float;
  static __device__ __forceinline__ Acc ident() { return -0.0f; }
  static __device__ __forceinline__ Acc mul(S a, S b) { return a * b; }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_add(p, v); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a + b; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};
template <>
struct Sem<kPlusTimes, uint8_t> {  // bool: OR / AND
  using S = uint8_t;
  using Acc = uint32_t;
  static __device__ __forceinline__ Acc ident() { return 0u; }
  static __device__ __forceinline__ Acc mul(S a, S b) { return (Acc)((a != 0) & (b != 0)); }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a | b; }
  static __device__ __forceinline__ S out(Acc a) { return (S)(a != 0); }
};

// -------------------------------------------------------------- MinPlus
template <class T>
__device__ __forceinline__ T inf_plus(T a, T b) {
  const T inf = std::numeric_limits<T>::max();
  if (a == inf || b == inf) return inf;
  return a + b;
}
template <>
struct Sem<kMinPlus, int64_t> {
  using S = int64_t;
  using Acc = int64_t;
  static __device__ __forceinline__ Acc ident() { return std::numeric_limits<int64_t>::max(); }
  static __device__ __forceinline__ Acc mul(S a, S b) {
    const S inf = std::numeric_limits<S>::max();
    if (a == inf || b == inf) return inf;
    return (S)((unsigned long long)a + (unsigned long long)b);
  }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_min(p, v); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return b < a ? b : a; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};
template <>
struct Sem<kMinPlus, double> {
  using S = double;
  using Acc = double;
  static __device__ __forceinline__ Acc ident() { return std::numeric_limits<double>::max(); }
  static __device__ __forceinline__ Acc mul(S a, S b) { return inf_plus(a, b); }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_cas_minmax<double, unsigned long long>(p, v, true); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return b < a ? b : a; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};
template <>
struct Sem<kMinPlus, float> {
  using S = float;
  using Acc = float;
  static __device__ __forceinline__ Acc ident() { return std::numeric_limits<float>::max(); }
  static __device__ __forceinline__ Acc mul(S a, S b) { return inf_plus(a, b); }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_cas_minmax<float, unsigned int>(p, v, true); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return b < a ? b : a; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};

// -------------------------------------------------------------- SelectMax
template <>
struct Sem<kSelectMax, int64_t> {
  using S = int64_t;
  using Acc = int64_t;
  static __device__ __forceinline__ Acc ident() { return std::numeric_limits<int64_t>::lowest(); }
  static __device__ __forceinline__ Acc mul(S a, S b) { return (S)((unsigned long long)a * (unsigned long long)b); }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_max(p, v); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a < b ? b : a; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};
template <>
struct Sem<kSelectMax, double> {
  using S = double;
  using Acc = double;
  static __device__ __forceinline__ Acc ident() { return -std::numeric_limits<double>::infinity(); }
  static __device__ __forceinline__ Acc mul(S a, S b) { return a * b; }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_cas_minmax<double, unsigned long long>(p, v, false); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a < b ? b : a; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};
template <>
struct Sem<kSelectMax, float> {
  using S = float;
  using Acc = float;
  static __device__ __forceinline__ Acc ident() { return -std::numeric_limits<float>::infinity(); }
  static __device__ __forceinline__ Acc mul(S a, S b) { return a * b; }
  static __device__ __forceinline__ void acc(Acc* p, Acc v) { lds_cas_minmax<float, unsigned int>(p, v, false); }
  static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a < b ? b : a; }
  static __device__ __forceinline__ S out(Acc a) { return a; }
};

}  // namespace cbx
