// Binary operators of the device-wide primitives and their identities.
//
// The operator set and its numbering follow the reference's group/algorithm
// builtin op enum (include/hipSYCL/sycl/libkernel/sscp/builtins/builtin_config.hpp:27-37):
//   plus, multiply, min, max, bit_and, bit_or, bit_xor, logical_and, logical_or.
//
// Each functor is parameterised on the "work" type W the kernel computes in
// and the "storage" type T the caller's data has.  For 8/16-bit integers the
// work type is a 32-bit integer: plus and multiply wrap modulo 2^32 and the
// result is truncated to T at the end, which equals wrapping in T because
// 2^bits(T) divides 2^32.  min/max never leave T's range.
//
// Identities:
//   plus      -> 0 for integers, -0.0 for floating point.  -0.0 is the exact
//                additive identity of IEEE-754 (x + -0.0 == x for every x,
//                including -0.0); the reference's T{} == +0.0 turns a lone
//                -0.0 into +0.0 (numeric.hpp:44).
//   multiply  -> 1
//   min       -> numeric_limits<T>::max(), +inf for floating point
//   max       -> numeric_limits<T>::lowest(), -inf for floating point.  The
//                reference's algorithms layer uses numeric_limits<T>::min()
//                (numeric.hpp:53), which for float is the smallest positive
//                normal and is not an identity of max; sycl::known_identity
//                uses lowest() (functional.hpp:234-244), which we follow.
//   bit_and   -> all ones;  bit_or, bit_xor -> 0
//   logical_and -> 1;       logical_or -> 0
#pragma once

#include <stdint.h>
#include <limits>
#include <type_traits>

#ifndef ACPP_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define ACPP_HD __host__ __device__ __forceinline__
#else
#define ACPP_HD inline
#endif
#endif

namespace acpp {
namespace ops {

// Floating-point element types: the standard ones plus _Float16 (the
// half-precision storage type of the f16 builtins), for which the standard
// library's is_floating_point and numeric_limits say nothing.
template <class W>
inline constexpr bool is_floating_v = std::is_floating_point_v<W>
#if defined(__clang__)
                                      || std::is_same_v<std::remove_cv_t<W>, _Float16>
#endif
    ;

template <class W>
ACPP_HD W infinity() {
#if defined(__clang__)
  if constexpr (std::is_same_v<W, _Float16>)
    return static_cast<_Float16>(__builtin_huge_valf());
  else
#endif
    return std::numeric_limits<W>::infinity();
}

enum algorithm_op : int32_t {
  op_plus = 0,
  op_multiply = 1,
  op_min = 2,
  op_max = 3,
  op_bit_and = 4,
  op_bit_or = 5,
  op_bit_xor = 6,
  op_logical_and = 7,
  op_logical_or = 8,
};

template <class W>
ACPP_HD W wrap_add(W a, W b) {
  if constexpr (std::is_integral_v<W>) {
    using U = std::make_unsigned_t<W>;
    return static_cast<W>(static_cast<U>(a) + static_cast<U>(b));
  } else {
    return a + b;
  }
}

template <class W>
ACPP_HD W wrap_mul(W a, W b) {
  if constexpr (std::is_integral_v<W>) {
    using U = std::make_unsigned_t<W>;
    return static_cast<W>(static_cast<U>(a) * static_cast<U>(b));
  } else {
    return a * b;
  }
}

template <class W, class T = W>
struct plus {
  static constexpr int id = op_plus;
  ACPP_HD static W apply(W a, W b) { return wrap_add(a, b); }
  ACPP_HD static W identity() {
    if constexpr (is_floating_v<W>)
      return W(-0.0);
    else
      return W(0);
  }
};

template <class W, class T = W>
struct multiply {
  static constexpr int id = op_multiply;
  ACPP_HD static W apply(W a, W b) { return wrap_mul(a, b); }
  ACPP_HD static W identity() { return W(1); }
};

// sycl::minimum: (x < y) ? x : y   (functional.hpp)
template <class W, class T = W>
struct minimum {
  static constexpr int id = op_min;
  ACPP_HD static W apply(W a, W b) { return (a < b) ? a : b; }
  ACPP_HD static W identity() {
    if constexpr (is_floating_v<W>)
      return infinity<W>();
    else
      return static_cast<W>(std::numeric_limits<T>::max());
  }
};

// sycl::maximum: (x > y) ? x : y   (functional.hpp)
template <class W, class T = W>
struct maximum {
  static constexpr int id = op_max;
  ACPP_HD static W apply(W a, W b) { return (a > b) ? a : b; }
  ACPP_HD static W identity() {
    if constexpr (is_floating_v<W>)
      return -infinity<W>();
    else
      return static_cast<W>(std::numeric_limits<T>::lowest());
  }
};

template <class W, class T = W>
struct bit_and {
  static constexpr int id = op_bit_and;
  ACPP_HD static W apply(W a, W b) { return a & b; }
  ACPP_HD static W identity() { return static_cast<W>(~static_cast<W>(0)); }
};

template <class W, class T = W>
struct bit_or {
  static constexpr int id = op_bit_or;
  ACPP_HD static W apply(W a, W b) { return a | b; }
  ACPP_HD static W identity() { return W(0); }
};

template <class W, class T = W>
struct bit_xor {
  static constexpr int id = op_bit_xor;
  ACPP_HD static W apply(W a, W b) { return a ^ b; }
  ACPP_HD static W identity() { return W(0); }
};

template <class W, class T = W>
struct logical_and {
  static constexpr int id = op_logical_and;
  ACPP_HD static W apply(W a, W b) { return static_cast<W>(a && b); }
  ACPP_HD static W identity() { return W(1); }
};

template <class W, class T = W>
struct logical_or {
  static constexpr int id = op_logical_or;
  ACPP_HD static W apply(W a, W b) { return static_cast<W>(a || b); }
  ACPP_HD static W identity() { return W(0); }
};

}  // namespace ops
}  // namespace acpp
