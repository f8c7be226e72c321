// Element transforms of the binary transform_reduce.
//
// The reference takes any pair of callables (reduce, transform) compiled
// into the kernel (include/hipSYCL/algorithms/numeric.hpp:170-191):
// *out = reduce(init, transform(x[i], y[i]) ...).  A C-ABI library cannot take
// device code, so the transform is one of a closed set of codes
// (acpp_binary_op), applied to the two elements in the element type T:
//   multiplies (x*y, the dot product), plus (x+y), minus (x-y), absolute
//   difference |x-y|, squared difference (x-y)^2, minimum (x < y ? x : y),
//   maximum (x > y ? x : y), and for integers bit_and, bit_or, bit_xor.
// Integer results wrap modulo 2^bits(T) (the absolute difference is
// computed as |x-y| of the wrapped difference, like std::abs of T(x-y));
// floating-point results are rounded once per operation, never fused.  The
// code is uniform across a launch, so a kernel applies it to a loaded
// vector with one scalar branch.
#pragma once

#include "internal.hpp"

namespace acpp_impl {
namespace binary_detail {

enum : int {
  b_multiplies = 0,
  b_plus = 1,
  b_minus = 2,
  b_abs_diff = 3,
  b_sq_diff = 4,
  b_min = 5,
  b_max = 6,
  b_bit_and = 7,
  b_bit_or = 8,
  b_bit_xor = 9,
  b_count = 10,
};

template <int C, class T>
__device__ __forceinline__ T apply(T x, T y) {
  if constexpr (std::is_integral_v<T>) {
    using U = std::conditional_t<(sizeof(T) > 4), uint64_t, uint32_t>;
    const U a = static_cast<U>(x), b = static_cast<U>(y);
    if constexpr (C == b_multiplies) return static_cast<T>(a * b);
    else if constexpr (C == b_plus) return static_cast<T>(a + b);
    else if constexpr (C == b_minus) return static_cast<T>(a - b);
    else if constexpr (C == b_abs_diff) {
      const T d = static_cast<T>(a - b);
      return (std::is_signed_v<T> && d < T(0)) ? static_cast<T>(U(0) - static_cast<U>(d)) : d;
    } else if constexpr (C == b_sq_diff) {
      const U d = a - b;
      return static_cast<T>(d * d);
    } else if constexpr (C == b_min) return (x < y) ? x : y;
    else if constexpr (C == b_max) return (x > y) ? x : y;
    else if constexpr (C == b_bit_and) return static_cast<T>(a & b);
    else if constexpr (C == b_bit_or) return static_cast<T>(a | b);
    else return static_cast<T>(a ^ b);
  } else {
#pragma clang fp contract(off)
    if constexpr (C == b_multiplies) return x * y;
    else if constexpr (C == b_plus) return x + y;
    else if constexpr (C == b_minus) return x - y;
    else if constexpr (C == b_abs_diff) {
      const T d = x - y;
      if constexpr (sizeof(T) == 4) return __builtin_fabsf(d);
      else return __builtin_fabs(d);
    } else if constexpr (C == b_sq_diff) {
      const T d = x - y;
      return d * d;
    } else if constexpr (C == b_min) return (x < y) ? x : y;  // sycl::minimum
    else return (x > y) ? x : y;  // b_max, sycl::maximum (bitwise codes are rejected on the host)
  }
}

template <int C, class T, class W, int VEC>
__device__ __forceinline__ void each(const W (&x)[VEC], const W (&y)[VEC], W (&out)[VEC]) {
#pragma unroll
  for (int k = 0; k < VEC; ++k)
    out[k] = elem_traits<T>::to_work(apply<C, T>(static_cast<T>(x[k]), static_cast<T>(y[k])));
}

// out[k] = transform `code` of (x[k], y[k]), one uniform branch per vector.
template <class T, class W, int VEC>
__device__ __forceinline__ void apply_row(int code, const W (&x)[VEC], const W (&y)[VEC], W (&out)[VEC]) {
  switch (code) {
    case b_plus: each<b_plus, T>(x, y, out); break;
    case b_minus: each<b_minus, T>(x, y, out); break;
    case b_abs_diff: each<b_abs_diff, T>(x, y, out); break;
    case b_sq_diff: each<b_sq_diff, T>(x, y, out); break;
    case b_min: each<b_min, T>(x, y, out); break;
    case b_max: each<b_max, T>(x, y, out); break;
    case b_bit_and:
      if constexpr (std::is_integral_v<T>) each<b_bit_and, T>(x, y, out);
      break;
    case b_bit_or:
      if constexpr (std::is_integral_v<T>) each<b_bit_or, T>(x, y, out);
      break;
    case b_bit_xor:
      if constexpr (std::is_integral_v<T>) each<b_bit_xor, T>(x, y, out);
      break;
    default: each<b_multiplies, T>(x, y, out); break;
  }
}

template <class T, class W>
__device__ __forceinline__ W apply_one(int code, W x, W y) {
  W a[1] = {x}, b[1] = {y}, r[1];
  apply_row<T, W, 1>(code, a, b, r);
  return r[0];
}

// Host-side validation: whether `code` is a binary transform of T.
template <class T>
inline bool valid_code(int code) {
  if (code < b_multiplies || code >= b_count) return false;
  if (code >= b_bit_and && !std::is_integral_v<T>) return false;
  return true;
}

}  // namespace binary_detail
}  // namespace acpp_impl
