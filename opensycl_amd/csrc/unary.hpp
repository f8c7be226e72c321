// Element transforms of transform_reduce (unary form) and of
// transform_inclusive_scan / transform_exclusive_scan.
//
// The reference takes the transform as an arbitrary callable compiled into
// the kernel (include/hipSYCL/algorithms/numeric.hpp:193-212, 299-327;
// scan/scan.hpp:85-132: the generator returns unary_op(*it)).  A C-ABI
// library cannot take device code, so the transform is one of a closed set
// of codes (acpp_unary_op):
//   identity, negate (-x), abs, square (x*x), twice (x+x), bit_not (~x,
//   integers), is_zero (x == 0 ? 1 : 0), is_nonzero (x != 0 ? 1 : 0).
// The result has the element type T: integer transforms wrap modulo
// 2^bits(T) (abs of the most negative value is itself), and the kernels then
// combine the transformed values exactly as they combine plain elements.
//
// The code is uniform across a launch.  A kernel applies it to a register
// tile or to one loaded row with one scalar branch (apply_tile, apply_row;
// apply_bytes for 16 packed 1-byte elements), so the transform costs one
// uniform switch per loaded batch, not per element.
#pragma once

#include "internal.hpp"

namespace acpp_impl {
namespace unary_detail {

enum : int {
  u_identity = 0,
  u_negate = 1,
  u_abs = 2,
  u_square = 3,
  u_twice = 4,
  u_bit_not = 5,
  u_is_zero = 6,
  u_is_nonzero = 7,
  u_count = 8,
};

// Transform x of element type T.
template <int C, class T>
__device__ __forceinline__ T apply(T x) {
  if constexpr (std::is_integral_v<T>) {
    // Arithmetic in an unsigned type at least 32 bits wide: wraps without
    // undefined behaviour and equals wrapping in T after the final cast.
    using U = std::conditional_t<(sizeof(T) > 4), uint64_t, uint32_t>;
    const U u = static_cast<U>(x);
    if constexpr (C == u_negate) return static_cast<T>(U(0) - u);
    else if constexpr (C == u_abs) return (std::is_signed_v<T> && x < T(0)) ? static_cast<T>(U(0) - u) : x;
    else if constexpr (C == u_square) return static_cast<T>(u * u);
    else if constexpr (C == u_twice) return static_cast<T>(u + u);
    else if constexpr (C == u_bit_not) return static_cast<T>(~u);
    else if constexpr (C == u_is_zero) return static_cast<T>(x == T(0) ? 1 : 0);
    else if constexpr (C == u_is_nonzero) return static_cast<T>(x != T(0) ? 1 : 0);
    else return x;
  } else {
    if constexpr (C == u_negate) return -x;
    else if constexpr (C == u_abs) {
      if constexpr (sizeof(T) == 4) return __builtin_fabsf(x);
      else return __builtin_fabs(x);
    } else if constexpr (C == u_square) {
#pragma clang fp contract(off)
      return x * x;
    } else if constexpr (C == u_twice) return x + x;
    else if constexpr (C == u_is_zero) return (x == T(0)) ? T(1) : T(0);
    else if constexpr (C == u_is_nonzero) return (x != T(0)) ? T(1) : T(0);
    else return x;  // bit_not is rejected on the host for floating point
  }
}

// Transform a value held in the work type W of element type T.
template <int C, class T, class W>
__device__ __forceinline__ W apply_work(W w) {
  return elem_traits<T>::to_work(apply<C, T>(static_cast<T>(w)));
}

template <int C, class T, class W, int R, int VEC>
__device__ __forceinline__ void each(W (&v)[R][VEC]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int k = 0; k < VEC; ++k) v[r][k] = apply_work<C, T>(v[r][k]);
  }
}

// Apply transform `code` to every value of a register tile.
template <class T, class W, int R, int VEC>
__device__ __forceinline__ void apply_tile(int code, W (&v)[R][VEC]) {
  switch (code) {
    case u_negate: each<u_negate, T>(v); break;
    case u_abs: each<u_abs, T>(v); break;
    case u_square: each<u_square, T>(v); break;
    case u_twice: each<u_twice, T>(v); break;
    case u_bit_not: each<u_bit_not, T>(v); break;
    case u_is_zero: each<u_is_zero, T>(v); break;
    case u_is_nonzero: each<u_is_nonzero, T>(v); break;
    default: break;
  }
}

// Apply transform `code` to one register vector.
template <class T, class W, int VEC>
__device__ __forceinline__ void apply_row(int code, W (&v)[VEC]) {
  W t[1][VEC];
#pragma unroll
  for (int k = 0; k < VEC; ++k) t[0][k] = v[k];
  apply_tile<T, W, 1, VEC>(code, t);
#pragma unroll
  for (int k = 0; k < VEC; ++k) v[k] = t[0][k];
}

// Apply transform `code` to one value (partial tiles, scalar loops).
template <class T, class W>
__device__ __forceinline__ W apply_one(int code, W w) {
  W v[1][1] = {{w}};
  apply_tile<T, W, 1, 1>(code, v);
  return v[0][0];
}

// Transform C of the four 1-byte elements (type T) packed in x, byte k of
// the result being apply<C, T> of byte k.  Plain 32-bit arithmetic whose
// carries and borrows never cross a byte (H: the top bit of every byte):
// negate computes (0x80 | 0) - (b & 0x7f) per byte and restores the top
// bit, is_zero / is_nonzero read the top bit of (b & 0x7f) + 0x7f | b, and
// square multiplies the even and the odd bytes as 16-bit halves.
template <int C, class T>
__device__ __forceinline__ uint32_t apply_bytes4(uint32_t x) {
  static_assert(sizeof(T) == 1 && std::is_integral_v<T>, "1-byte integer elements");
  typedef unsigned short half2 __attribute__((ext_vector_type(2)));
  constexpr uint32_t H = 0x80808080u;
  constexpr uint32_t L = 0x7f7f7f7fu;
  if constexpr (C == u_negate) {
    return (H - (x & L)) ^ (~x & H);
  } else if constexpr (C == u_abs) {
    if constexpr (std::is_signed_v<T>) {
      const uint32_t neg = (H - (x & L)) ^ (~x & H);
      const uint32_t m = ((x & H) >> 7) * 0xffu;  // 0xff in every negative byte
      return (neg & m) | (x & ~m);
    } else {
      return x;
    }
  } else if constexpr (C == u_square) {
    const half2 e = __builtin_bit_cast(half2, x & 0x00ff00ffu);
    const half2 o = __builtin_bit_cast(half2, (x >> 8) & 0x00ff00ffu);
    return (__builtin_bit_cast(uint32_t, e * e) & 0x00ff00ffu) |
           ((__builtin_bit_cast(uint32_t, o * o) & 0x00ff00ffu) << 8);
  } else if constexpr (C == u_twice) {
    return (x << 1) & 0xfefefefeu;
  } else if constexpr (C == u_bit_not) {
    return ~x;
  } else if constexpr (C == u_is_zero || C == u_is_nonzero) {
    const uint32_t y = ((x & L) + L) | x;  // top bit of a byte set: byte != 0
    return C == u_is_nonzero ? (y & H) >> 7 : (~y & H) >> 7;
  } else {
    return x;
  }
}

template <int C, class T>
__device__ __forceinline__ void each_bytes(uint32_t (&d)[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) d[k] = apply_bytes4<C, T>(d[k]);
}

// Apply transform `code` to 16 packed 1-byte elements (one uniform branch).
template <class T>
__device__ __forceinline__ void apply_bytes(int code, uint32_t (&d)[4]) {
  switch (code) {
    case u_negate: each_bytes<u_negate, T>(d); break;
    case u_abs: each_bytes<u_abs, T>(d); break;
    case u_square: each_bytes<u_square, T>(d); break;
    case u_twice: each_bytes<u_twice, T>(d); break;
    case u_bit_not: each_bytes<u_bit_not, T>(d); break;
    case u_is_zero: each_bytes<u_is_zero, T>(d); break;
    case u_is_nonzero: each_bytes<u_is_nonzero, T>(d); break;
    default: break;
  }
}

// Host-side validation: whether `code` is a transform of element type T.
template <class T>
inline bool valid_code(int code) {
  if (code < u_identity || code >= u_count) return false;
  if (code == u_bit_not && !std::is_integral_v<T>) return false;
  return true;
}

}  // namespace unary_detail
}  // namespace acpp_impl
