// Operator functors shared by host code (the algorithms facade) and device
// code (group collectives): acpp::plus, multiplies, minimum, maximum,
// bit_and, bit_or, bit_xor, logical_and, logical_or.
//
// Same names and meaning as sycl::plus ... sycl::logical_or
// (include/hipSYCL/sycl/libkernel/functional.hpp).  Each functor also carries
// the operator code of the C-ABI (acpp_algorithm_op, numbered like
// __acpp_sscp_algorithm_op in
// include/hipSYCL/sycl/libkernel/sscp/builtins/builtin_config.hpp:27-37), and
// algorithm_op_of<Op> maps std:: functors onto the same codes.
#pragma once

#include <functional>
#include <type_traits>

#if defined(__HIP__) || defined(__HIPCC__)
#define ACPP_HOST_DEVICE __host__ __device__
#else
#define ACPP_HOST_DEVICE
#endif

namespace acpp {

template <class T = void>
struct plus {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return a + b; }
};
template <class T = void>
struct multiplies {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return a * b; }
};
template <class T = void>
struct minimum {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return (a < b) ? a : b; }
};
template <class T = void>
struct maximum {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return (a > b) ? a : b; }
};
template <class T = void>
struct bit_and {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return a & b; }
};
template <class T = void>
struct bit_or {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return a | b; }
};
template <class T = void>
struct bit_xor {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return a ^ b; }
};
template <class T = void>
struct logical_and {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return T(a && b); }
};
template <class T = void>
struct logical_or {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return T(a || b); }
};

// Transparent (T = void) forms, as std::plus<> is.
#define ACPP_TRANSPARENT(NAME, EXPR)                                              \
  template <>                                                                     \
  struct NAME<void> {                                                             \
    template <class A, class B>                                                   \
    ACPP_HOST_DEVICE auto operator()(const A &a, const B &b) const { return EXPR; } \
  };
ACPP_TRANSPARENT(plus, a + b)
ACPP_TRANSPARENT(multiplies, a * b)
ACPP_TRANSPARENT(minimum, (a < b) ? a : b)
ACPP_TRANSPARENT(maximum, (a > b) ? a : b)
ACPP_TRANSPARENT(bit_and, a & b)
ACPP_TRANSPARENT(bit_or, a | b)
ACPP_TRANSPARENT(bit_xor, a ^ b)
ACPP_TRANSPARENT(logical_and, a && b)
ACPP_TRANSPARENT(logical_or, a || b)
#undef ACPP_TRANSPARENT

// Operator code of a functor type; value == -1 for functors the library
// has no kernel for.
template <class Op>
struct algorithm_op_of {
  static constexpr int value = -1;
};
#define ACPP_OP_CODE(TMPL, CODE)                                                  \
  template <class T>                                                              \
  struct algorithm_op_of<TMPL<T>> {                                               \
    static constexpr int value = CODE;                                            \
  };
ACPP_OP_CODE(plus, 0)
ACPP_OP_CODE(multiplies, 1)
ACPP_OP_CODE(minimum, 2)
ACPP_OP_CODE(maximum, 3)
ACPP_OP_CODE(bit_and, 4)
ACPP_OP_CODE(bit_or, 5)
ACPP_OP_CODE(bit_xor, 6)
ACPP_OP_CODE(logical_and, 7)
ACPP_OP_CODE(logical_or, 8)
ACPP_OP_CODE(std::plus, 0)
ACPP_OP_CODE(std::multiplies, 1)
ACPP_OP_CODE(std::bit_and, 4)
ACPP_OP_CODE(std::bit_or, 5)
ACPP_OP_CODE(std::bit_xor, 6)
ACPP_OP_CODE(std::logical_and, 7)
ACPP_OP_CODE(std::logical_or, 8)
#undef ACPP_OP_CODE

// ---------------------------------------------------------------------------
// Element transforms for transform_reduce and the transform scans.  The
// library runs a closed set of transforms (acpp_unary_op in acpp_mi355x.h);
// these functors name them.  Each returns the element type: integer
// results wrap, as the kernels compute them.
// ---------------------------------------------------------------------------
#define ACPP_UNARY_FUNCTOR(NAME, EXPR)                                            \
  template <class T = void>                                                       \
  struct NAME {                                                                   \
    ACPP_HOST_DEVICE T operator()(const T &x) const { return static_cast<T>(EXPR); } \
  };                                                                              \
  template <>                                                                     \
  struct NAME<void> {                                                             \
    template <class T>                                                            \
    ACPP_HOST_DEVICE T operator()(const T &x) const { return static_cast<T>(EXPR); } \
  };
ACPP_UNARY_FUNCTOR(identity, x)
ACPP_UNARY_FUNCTOR(negate, -x)
ACPP_UNARY_FUNCTOR(absolute, (x < T(0)) ? -x : x)
ACPP_UNARY_FUNCTOR(square, x * x)
ACPP_UNARY_FUNCTOR(twice, x + x)
ACPP_UNARY_FUNCTOR(bit_not, ~x)
ACPP_UNARY_FUNCTOR(is_zero, x == T(0) ? 1 : 0)
ACPP_UNARY_FUNCTOR(is_nonzero, x != T(0) ? 1 : 0)
#undef ACPP_UNARY_FUNCTOR

// op(x, x) for a binary operator: the transform the reference's own transform
// scan tests use (tests/pstl/transform_inclusive_scan.cpp:77-79).
template <class Op>
struct self_op {
  Op op;
  template <class T>
  ACPP_HOST_DEVICE T operator()(const T &x) const { return static_cast<T>(op(x, x)); }
};

// Transform code of a functor type; value == -1 for functors the library has
// no kernel for.
template <class F>
struct unary_op_of {
  static constexpr int value = -1;
};
#define ACPP_UNARY_CODE(TMPL, CODE)                                               \
  template <class T>                                                              \
  struct unary_op_of<TMPL<T>> {                                                   \
    static constexpr int value = CODE;                                            \
  };
ACPP_UNARY_CODE(identity, 0)
ACPP_UNARY_CODE(negate, 1)
ACPP_UNARY_CODE(std::negate, 1)
ACPP_UNARY_CODE(absolute, 2)
ACPP_UNARY_CODE(square, 3)
ACPP_UNARY_CODE(twice, 4)
ACPP_UNARY_CODE(bit_not, 5)
ACPP_UNARY_CODE(std::bit_not, 5)
ACPP_UNARY_CODE(is_zero, 6)
ACPP_UNARY_CODE(is_nonzero, 7)
#undef ACPP_UNARY_CODE

// self_op<Op>: plus -> twice, multiplies -> square, minimum / maximum /
// bit_and / bit_or -> identity, logical_and / logical_or -> is_nonzero.
// bit_xor (always 0) has no code.
template <class Op>
struct unary_op_of<self_op<Op>> {
 private:
  static constexpr int op = algorithm_op_of<Op>::value;

 public:
  static constexpr int value = op == 0   ? 4
                               : op == 1 ? 3
                               : (op == 2 || op == 3 || op == 4 || op == 5) ? 0
                               : (op == 7 || op == 8) ? 7
                                                      : -1;
};

// ---------------------------------------------------------------------------
// Pair transforms of the binary transform_reduce.  The library runs a closed
// set (acpp_binary_op in acpp_mi355x.h); these functors name the ones that
// have no std:: counterpart.  binary_op_of<F> maps functor types (also
// std::multiplies, plus, minus, bit_and, bit_or, bit_xor and acpp::minimum /
// maximum) onto the codes.
// ---------------------------------------------------------------------------
template <class T = void>
struct minus {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return a - b; }
};
template <>
struct minus<void> {
  template <class A, class B>
  ACPP_HOST_DEVICE auto operator()(const A &a, const B &b) const { return a - b; }
};
template <class T = void>
struct absolute_difference {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const {
    const T d = static_cast<T>(a - b);
    return (d < T(0)) ? static_cast<T>(-d) : d;
  }
};
template <>
struct absolute_difference<void> {
  template <class T>
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return absolute_difference<T>{}(a, b); }
};
template <class T = void>
struct squared_difference {
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const {
    const T d = static_cast<T>(a - b);
    return static_cast<T>(d * d);
  }
};
template <>
struct squared_difference<void> {
  template <class T>
  ACPP_HOST_DEVICE T operator()(const T &a, const T &b) const { return squared_difference<T>{}(a, b); }
};

template <class F>
struct binary_op_of {
  static constexpr int value = -1;
};
#define ACPP_BINARY_CODE(TMPL, CODE)                                              \
  template <class T>                                                              \
  struct binary_op_of<TMPL<T>> {                                                  \
    static constexpr int value = CODE;                                            \
  };
ACPP_BINARY_CODE(multiplies, 0)
ACPP_BINARY_CODE(std::multiplies, 0)
ACPP_BINARY_CODE(plus, 1)
ACPP_BINARY_CODE(std::plus, 1)
ACPP_BINARY_CODE(minus, 2)
ACPP_BINARY_CODE(std::minus, 2)
ACPP_BINARY_CODE(absolute_difference, 3)
ACPP_BINARY_CODE(squared_difference, 4)
ACPP_BINARY_CODE(minimum, 5)
ACPP_BINARY_CODE(maximum, 6)
ACPP_BINARY_CODE(bit_and, 7)
ACPP_BINARY_CODE(std::bit_and, 7)
ACPP_BINARY_CODE(bit_or, 8)
ACPP_BINARY_CODE(std::bit_or, 8)
ACPP_BINARY_CODE(bit_xor, 9)
ACPP_BINARY_CODE(std::bit_xor, 9)
#undef ACPP_BINARY_CODE

// ---------------------------------------------------------------------------
// Predicates of copy_if.  The library evaluates a closed set of predicates
// (the ACPP_PRED_* codes of acpp_copy_if_<t>); these functors name them,
// carry their operand and can be called on the host too (for example with
// std::copy_if to compute an expected result):
//   is_nonzero, is_zero (the element transforms above, read as truth values),
//   is_equal_to(v), is_not_equal_to(v), is_less_than(v), is_less_equal(v),
//   is_greater_than(v), is_greater_equal(v), has_bits(mask, v) ((x & mask) == v,
//   integers).
// ---------------------------------------------------------------------------
template <int Code, class T>
struct value_predicate {
  static constexpr int code = Code;
  T operand;
  T mask;
  ACPP_HOST_DEVICE bool operator()(const T &x) const {
    if constexpr (Code == 1) return x == operand;
    else if constexpr (Code == 2) return x != operand;
    else if constexpr (Code == 3) return x < operand;
    else if constexpr (Code == 4) return x <= operand;
    else if constexpr (Code == 5) return x > operand;
    else if constexpr (Code == 6) return x >= operand;
    else return static_cast<T>(x & mask) == operand;
  }
};
template <class T> constexpr value_predicate<1, T> is_equal_to(T v) { return {v, T{}}; }
template <class T> constexpr value_predicate<2, T> is_not_equal_to(T v) { return {v, T{}}; }
template <class T> constexpr value_predicate<3, T> is_less_than(T v) { return {v, T{}}; }
template <class T> constexpr value_predicate<4, T> is_less_equal(T v) { return {v, T{}}; }
template <class T> constexpr value_predicate<5, T> is_greater_than(T v) { return {v, T{}}; }
template <class T> constexpr value_predicate<6, T> is_greater_equal(T v) { return {v, T{}}; }
template <class T> constexpr value_predicate<7, T> has_bits(T mask, T v) {
  static_assert(std::is_integral<T>::value, "has_bits needs an integer type");
  return {v, mask};
}

// Predicate code, operand and mask of a predicate type for element type T;
// code == -1 for predicates the library has no kernel for.
template <class P, class T>
struct predicate_of {
  static constexpr int code = -1;
};
template <int Code, class U, class T>
struct predicate_of<value_predicate<Code, U>, T> {
  static constexpr int code = std::is_same<U, T>::value ? Code : -1;
  static T operand(const value_predicate<Code, U> &p) { return p.operand; }
  static T mask(const value_predicate<Code, U> &p) { return p.mask; }
};
template <class U, class T>
struct predicate_of<is_nonzero<U>, T> {
  static constexpr int code = 0;
  static T operand(const is_nonzero<U> &) { return T{}; }
  static T mask(const is_nonzero<U> &) { return T{}; }
};
template <class U, class T>
struct predicate_of<is_zero<U>, T> {
  static constexpr int code = 1;  // x == 0
  static T operand(const is_zero<U> &) { return T{}; }
  static T mask(const is_zero<U> &) { return T{}; }
};

}  // namespace acpp
