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

}  // namespace acpp
