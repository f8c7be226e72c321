// acpp::algorithms numeric entry points, backed by the gfx950 kernels of
// libacpp_mi355x.so through its C-ABI.
//
// Same templates, parameter lists and rules as
// include/hipSYCL/algorithms/numeric.hpp:170-327 (contract in
// doc/algorithms.md):
//   transform_reduce(q, scratch, first1, last1, first2, out, init[, reduce, transform][, deps])
//   reduce(q, scratch, first, last, out[, init[, op]][, deps])
//   transform_reduce(q, scratch, first, last, out, init, reduce, transform[, deps])
//   inclusive_scan(q, scratch, first, last, d_first[, op[, init]][, deps])
//   exclusive_scan(q, scratch, first, last, d_first, init[, op][, deps])
//   transform_inclusive_scan(q, scratch, first, last, d_first, op, unary[, init][, deps])
//   transform_exclusive_scan(q, scratch, first, last, d_first, init, op, unary[, deps])
// Every call is asynchronous on the queue's stream and returns an event;
// an empty range returns a complete default event and leaves *out untouched
// (numeric.hpp:165-169).  `deps` are waited for on the stream.
//
// What the library provides behind the templates (and what a call must
// therefore name) is checked at compile time, with no generic-functor
// fallback:
//   - iterators are pointers to device-accessible memory (or contiguous
//     std::vector iterators over such memory);
//   - element types: int8/16/32/64, uint8/16/32/64, float, double; the
//     value type of the input equals T (and the output's value type);
//   - reduction / scan operators: acpp::plus ... acpp::logical_or and the
//     std:: functors of the same operations (acpp/functional.hpp);
//     floating-point types take plus, multiplies, minimum and maximum;
//   - binary transform_reduce: reduce = plus, multiplies, minimum or maximum
//     with a pair transform (multiplies -- with plus, the dot product,
//     numeric.hpp:215-223 -- plus, minus, absolute_difference,
//     squared_difference, minimum, maximum, bit_and/or/xor) for int32, int64,
//     uint32, uint64, float and double;
//   - unary transform_reduce and transform_inclusive_scan /
//     transform_exclusive_scan (numeric.hpp:193-212, 299-327): any of the
//     operators above with an element transform from acpp/functional.hpp
//     (acpp::identity, negate, absolute, square, twice, bit_not, is_zero,
//     is_nonzero, self_op<Op>; std::negate, std::bit_not).
// Segmented forms (no counterpart in the reference): segmented_inclusive_scan
// and segmented_reduce over CSR-style offsets.
#pragma once

#include <cstdint>
#include <iterator>
#include <type_traits>
#include <vector>

#include "acpp/algorithms/util/allocation_cache.hpp"
#include "acpp/functional.hpp"
#include "acpp/queue.hpp"
#include "acpp_mi355x.h"

namespace acpp {
namespace algorithms {
namespace detail {

template <class T>
struct abi {
  static constexpr bool supported = false;
};

#define ACPP_FACADE_ABI(CT, S, AT)                                                                 \
  template <>                                                                                      \
  struct abi<CT> {                                                                                 \
    static constexpr bool supported = true;                                                        \
    static acpp_status reduce(acpp_allocation_group *g, hipStream_t s, const CT *f, uint64_t n,   \
                              CT *out, CT init, int op) {                                          \
      return acpp_reduce_##S(g, s, reinterpret_cast<const AT *>(f), n,                             \
                             reinterpret_cast<AT *>(out), static_cast<AT>(init), op, nullptr);     \
    }                                                                                              \
    static acpp_status inclusive_scan(acpp_allocation_group *g, hipStream_t s, const CT *f,       \
                                      uint64_t n, CT *d, int op, const CT *init) {                 \
      return acpp_inclusive_scan_##S(g, s, reinterpret_cast<const AT *>(f), n,                     \
                                     reinterpret_cast<AT *>(d), op,                                \
                                     reinterpret_cast<const AT *>(init), nullptr);                 \
    }                                                                                              \
    static acpp_status exclusive_scan(acpp_allocation_group *g, hipStream_t s, const CT *f,       \
                                      uint64_t n, CT *d, CT init, int op) {                        \
      return acpp_exclusive_scan_##S(g, s, reinterpret_cast<const AT *>(f), n,                     \
                                     reinterpret_cast<AT *>(d), static_cast<AT>(init), op,         \
                                     nullptr);                                                     \
    }                                                                                              \
    static acpp_status transform_reduce(acpp_allocation_group *g, hipStream_t s, const CT *f,     \
                                        uint64_t n, CT *out, CT init, int op, int t) {             \
      return acpp_transform_reduce_##S(g, s, reinterpret_cast<const AT *>(f), n,                   \
                                       reinterpret_cast<AT *>(out), static_cast<AT>(init), op, t,  \
                                       nullptr);                                                   \
    }                                                                                              \
    static acpp_status transform_inclusive_scan(acpp_allocation_group *g, hipStream_t s,          \
                                                const CT *f, uint64_t n, CT *d, int op, int t,     \
                                                const CT *init) {                                  \
      return acpp_transform_inclusive_scan_##S(g, s, reinterpret_cast<const AT *>(f), n,           \
                                               reinterpret_cast<AT *>(d), op, t,                   \
                                               reinterpret_cast<const AT *>(init), nullptr);       \
    }                                                                                              \
    static acpp_status transform_exclusive_scan(acpp_allocation_group *g, hipStream_t s,          \
                                                const CT *f, uint64_t n, CT *d, CT init, int op,   \
                                                int t) {                                           \
      return acpp_transform_exclusive_scan_##S(g, s, reinterpret_cast<const AT *>(f), n,           \
                                               reinterpret_cast<AT *>(d), static_cast<AT>(init),   \
                                               op, t, nullptr);                                    \
    }                                                                                              \
  };
ACPP_FACADE_ABI(signed char, i8, int8_t)
ACPP_FACADE_ABI(short, i16, int16_t)
ACPP_FACADE_ABI(int, i32, int32_t)
ACPP_FACADE_ABI(long, i64, int64_t)
ACPP_FACADE_ABI(long long, i64, int64_t)
ACPP_FACADE_ABI(unsigned char, u8, uint8_t)
ACPP_FACADE_ABI(unsigned short, u16, uint16_t)
ACPP_FACADE_ABI(unsigned int, u32, uint32_t)
ACPP_FACADE_ABI(unsigned long, u64, uint64_t)
ACPP_FACADE_ABI(unsigned long long, u64, uint64_t)
ACPP_FACADE_ABI(float, f32, float)
ACPP_FACADE_ABI(double, f64, double)
#undef ACPP_FACADE_ABI

static_assert(sizeof(long) == 8 && sizeof(long long) == 8, "LP64 expected");

template <class T>
struct dot_abi {
  static constexpr bool supported = false;
};
#define ACPP_FACADE_DOT(CT, S, AT)                                                                 \
  template <>                                                                                      \
  struct dot_abi<CT> {                                                                             \
    static constexpr bool supported = true;                                                        \
    static acpp_status dot(acpp_allocation_group *g, hipStream_t s, const CT *a, uint64_t n,      \
                           const CT *b, CT *out, CT init) {                                        \
      return acpp_transform_reduce_dot_##S(g, s, reinterpret_cast<const AT *>(a), n,               \
                                           reinterpret_cast<const AT *>(b),                        \
                                           reinterpret_cast<AT *>(out), static_cast<AT>(init),     \
                                           nullptr);                                               \
    }                                                                                              \
  };
ACPP_FACADE_DOT(int, i32, int32_t)
ACPP_FACADE_DOT(long, i64, int64_t)
ACPP_FACADE_DOT(long long, i64, int64_t)
ACPP_FACADE_DOT(unsigned int, u32, uint32_t)
ACPP_FACADE_DOT(unsigned long, u64, uint64_t)
ACPP_FACADE_DOT(unsigned long long, u64, uint64_t)
ACPP_FACADE_DOT(float, f32, float)
ACPP_FACADE_DOT(double, f64, double)
#undef ACPP_FACADE_DOT

template <class T>
struct binary_abi {
  static constexpr bool supported = false;
};
#define ACPP_FACADE_BINARY(CT, S, AT)                                                              \
  template <>                                                                                      \
  struct binary_abi<CT> {                                                                          \
    static constexpr bool supported = true;                                                        \
    static acpp_status run(acpp_allocation_group *g, hipStream_t s, const CT *a, uint64_t n,      \
                           const CT *b, CT *out, CT init, int op, int code) {                      \
      return acpp_transform_reduce_binary_##S(g, s, reinterpret_cast<const AT *>(a), n,            \
                                              reinterpret_cast<const AT *>(b),                     \
                                              reinterpret_cast<AT *>(out), static_cast<AT>(init),  \
                                              op, code, nullptr);                                  \
    }                                                                                              \
  };
ACPP_FACADE_BINARY(int, i32, int32_t)
ACPP_FACADE_BINARY(long, i64, int64_t)
ACPP_FACADE_BINARY(long long, i64, int64_t)
ACPP_FACADE_BINARY(unsigned int, u32, uint32_t)
ACPP_FACADE_BINARY(unsigned long, u64, uint64_t)
ACPP_FACADE_BINARY(unsigned long long, u64, uint64_t)
ACPP_FACADE_BINARY(float, f32, float)
ACPP_FACADE_BINARY(double, f64, double)
#undef ACPP_FACADE_BINARY

// Contiguous iterators: pointers, and the pointer-wrapping iterators of
// contiguous standard containers (std::vector with any allocator).
template <class It>
struct wraps_pointer : std::false_type {};
#if defined(__GLIBCXX__)
template <class P, class C>
struct wraps_pointer<__gnu_cxx::__normal_iterator<P, C>> : std::is_pointer<P> {};
#endif

template <class It>
struct contiguous {
  using value = std::remove_cv_t<typename std::iterator_traits<It>::value_type>;
  static constexpr bool pointer = std::is_pointer<It>::value;
  static constexpr bool vector_it =
      std::is_same<It, typename std::vector<value>::iterator>::value ||
      std::is_same<It, typename std::vector<value>::const_iterator>::value || wraps_pointer<It>::value;
  static constexpr bool value_ok = pointer || vector_it;
};

template <class It>
auto to_ptr(It it) {
  static_assert(contiguous<It>::value_ok,
                "acpp::algorithms: iterators must be pointers (or std::vector iterators) to "
                "device-accessible memory");
  return &*it;
}

template <class Op, class T>
constexpr int op_code() {
  constexpr int code = algorithm_op_of<std::decay_t<Op>>::value;
  static_assert(code >= 0,
                "acpp::algorithms: operator must be acpp::plus/multiplies/minimum/maximum/"
                "bit_and/bit_or/bit_xor/logical_and/logical_or or the std:: functor of one of "
                "these operations");
  static_assert(!std::is_floating_point<T>::value || code <= 3,
                "acpp::algorithms: floating-point types take plus, multiplies, minimum or maximum");
  return code;
}

template <class F, class T>
constexpr int unary_code() {
  constexpr int code = unary_op_of<std::decay_t<F>>::value;
  static_assert(code >= 0,
                "acpp::algorithms: the element transform must be acpp::identity/negate/absolute/"
                "square/twice/bit_not/is_zero/is_nonzero, std::negate, std::bit_not or "
                "acpp::self_op<Op>");
  static_assert(std::is_integral<T>::value || code != 5,
                "acpp::algorithms: bit_not needs an integer element type");
  return code;
}

template <class T>
constexpr void check_type() {
  static_assert(abi<T>::supported,
                "acpp::algorithms: element type must be an 8/16/32/64-bit integer, float or double");
}

inline event finish(queue &q, acpp_status st) {
  acpp::detail::check_status(st);
  return q.record();
}

}  // namespace detail

// ------------------------------------------------------------------ reduce

// Binary form (numeric.hpp:170-191): *out = reduce(init, transform(x[i], y[i])...).
// reduce: plus, multiplies, minimum or maximum; transform: a pair transform
// of acpp/functional.hpp (multiplies, plus, minus, absolute_difference,
// squared_difference, minimum, maximum, bit_and, bit_or, bit_xor; std::
// functors of the same operations).  plus with multiplies is the dot product.
template <class ForwardIt1, class ForwardIt2, class T, class BinaryReductionOp,
          class BinaryTransformOp>
event transform_reduce(queue &q, util::allocation_group &scratch, ForwardIt1 first1,
                       ForwardIt1 last1, ForwardIt2 first2, T *out, T init,
                       BinaryReductionOp reduce, BinaryTransformOp transform,
                       const std::vector<event> &deps = {}) {
  static_assert(detail::binary_abi<T>::supported,
                "acpp::algorithms::transform_reduce: the binary form takes int32/int64/uint32/"
                "uint64/float/double");
  constexpr int op = detail::op_code<BinaryReductionOp, T>();
  static_assert(op <= 3,
                "acpp::algorithms::transform_reduce: the binary form reduces with plus, multiplies, "
                "minimum or maximum");
  constexpr int code = binary_op_of<std::decay_t<BinaryTransformOp>>::value;
  static_assert(code >= 0,
                "acpp::algorithms::transform_reduce: the pair transform must be multiplies, plus, "
                "minus, absolute_difference, squared_difference, minimum, maximum, bit_and, bit_or "
                "or bit_xor");
  static_assert(std::is_integral<T>::value || code <= 6,
                "acpp::algorithms::transform_reduce: bitwise pair transforms need an integer type");
  static_assert(std::is_same<typename detail::contiguous<ForwardIt1>::value, T>::value &&
                    std::is_same<typename detail::contiguous<ForwardIt2>::value, T>::value,
                "acpp::algorithms::transform_reduce: input value types must equal T");
  (void)reduce;
  (void)transform;
  if (first1 == last1) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first1, last1));
  q.depends_on(deps);
  if constexpr (op == 0 && code == 0)
    return detail::finish(q, detail::dot_abi<T>::dot(scratch.native(), q.get_native(),
                                                     detail::to_ptr(first1), n,
                                                     detail::to_ptr(first2), out, init));
  else
    return detail::finish(q, detail::binary_abi<T>::run(scratch.native(), q.get_native(),
                                                        detail::to_ptr(first1), n,
                                                        detail::to_ptr(first2), out, init, op, code));
}

template <class ForwardIt1, class ForwardIt2, class T>
event transform_reduce(queue &q, util::allocation_group &scratch, ForwardIt1 first1,
                       ForwardIt1 last1, ForwardIt2 first2, T *out, T init,
                       const std::vector<event> &deps = {}) {
  return transform_reduce(q, scratch, first1, last1, first2, out, init, std::plus<T>{},
                          std::multiplies<T>{}, deps);
}

// Unary form (numeric.hpp:193-212): *out = reduce(init, transform(x[i])...).
template <class ForwardIt, class T, class BinaryReductionOp, class UnaryTransformOp>
event transform_reduce(queue &q, util::allocation_group &scratch, ForwardIt first, ForwardIt last,
                       T *out, T init, BinaryReductionOp reduce, UnaryTransformOp transform,
                       const std::vector<event> &deps = {}) {
  detail::check_type<T>();
  static_assert(std::is_same<typename detail::contiguous<ForwardIt>::value, T>::value,
                "acpp::algorithms::transform_reduce: the input value type must equal T");
  constexpr int op = detail::op_code<BinaryReductionOp, T>();
  constexpr int t = detail::unary_code<UnaryTransformOp, T>();
  (void)reduce;
  (void)transform;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  return detail::finish(q, detail::abi<T>::transform_reduce(scratch.native(), q.get_native(),
                                                            detail::to_ptr(first), n, out, init, op,
                                                            t));
}

template <class ForwardIt, class T, class BinaryOp>
event reduce(queue &q, util::allocation_group &scratch, ForwardIt first, ForwardIt last, T *out,
             T init, BinaryOp binary_op, const std::vector<event> &deps = {}) {
  detail::check_type<T>();
  static_assert(std::is_same<typename detail::contiguous<ForwardIt>::value, T>::value,
                "acpp::algorithms::reduce: the input value type must equal T");
  constexpr int op = detail::op_code<BinaryOp, T>();
  (void)binary_op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  return detail::finish(q, detail::abi<T>::reduce(scratch.native(), q.get_native(),
                                                  detail::to_ptr(first), n, out, init, op));
}

template <class ForwardIt, class T>
event reduce(queue &q, util::allocation_group &scratch, ForwardIt first, ForwardIt last, T *out,
             T init, const std::vector<event> &deps = {}) {
  return reduce(q, scratch, first, last, out, init, std::plus<T>{}, deps);
}

template <class ForwardIt>
event reduce(queue &q, util::allocation_group &scratch, ForwardIt first, ForwardIt last,
             typename std::iterator_traits<ForwardIt>::value_type *out,
             const std::vector<event> &deps = {}) {
  using T = typename std::iterator_traits<ForwardIt>::value_type;
  return reduce(q, scratch, first, last, out, T{}, deps);
}

// ------------------------------------------------------------------- scans

namespace detail {
template <class InputIt, class OutputIt>
using scan_value_t = typename contiguous<InputIt>::value;

template <class InputIt, class OutputIt>
constexpr void check_scan_types() {
  using V = scan_value_t<InputIt, OutputIt>;
  check_type<V>();
  static_assert(std::is_same<V, typename contiguous<OutputIt>::value>::value,
                "acpp::algorithms: input and output value types must be equal");
}
}  // namespace detail

template <class InputIt, class OutputIt, class BinaryOp>
event inclusive_scan(queue &q, util::allocation_group &scratch, InputIt first, InputIt last,
                     OutputIt d_first, BinaryOp op, const std::vector<event> &deps = {}) {
  detail::check_scan_types<InputIt, OutputIt>();
  using V = detail::scan_value_t<InputIt, OutputIt>;
  constexpr int code = detail::op_code<BinaryOp, V>();
  (void)op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  return detail::finish(q, detail::abi<V>::inclusive_scan(scratch.native(), q.get_native(),
                                                          detail::to_ptr(first), n,
                                                          detail::to_ptr(d_first), code, nullptr));
}

template <class InputIt, class OutputIt, class BinaryOp, class T>
event inclusive_scan(queue &q, util::allocation_group &scratch, InputIt first, InputIt last,
                     OutputIt d_first, BinaryOp op, T init, const std::vector<event> &deps = {}) {
  detail::check_scan_types<InputIt, OutputIt>();
  using V = detail::scan_value_t<InputIt, OutputIt>;
  constexpr int code = detail::op_code<BinaryOp, V>();
  (void)op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  const V init_v = static_cast<V>(init);
  q.depends_on(deps);
  return detail::finish(q, detail::abi<V>::inclusive_scan(scratch.native(), q.get_native(),
                                                          detail::to_ptr(first), n,
                                                          detail::to_ptr(d_first), code, &init_v));
}

template <class InputIt, class OutputIt>
event inclusive_scan(queue &q, util::allocation_group &scratch, InputIt first, InputIt last,
                     OutputIt d_first, const std::vector<event> &deps = {}) {
  return inclusive_scan(q, scratch, first, last, d_first, std::plus<>{}, deps);
}

template <class InputIt, class OutputIt, class T, class BinaryOp>
event exclusive_scan(queue &q, util::allocation_group &scratch, InputIt first, InputIt last,
                     OutputIt d_first, T init, BinaryOp op, const std::vector<event> &deps = {}) {
  detail::check_scan_types<InputIt, OutputIt>();
  using V = detail::scan_value_t<InputIt, OutputIt>;
  constexpr int code = detail::op_code<BinaryOp, V>();
  (void)op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  return detail::finish(q, detail::abi<V>::exclusive_scan(scratch.native(), q.get_native(),
                                                          detail::to_ptr(first), n,
                                                          detail::to_ptr(d_first),
                                                          static_cast<V>(init), code));
}

template <class InputIt, class OutputIt, class T>
event exclusive_scan(queue &q, util::allocation_group &scratch, InputIt first, InputIt last,
                     OutputIt d_first, T init, const std::vector<event> &deps = {}) {
  return exclusive_scan(q, scratch, first, last, d_first, init, std::plus<>{}, deps);
}

// scan.hpp:85-132: the scans of transform(x[i]).
template <class InputIt, class OutputIt, class BinaryOp, class UnaryOp>
event transform_inclusive_scan(queue &q, util::allocation_group &scratch, InputIt first,
                               InputIt last, OutputIt d_first, BinaryOp binary_op, UnaryOp unary_op,
                               const std::vector<event> &deps = {}) {
  detail::check_scan_types<InputIt, OutputIt>();
  using V = detail::scan_value_t<InputIt, OutputIt>;
  constexpr int code = detail::op_code<BinaryOp, V>();
  constexpr int t = detail::unary_code<UnaryOp, V>();
  (void)binary_op;
  (void)unary_op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  return detail::finish(q, detail::abi<V>::transform_inclusive_scan(
                               scratch.native(), q.get_native(), detail::to_ptr(first), n,
                               detail::to_ptr(d_first), code, t, nullptr));
}

template <class InputIt, class OutputIt, class BinaryOp, class UnaryOp, class T>
event transform_inclusive_scan(queue &q, util::allocation_group &scratch, InputIt first,
                               InputIt last, OutputIt d_first, BinaryOp binary_op, UnaryOp unary_op,
                               T init, const std::vector<event> &deps = {}) {
  detail::check_scan_types<InputIt, OutputIt>();
  using V = detail::scan_value_t<InputIt, OutputIt>;
  constexpr int code = detail::op_code<BinaryOp, V>();
  constexpr int t = detail::unary_code<UnaryOp, V>();
  (void)binary_op;
  (void)unary_op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  const V init_v = static_cast<V>(init);
  q.depends_on(deps);
  return detail::finish(q, detail::abi<V>::transform_inclusive_scan(
                               scratch.native(), q.get_native(), detail::to_ptr(first), n,
                               detail::to_ptr(d_first), code, t, &init_v));
}

template <class InputIt, class OutputIt, class T, class BinaryOp, class UnaryOp>
event transform_exclusive_scan(queue &q, util::allocation_group &scratch, InputIt first,
                               InputIt last, OutputIt d_first, T init, BinaryOp binary_op,
                               UnaryOp unary_op, const std::vector<event> &deps = {}) {
  detail::check_scan_types<InputIt, OutputIt>();
  using V = detail::scan_value_t<InputIt, OutputIt>;
  constexpr int code = detail::op_code<BinaryOp, V>();
  constexpr int t = detail::unary_code<UnaryOp, V>();
  (void)binary_op;
  (void)unary_op;
  if (first == last) return event{};
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  return detail::finish(q, detail::abi<V>::transform_exclusive_scan(
                               scratch.native(), q.get_native(), detail::to_ptr(first), n,
                               detail::to_ptr(d_first), static_cast<V>(init), code, t));
}

// ------------------------------------------------ segmented (extension)

namespace detail {
template <class T>
struct seg_abi {
  static constexpr bool supported = false;
};
#define ACPP_FACADE_SEG(CT, S, AT)                                                                 \
  template <>                                                                                      \
  struct seg_abi<CT> {                                                                             \
    static constexpr bool supported = true;                                                        \
    static acpp_status scan(acpp_allocation_group *g, hipStream_t s, const CT *f, uint64_t n,     \
                            CT *d, const uint64_t *off, uint64_t segs, int op) {                   \
      return acpp_segmented_inclusive_scan_##S(g, s, reinterpret_cast<const AT *>(f), n,          \
                                               reinterpret_cast<AT *>(d), off, segs, op, nullptr);  \
    }                                                                                              \
    static acpp_status reduce(acpp_allocation_group *g, hipStream_t s, const CT *f, uint64_t n,   \
                              const uint64_t *off, uint64_t segs, CT *out, CT init, int op) {      \
      return acpp_segmented_reduce_##S(g, s, reinterpret_cast<const AT *>(f), n, off, segs,       \
                                       reinterpret_cast<AT *>(out), static_cast<AT>(init), op,     \
                                       nullptr);                                                   \
    }                                                                                              \
  };
ACPP_FACADE_SEG(int, i32, int32_t)
ACPP_FACADE_SEG(long, i64, int64_t)
ACPP_FACADE_SEG(long long, i64, int64_t)
ACPP_FACADE_SEG(float, f32, float)
ACPP_FACADE_SEG(double, f64, double)
#undef ACPP_FACADE_SEG
}  // namespace detail

// Inclusive scan of every segment [first + offsets[s], first + offsets[s+1])
// of [first, last); `offsets` (device memory, num_segments + 1 entries,
// non-decreasing, offsets[0] = 0, offsets[num_segments] = last - first).
template <class T, class BinaryOp = std::plus<T>>
event segmented_inclusive_scan(queue &q, util::allocation_group &scratch, const T *first,
                               const T *last, T *d_first, const uint64_t *offsets,
                               uint64_t num_segments, BinaryOp op = {},
                               const std::vector<event> &deps = {}) {
  static_assert(detail::seg_abi<T>::supported, "segmented forms: int32, int64, float, double");
  constexpr int code = detail::op_code<BinaryOp, T>();
  (void)op;
  q.depends_on(deps);
  return detail::finish(q, detail::seg_abi<T>::scan(scratch.native(), q.get_native(), first,
                                                    static_cast<uint64_t>(last - first), d_first,
                                                    offsets, num_segments, code));
}

// out[s] = op(init, fold of segment s); empty segments give init.
template <class T, class BinaryOp = std::plus<T>>
event segmented_reduce(queue &q, util::allocation_group &scratch, const T *first, const T *last,
                       const uint64_t *offsets, uint64_t num_segments, T *out, T init,
                       BinaryOp op = {}, const std::vector<event> &deps = {}) {
  static_assert(detail::seg_abi<T>::supported, "segmented forms: int32, int64, float, double");
  constexpr int code = detail::op_code<BinaryOp, T>();
  (void)op;
  q.depends_on(deps);
  return detail::finish(q, detail::seg_abi<T>::reduce(scratch.native(), q.get_native(), first,
                                                      static_cast<uint64_t>(last - first),
                                                      offsets, num_segments, out, init, code));
}

}  // namespace algorithms
}  // namespace acpp
