// acpp::algorithms::copy_if, backed by the gfx950 stream-compaction kernel of
// libacpp_mi355x.so (acpp_copy_if_<t>).
//
// Same template and parameter list as
// include/hipSYCL/algorithms/algorithm.hpp:178-243:
//   copy_if(q, scratch, first, last, d_first, pred
//           [, std::size_t *num_elements_copied = nullptr][, deps])
// d_first receives, in order, the elements for which pred holds;
// *num_elements_copied (device-accessible memory) receives their number.
// An empty range returns a complete event; the count is then set to 0 (the
// reference writes it on the host, algorithm.hpp:183-187; here a kernel
// writes it, so it may live in device memory).
//
// The predicate must be one the library evaluates (acpp/functional.hpp):
// acpp::is_nonzero, is_zero, is_equal_to(v), is_not_equal_to(v),
// is_less_than(v), is_less_equal(v), is_greater_than(v), is_greater_equal(v)
// or has_bits(mask, v); anything else (a lambda, say) fails to compile.
// Element types: int32, uint32, int64, uint64, float, double.
#pragma once

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "acpp/algorithms/numeric.hpp"

namespace acpp {
namespace algorithms {
namespace detail {

template <class T>
struct copy_if_abi {
  static constexpr bool supported = false;
};
#define ACPP_FACADE_COPY_IF(CT, S, AT)                                                             \
  template <>                                                                                      \
  struct copy_if_abi<CT> {                                                                         \
    static constexpr bool supported = true;                                                        \
    static acpp_status run(acpp_allocation_group *g, hipStream_t s, const CT *f, uint64_t n, CT *d, \
                           int pred, CT operand, CT mask, uint64_t *count) {                       \
      return acpp_copy_if_##S(g, s, reinterpret_cast<const AT *>(f), n, reinterpret_cast<AT *>(d), \
                              pred, static_cast<AT>(operand), static_cast<AT>(mask), count,        \
                              nullptr);                                                            \
    }                                                                                              \
  };
ACPP_FACADE_COPY_IF(signed char, i8, int8_t)
ACPP_FACADE_COPY_IF(unsigned char, u8, uint8_t)
ACPP_FACADE_COPY_IF(short, i16, int16_t)
ACPP_FACADE_COPY_IF(unsigned short, u16, uint16_t)
ACPP_FACADE_COPY_IF(int, i32, int32_t)
ACPP_FACADE_COPY_IF(unsigned int, u32, uint32_t)
ACPP_FACADE_COPY_IF(long, i64, int64_t)
ACPP_FACADE_COPY_IF(long long, i64, int64_t)
ACPP_FACADE_COPY_IF(unsigned long, u64, uint64_t)
ACPP_FACADE_COPY_IF(unsigned long long, u64, uint64_t)
ACPP_FACADE_COPY_IF(float, f32, float)
ACPP_FACADE_COPY_IF(double, f64, double)
#undef ACPP_FACADE_COPY_IF

static_assert(sizeof(std::size_t) == sizeof(uint64_t), "the element count is written as 64 bits");

}  // namespace detail

template <class ForwardIt1, class ForwardIt2, class UnaryPredicate>
event copy_if(queue &q, util::allocation_group &scratch, ForwardIt1 first, ForwardIt1 last,
              ForwardIt2 d_first, UnaryPredicate pred, std::size_t *num_elements_copied = nullptr,
              const std::vector<event> &deps = {}) {
  using T = typename detail::contiguous<ForwardIt1>::value;
  static_assert(detail::copy_if_abi<T>::supported,
                "acpp::algorithms::copy_if: element type must be an 8-, 16-, 32- or 64-bit integer, "
                "float or double");
  static_assert(std::is_same<T, typename detail::contiguous<ForwardIt2>::value>::value,
                "acpp::algorithms::copy_if: input and output value types must be equal");
  using P = predicate_of<std::decay_t<UnaryPredicate>, T>;
  static_assert(P::code >= 0,
                "acpp::algorithms::copy_if: the predicate must be acpp::is_nonzero, is_zero, "
                "is_equal_to, is_not_equal_to, is_less_than, is_less_equal, is_greater_than, "
                "is_greater_equal or has_bits for the element type");
  const uint64_t n = static_cast<uint64_t>(std::distance(first, last));
  q.depends_on(deps);
  // n == 0 enqueues only the zeroing of the count (nothing without one).
  if (n == 0 && !num_elements_copied) return event{};
  return detail::finish(q, detail::copy_if_abi<T>::run(scratch.native(), q.get_native(),
                                                       n ? detail::to_ptr(first) : nullptr, n,
                                                       n ? detail::to_ptr(d_first) : nullptr, P::code,
                                                       P::operand(pred), P::mask(pred),
                                                       reinterpret_cast<uint64_t *>(num_elements_copied)));
}

}  // namespace algorithms
}  // namespace acpp
