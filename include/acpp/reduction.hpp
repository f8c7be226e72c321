// sycl::reduction for the algorithms facade: reduction objects with the
// reference's factories and properties, and the bulk parallel_for that runs
// them on the gfx950 reduction kernels.
//
// Factories and rules as include/hipSYCL/sycl/libkernel/reduction.hpp:
// 165-236 (pointer forms; property::reduction::initialize_to_identity):
//   acpp::reduction(T *var, combiner[, property_list])
//   acpp::reduction(T *var, const T &identity, combiner[, property_list])
// Result (algorithms/reduction/reduction_descriptor.hpp:36-53): with
// initialize_to_identity *var = combiner(identity, fold), an empty range
// leaving the identity; otherwise *var = combiner(*var, fold).
//
// The reference runs a reduction through handler::parallel_for(range,
// reductions..., kernel) (include/hipSYCL/sycl/handler.hpp:799-975), whose
// kernel calls reducer.combine(value).  A kernel cannot cross the C-ABI, so
// the bulk form fixes the kernel: work-item i combines x[i]:
//   acpp::parallel_for(q, scratch, n, acpp::combine_each(x, reduction)[, deps])
//   acpp::parallel_for(q, scratch, n, combine_each(x1, r1), ..., combine_each(x4, r4))
// Two to four reductions over inputs of one length and element type run
// fused in one kernel (acpp_reduction_multi_identity_<t>).  Combiners are
// the library's operators (acpp/functional.hpp); anything else fails to
// compile.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

#include "acpp/algorithms/numeric.hpp"

namespace acpp {

namespace property {
namespace reduction {
struct initialize_to_identity {};
}  // namespace reduction
}  // namespace property

class property_list {
 public:
  property_list() = default;
  property_list(property::reduction::initialize_to_identity) : init_to_identity_(true) {}

  template <class P>
  bool has_property() const {
    return std::is_same<P, property::reduction::initialize_to_identity>::value && init_to_identity_;
  }

 private:
  bool init_to_identity_ = false;
};

template <class T, class BinaryOperation>
struct reduction_descriptor {
  T *var;
  BinaryOperation combiner;
  bool initialize_to_identity;
  bool has_identity;
  T identity;
};

template <class T, class BinaryOperation>
reduction_descriptor<T, BinaryOperation> reduction(T *var, BinaryOperation combiner,
                                                   const property_list &props = {}) {
  return {var, combiner, props.has_property<property::reduction::initialize_to_identity>(), false, T{}};
}

template <class T, class BinaryOperation>
reduction_descriptor<T, BinaryOperation> reduction(T *var, const T &identity, BinaryOperation combiner,
                                                   const property_list &props = {}) {
  return {var, combiner, props.has_property<property::reduction::initialize_to_identity>(), true, identity};
}

// Work-item i of the bulk parallel_for calls reducer.combine(x[i]).
template <class T, class BinaryOperation>
struct combine_from {
  const T *x;
  reduction_descriptor<T, BinaryOperation> red;
};

template <class T, class BinaryOperation>
combine_from<T, BinaryOperation> combine_each(const T *x, const reduction_descriptor<T, BinaryOperation> &red) {
  return {x, red};
}

namespace detail {

template <class T>
struct reduction_abi {
  static constexpr bool supported = false;
  static constexpr bool fused = false;
};
#define ACPP_FACADE_RED(CT, S, AT, FUSED)                                                          \
  template <>                                                                                      \
  struct reduction_abi<CT> {                                                                       \
    static constexpr bool supported = true;                                                        \
    static constexpr bool fused = FUSED;                                                           \
    static acpp_status run(acpp_allocation_group *g, hipStream_t s, const CT *x, uint64_t n, CT *out, \
                           int op, int init, bool has_identity, CT identity) {                     \
      if (has_identity)                                                                            \
        return acpp_reduction_identity_##S(g, s, reinterpret_cast<const AT *>(x), n,               \
                                           reinterpret_cast<AT *>(out), static_cast<AT>(identity), \
                                           op, init, nullptr);                                     \
      return acpp_reduction_##S(g, s, reinterpret_cast<const AT *>(x), n, reinterpret_cast<AT *>(out), \
                                op, init, nullptr);                                                \
    }                                                                                              \
    ACPP_FACADE_RED_MULTI_##FUSED(CT, S, AT)                                                       \
  };
#define ACPP_FACADE_RED_MULTI_false(CT, S, AT)
#define ACPP_FACADE_RED_MULTI_true(CT, S, AT)                                                      \
  static acpp_status multi(acpp_allocation_group *g, hipStream_t s, uint32_t count, const CT *const *x, \
                           uint64_t n, CT *const *outs, const acpp_algorithm_op *ops, const CT *ids, \
                           const int *init) {                                                      \
    return acpp_reduction_multi_identity_##S(g, s, count, reinterpret_cast<const AT *const *>(x), n, \
                                             reinterpret_cast<AT *const *>(outs), ops,              \
                                             reinterpret_cast<const AT *>(ids), init, nullptr);     \
  }
ACPP_FACADE_RED(signed char, i8, int8_t, true)
ACPP_FACADE_RED(short, i16, int16_t, true)
ACPP_FACADE_RED(int, i32, int32_t, true)
ACPP_FACADE_RED(long, i64, int64_t, true)
ACPP_FACADE_RED(long long, i64, int64_t, true)
ACPP_FACADE_RED(unsigned char, u8, uint8_t, true)
ACPP_FACADE_RED(unsigned short, u16, uint16_t, true)
ACPP_FACADE_RED(unsigned int, u32, uint32_t, true)
ACPP_FACADE_RED(unsigned long, u64, uint64_t, true)
ACPP_FACADE_RED(unsigned long long, u64, uint64_t, true)
ACPP_FACADE_RED(float, f32, float, true)
ACPP_FACADE_RED(double, f64, double, true)
#undef ACPP_FACADE_RED
#undef ACPP_FACADE_RED_MULTI_false
#undef ACPP_FACADE_RED_MULTI_true

// The library's identity of operator `op` for T (the device kernels' own,
// acpp/device/ops.hpp): -0.0 for floating-point sums, +-infinity for
// floating-point minimum / maximum.
template <class T>
T identity_of(int op) {
  switch (op) {
    case 0:
      if (std::is_floating_point<T>::value) return static_cast<T>(-0.0);
      return T(0);
    case 1: return T(1);
    case 2:
      if (std::is_floating_point<T>::value) return std::numeric_limits<T>::infinity();
      return std::numeric_limits<T>::max();
    case 3:
      if (std::is_floating_point<T>::value) return -std::numeric_limits<T>::infinity();
      return std::numeric_limits<T>::lowest();
    case 4: return static_cast<T>(~static_cast<std::make_unsigned_t<std::conditional_t<
                                      std::is_integral<T>::value, T, int>>>(0));
    case 7: return T(1);
    default: return T(0);
  }
}

template <class T>
constexpr void check_reduction_type() {
  static_assert(reduction_abi<T>::supported,
                "acpp::reduction: element type must be an 8/16/32/64-bit integer, float or double");
}

}  // namespace detail

// One reduction: handler::parallel_for(range<1>{n}, reduction, kernel).
template <class T, class BinaryOperation>
event parallel_for(queue &q, algorithms::util::allocation_group &scratch, std::size_t n,
                   const combine_from<T, BinaryOperation> &item, const std::vector<event> &deps = {}) {
  detail::check_reduction_type<T>();
  constexpr int op = algorithms::detail::op_code<BinaryOperation, T>();
  const auto &r = item.red;
  q.depends_on(deps);
  return algorithms::detail::finish(
      q, detail::reduction_abi<T>::run(scratch.native(), q.get_native(), item.x, static_cast<uint64_t>(n), r.var,
                                       op, r.initialize_to_identity ? 1 : 0, r.has_identity, r.identity));
}

// Two to four reductions fused in one kernel (handler.hpp:799-975).
template <class T, class... Ops>
event parallel_for(queue &q, algorithms::util::allocation_group &scratch, std::size_t n,
                   const combine_from<T, Ops> &...items) {
  constexpr std::size_t k = sizeof...(Ops);
  static_assert(k >= 2 && k <= 4, "acpp::parallel_for: one to four reductions");
  detail::check_reduction_type<T>();
  static_assert(detail::reduction_abi<T>::fused,
                "acpp::parallel_for: fused reductions take 8/16/32/64-bit integers, float or double");
  const T *inputs[k] = {items.x...};
  T *outs[k] = {items.red.var...};
  const acpp_algorithm_op ops[k] = {algorithms::detail::op_code<Ops, T>()...};
  const int init[k] = {(items.red.initialize_to_identity ? 1 : 0)...};
  const bool given[k] = {items.red.has_identity...};
  // A reduction without a given identity uses its operator's own; the
  // library takes the identities of all reductions or none, so the missing
  // ones are filled in with the library's identities.
  bool any_given = false;
  for (bool g : given) any_given = any_given || g;
  T ids[k] = {items.red.identity...};
  if (any_given) {
    for (std::size_t j = 0; j < k; ++j)
      if (!given[j]) ids[j] = detail::identity_of<T>(ops[j]);
  }
  return algorithms::detail::finish(
      q, detail::reduction_abi<T>::multi(scratch.native(), q.get_native(), static_cast<uint32_t>(k), inputs,
                                         static_cast<uint64_t>(n), outs, ops, any_given ? ids : nullptr, init));
}

}  // namespace acpp
