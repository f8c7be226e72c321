// Device side of sycl::reduction for user HIP kernels: a per-work-item
// reducer and the end of the reduction, where a work-group's values reach
// the result.
//
// Reference: the reducer handed to a kernel by
// handler::parallel_for(range, reduction, kernel)
// (include/hipSYCL/sycl/libkernel/reduction.hpp:61-160: combine, identity,
// the operators +=, *=, &=, |=, ^= and ++ for the operators they name), fed
// by the algorithms library's per-work-item sequential_reducer
// (include/hipSYCL/algorithms/reduction/wg_model/wi_reducer.hpp:36-88) and
// finished by group_horizontal_reducer::finalize plus the engine's stages
// (wg_model/group_horizontal_reducer.hpp:54-80, reduction_engine.hpp:116-390).
//
// Here the end of the reduction is one call inside the user kernel:
//   acpp::reducer<int, acpp::plus<int>> r;            // starts at the identity
//   for (i = ...; i < n; i += stride) r += x[i];        // or r.combine(v)
//   acpp::combine_into(acpp::this_group(), r, out);     // *out = op(*out, group total)
// combine_into reduces the work-group's values with reduce_over_group (DPP
// within each wave, one LDS exchange across waves) and one work-item folds
// the total into *out with a device-scope atomic read-modify-write:
// fetch_add / fetch_min / fetch_max / fetch_and / fetch_or / fetch_xor where
// the operator has one, a compare-and-swap loop for multiplies and the
// logical operators.  Every work-item of the group must call it (it
// contains a barrier).  The groups' totals arrive in any order: the result
// is exact for integers and the bitwise operators, and for floating-point
// plus it is the fold of the same group totals in the order the atomics
// happen to land.  A result that starts from the identity
// (property::reduction::initialize_to_identity) has *out set to the
// identity before the kernel runs.
#pragma once

#include "acpp/device/atomic_ref.hpp"
#include "acpp/device/group_functions.hpp"
#include "acpp/functional.hpp"

namespace acpp {

template <class T, class BinaryOperation>
class reducer {
 public:
  using value_type = T;
  using binary_operation = BinaryOperation;

  // Starts at the operator's known identity (sycl::known_identity).
  __device__ reducer() : value_(known_identity<BinaryOperation, T>::get()), identity_(value_) {
    static_assert(known_identity<BinaryOperation, T>::value,
                  "acpp::reducer: no known identity for this operator and type; pass one");
  }
  // Starts at a user identity (the reduction(var, identity, op) form).
  __device__ explicit reducer(const T &identity) : value_(identity), identity_(identity) {}

  __device__ reducer &combine(const T &x) {
    value_ = BinaryOperation{}(value_, x);
    return *this;
  }
  __device__ T identity() const { return identity_; }
  __device__ T value() const { return value_; }

  // The operators of the reference's reducer, each for the operator it
  // names (reduction.hpp:104-160).
  template <class Op = BinaryOperation, class = std::enable_if_t<std::is_same_v<Op, plus<T>>>>
  __device__ reducer &operator+=(const T &x) {
    return combine(x);
  }
  template <class Op = BinaryOperation, class = std::enable_if_t<std::is_same_v<Op, multiplies<T>>>>
  __device__ reducer &operator*=(const T &x) {
    return combine(x);
  }
  template <class Op = BinaryOperation, class = std::enable_if_t<std::is_same_v<Op, bit_and<T>>>>
  __device__ reducer &operator&=(const T &x) {
    return combine(x);
  }
  template <class Op = BinaryOperation, class = std::enable_if_t<std::is_same_v<Op, bit_or<T>>>>
  __device__ reducer &operator|=(const T &x) {
    return combine(x);
  }
  template <class Op = BinaryOperation, class = std::enable_if_t<std::is_same_v<Op, bit_xor<T>>>>
  __device__ reducer &operator^=(const T &x) {
    return combine(x);
  }
  template <class Op = BinaryOperation,
            class = std::enable_if_t<std::is_same_v<Op, plus<T>> && std::is_integral_v<T>>>
  __device__ reducer &operator++() {
    return combine(T(1));
  }

 private:
  T value_;
  T identity_;
};

namespace detail {

// *out = op(*out, v), atomically at device scope.
template <class T, class Op>
__device__ void atomic_combine(T *out, T v) {
  atomic_ref<T, memory_order::relaxed, memory_scope::device, address_space::global_space> a(*out);
  if constexpr (std::is_same_v<Op, plus<T>>) {
    a.fetch_add(v);
  } else if constexpr (std::is_same_v<Op, minimum<T>>) {
    a.fetch_min(v);
  } else if constexpr (std::is_same_v<Op, maximum<T>>) {
    a.fetch_max(v);
  } else if constexpr (std::is_integral_v<T> && std::is_same_v<Op, bit_and<T>>) {
    a.fetch_and(v);
  } else if constexpr (std::is_integral_v<T> && std::is_same_v<Op, bit_or<T>>) {
    a.fetch_or(v);
  } else if constexpr (std::is_integral_v<T> && std::is_same_v<Op, bit_xor<T>>) {
    a.fetch_xor(v);
  } else {
    T expected = a.load();
    while (!a.compare_exchange_strong(expected, Op{}(expected, v))) {
    }
  }
}

}  // namespace detail

// The end of a reduction inside a user kernel: *out = op(*out, fold of every
// work-item's reducer value in the group).  Called by every work-item of the
// work-group; one atomic per group.
template <class T, class BinaryOperation>
__device__ void combine_into(group g, const reducer<T, BinaryOperation> &r, T *out) {
  const T total = reduce_over_group(g, r.value(), BinaryOperation{});
  if (g.leader()) detail::atomic_combine<T, BinaryOperation>(out, total);
}

}  // namespace acpp
