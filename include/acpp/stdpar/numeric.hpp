// Standard-parallelism entry points for the numeric algorithms: reduce,
// transform_reduce, inclusive_scan, exclusive_scan, transform_inclusive_scan
// and transform_exclusive_scan with the par_unseq policy
// (std::execution::parallel_unsequenced_policy), offloaded to the MI355X
// kernels through the acpp::algorithms facade or run on the host.
//
// Same entry points, argument lists and offload structure as the reference's
// stdpar layer (include/hipSYCL/std/stdpar/pstl-impl/numeric.hpp:28-580):
//   - a thread-local in-order queue and thread-local scratch caches
//     (include/hipSYCL/std/stdpar/detail/sycl_glue.hpp:54-90);
//   - reductions block until their result is on the host
//     (HIPSYCL_STDPAR_BLOCKING_OFFLOAD); scans return as soon as their work
//     is enqueued (HIPSYCL_STDPAR_OFFLOAD, numeric.hpp:298-358).  The
//     reference's compiler then inserts a synchronisation before the host
//     touches the data; without that compiler pass the caller synchronises
//     explicitly: acpp::stdpar::sync(), or any blocking call of this header
//     (the queue is in order).
//   - every call decides between the GPU and the host first
//     (acpp_stdpar_should_offload: memory only the device can read forces
//     the GPU, pageable host memory forces the host, and for pinned or
//     managed memory the library's learned cost model picks the faster
//     side; the reference interpolates sampled run times instead,
//     offload.hpp:384-594, offload_heuristic_db.hpp:247-318).  The host side
//     is the sequential C++ standard library algorithm, run after the
//     thread's offloaded work has completed.
// Element types, operators and transforms are those of the acpp::algorithms
// facade (acpp/algorithms/numeric.hpp); others fail to compile.
//
// To make std::reduce(std::execution::par_unseq, ...) itself offload, put
// include/acpp/stdpar/std before the system include directories: its
// <numeric> adds std:: overloads that forward here, as the reference's
// include/hipSYCL/std/stdpar/numeric does.
#pragma once

#include <cstddef>
#include <execution>
#include <functional>
#include <iterator>
#include <memory>
#include <numeric>
#include <type_traits>

#include "acpp/algorithms/numeric.hpp"

namespace acpp {
namespace stdpar {

using par_unseq_policy = std::execution::parallel_unsequenced_policy;
inline constexpr const par_unseq_policy &par_unseq = std::execution::par_unseq;

// Process-wide override of the learned offload choice.
enum class offload_policy : int {
  automatic = ACPP_STDPAR_OFFLOAD_AUTO,
  always = ACPP_STDPAR_OFFLOAD_ALWAYS,
  never = ACPP_STDPAR_OFFLOAD_NEVER,
};

inline void set_offload_policy(offload_policy p) {
  acpp::detail::check_status(acpp_stdpar_set_offload_policy(static_cast<int>(p)));
}

namespace detail {

// Thread-local queue and scratch caches (sycl_glue.hpp:54-90).
struct tls_runtime {
  acpp::queue q{0};
  std::size_t offloaded = 0;  // calls of this thread that ran on the GPU
  algorithms::util::allocation_cache device_cache{algorithms::util::allocation_type::device};
  algorithms::util::allocation_cache host_cache{algorithms::util::allocation_type::host};

  static tls_runtime &get() {
    thread_local tls_runtime rt;
    return rt;
  }
};

template <class It>
const void *address_of(It it) {
  return static_cast<const void *>(std::addressof(*it));
}

// Whether a call of operation `op` over n elements of T touching the memory
// of `its` runs on the GPU.  n > 0.
template <class T, class... Its>
bool offload(int op, std::size_t n, Its... its) {
  const void *ptrs[] = {address_of(its)...};
  int dec = 1;
  acpp::detail::check_status(acpp_stdpar_should_offload(
      tls_runtime::get().q.get_device(), op, static_cast<int>(sizeof(T)), std::is_floating_point<T>::value ? 1 : 0,
      static_cast<uint64_t>(n), ptrs, static_cast<uint32_t>(sizeof...(Its)), &dec));
  if (dec) ++tls_runtime::get().offloaded;
  return dec != 0;
}

// The host side runs once the thread's offloaded work has completed.
inline void before_host() { tls_runtime::get().q.wait(); }

template <class It>
using value_t = typename std::iterator_traits<It>::value_type;

}  // namespace detail

// Waits for this thread's offloaded work (the scans return without waiting)
// and throws its asynchronous errors.
inline void sync() { detail::tls_runtime::get().q.wait(); }

// Number of this thread's calls that ran on the GPU so far.
inline std::size_t offloaded_calls() { return detail::tls_runtime::get().offloaded; }

// ---------------------------------------------------------------- reduce

template <class ForwardIt1, class ForwardIt2, class T>
T transform_reduce(par_unseq_policy, ForwardIt1 first1, ForwardIt1 last1, ForwardIt2 first2, T init) {
  if (first1 == last1) return init;
  const std::size_t n = static_cast<std::size_t>(std::distance(first1, last1));
  if (!detail::offload<T>(ACPP_STDPAR_OP_BINARY_TRANSFORM_REDUCE, n, first1, first2)) {
    detail::before_host();
    return std::transform_reduce(first1, last1, first2, init);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group out_group{&rt.host_cache, rt.q};
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  T *out = out_group.template obtain<T>(1);
  algorithms::transform_reduce(rt.q, scratch, first1, last1, first2, out, init);
  rt.q.wait();
  return *out;
}

template <class ForwardIt1, class ForwardIt2, class T, class BinaryReductionOp, class BinaryTransformOp>
T transform_reduce(par_unseq_policy, ForwardIt1 first1, ForwardIt1 last1, ForwardIt2 first2, T init,
                   BinaryReductionOp reduce, BinaryTransformOp transform) {
  if (first1 == last1) return init;
  const std::size_t n = static_cast<std::size_t>(std::distance(first1, last1));
  if (!detail::offload<T>(ACPP_STDPAR_OP_BINARY_TRANSFORM_REDUCE, n, first1, first2)) {
    detail::before_host();
    return std::transform_reduce(first1, last1, first2, init, reduce, transform);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group out_group{&rt.host_cache, rt.q};
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  T *out = out_group.template obtain<T>(1);
  algorithms::transform_reduce(rt.q, scratch, first1, last1, first2, out, init, reduce, transform);
  rt.q.wait();
  return *out;
}

template <class ForwardIt, class T, class BinaryReductionOp, class UnaryTransformOp>
T transform_reduce(par_unseq_policy, ForwardIt first, ForwardIt last, T init, BinaryReductionOp reduce,
                   UnaryTransformOp transform) {
  if (first == last) return init;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<T>(ACPP_STDPAR_OP_TRANSFORM_REDUCE, n, first)) {
    detail::before_host();
    return std::transform_reduce(first, last, init, reduce, transform);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group out_group{&rt.host_cache, rt.q};
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  T *out = out_group.template obtain<T>(1);
  algorithms::transform_reduce(rt.q, scratch, first, last, out, init, reduce, transform);
  rt.q.wait();
  return *out;
}

template <class ForwardIt, class T, class BinaryOp>
T reduce(par_unseq_policy, ForwardIt first, ForwardIt last, T init, BinaryOp op) {
  if (first == last) return init;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<T>(ACPP_STDPAR_OP_REDUCE, n, first)) {
    detail::before_host();
    return std::reduce(first, last, init, op);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group out_group{&rt.host_cache, rt.q};
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  T *out = out_group.template obtain<T>(1);
  algorithms::reduce(rt.q, scratch, first, last, out, init, op);
  rt.q.wait();
  return *out;
}

template <class ForwardIt, class T>
T reduce(par_unseq_policy p, ForwardIt first, ForwardIt last, T init) {
  return ::acpp::stdpar::reduce(p, first, last, init, std::plus<T>{});
}

template <class ForwardIt>
detail::value_t<ForwardIt> reduce(par_unseq_policy p, ForwardIt first, ForwardIt last) {
  using T = detail::value_t<ForwardIt>;
  return ::acpp::stdpar::reduce(p, first, last, T{}, std::plus<T>{});
}

// ------------------------------------------------------------------ scans
// Offloaded scans return once their work is enqueued; see sync().

template <class InputIt, class OutputIt, class BinaryOp>
OutputIt inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first, BinaryOp op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<detail::value_t<InputIt>>(ACPP_STDPAR_OP_INCLUSIVE_SCAN, n, first, d_first)) {
    detail::before_host();
    return std::inclusive_scan(first, last, d_first, op);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::inclusive_scan(rt.q, scratch, first, last, d_first, op);
  return result;
}

template <class InputIt, class OutputIt, class BinaryOp, class T>
OutputIt inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first, BinaryOp op,
                        T init) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<detail::value_t<InputIt>>(ACPP_STDPAR_OP_INCLUSIVE_SCAN, n, first, d_first)) {
    detail::before_host();
    return std::inclusive_scan(first, last, d_first, op, init);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::inclusive_scan(rt.q, scratch, first, last, d_first, op, init);
  return result;
}

template <class InputIt, class OutputIt>
OutputIt inclusive_scan(par_unseq_policy p, InputIt first, InputIt last, OutputIt d_first) {
  return ::acpp::stdpar::inclusive_scan(p, first, last, d_first, std::plus<>{});
}

template <class InputIt, class OutputIt, class T, class BinaryOp>
OutputIt exclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first, T init,
                        BinaryOp op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<detail::value_t<InputIt>>(ACPP_STDPAR_OP_EXCLUSIVE_SCAN, n, first, d_first)) {
    detail::before_host();
    return std::exclusive_scan(first, last, d_first, init, op);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::exclusive_scan(rt.q, scratch, first, last, d_first, init, op);
  return result;
}

template <class InputIt, class OutputIt, class T>
OutputIt exclusive_scan(par_unseq_policy p, InputIt first, InputIt last, OutputIt d_first, T init) {
  return ::acpp::stdpar::exclusive_scan(p, first, last, d_first, init, std::plus<>{});
}

// pstl-impl/numeric.hpp:471-580.
template <class InputIt, class OutputIt, class BinaryOp, class UnaryOp>
OutputIt transform_inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first,
                                  BinaryOp binary_op, UnaryOp unary_op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<detail::value_t<InputIt>>(ACPP_STDPAR_OP_INCLUSIVE_SCAN, n, first, d_first)) {
    detail::before_host();
    return std::transform_inclusive_scan(first, last, d_first, binary_op, unary_op);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::transform_inclusive_scan(rt.q, scratch, first, last, d_first, binary_op, unary_op);
  return result;
}

template <class InputIt, class OutputIt, class BinaryOp, class UnaryOp, class T>
OutputIt transform_inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first,
                                  BinaryOp binary_op, UnaryOp unary_op, T init) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<detail::value_t<InputIt>>(ACPP_STDPAR_OP_INCLUSIVE_SCAN, n, first, d_first)) {
    detail::before_host();
    return std::transform_inclusive_scan(first, last, d_first, binary_op, unary_op, init);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::transform_inclusive_scan(rt.q, scratch, first, last, d_first, binary_op, unary_op, init);
  return result;
}

template <class InputIt, class OutputIt, class T, class BinaryOp, class UnaryOp>
OutputIt transform_exclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first,
                                  T init, BinaryOp binary_op, UnaryOp unary_op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  const std::size_t n = static_cast<std::size_t>(std::distance(first, last));
  if (!detail::offload<detail::value_t<InputIt>>(ACPP_STDPAR_OP_EXCLUSIVE_SCAN, n, first, d_first)) {
    detail::before_host();
    return std::transform_exclusive_scan(first, last, d_first, init, binary_op, unary_op);
  }
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::transform_exclusive_scan(rt.q, scratch, first, last, d_first, init, binary_op, unary_op);
  return result;
}

}  // namespace stdpar
}  // namespace acpp
