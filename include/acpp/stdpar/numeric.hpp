// Standard-parallelism entry points for the numeric algorithms: reduce,
// transform_reduce, inclusive_scan, exclusive_scan, transform_inclusive_scan
// and transform_exclusive_scan with a par_unseq policy, offloaded to the MI355X kernels through the acpp::algorithms
// facade.
//
// Same entry points and offload structure as the reference's stdpar layer
// (include/hipSYCL/std/stdpar/pstl-impl/numeric.hpp:28-580): a thread-local
// in-order queue and thread-local scratch caches
// (include/hipSYCL/std/stdpar/detail/sycl_glue.hpp:54-90), device scratch
// for the algorithm and host scratch for a reduction's result, and the
// reduction returns once its result is on the host (the reference's
// HIPSYCL_STDPAR_BLOCKING_OFFLOAD).
//
// Two differences follow from having no compiler support here:
//   - every call, scans included, completes before it returns.  The
//     reference returns from a scan without waiting and relies on
//     synchronisation points its compiler inserts before the host touches
//     the data;
//   - there is no host fallback and no offload heuristic: the data must be
//     device-accessible (hipMalloc, hipMallocManaged or hipHostMalloc
//     memory; the reference gets there by replacing the allocator), and
//     every call runs on the GPU.
#pragma once

#include <cstddef>
#include <functional>
#include <iterator>
#include <memory>

#include "acpp/algorithms/numeric.hpp"

namespace acpp {
namespace stdpar {

struct par_unseq_policy {};
inline constexpr par_unseq_policy par_unseq{};

namespace detail {

// Thread-local queue and scratch caches (sycl_glue.hpp:54-90).
struct tls_runtime {
  acpp::queue q{0};
  algorithms::util::allocation_cache device_cache{algorithms::util::allocation_type::device};
  algorithms::util::allocation_cache host_cache{algorithms::util::allocation_type::host};

  static tls_runtime &get() {
    thread_local tls_runtime rt;
    return rt;
  }
};

}  // namespace detail

// ---------------------------------------------------------------- reduce

template <class ForwardIt1, class ForwardIt2, class T>
T transform_reduce(par_unseq_policy, ForwardIt1 first1, ForwardIt1 last1, ForwardIt2 first2, T init) {
  if (first1 == last1) return init;
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
  return reduce(p, first, last, init, std::plus<T>{});
}

template <class ForwardIt>
typename std::iterator_traits<ForwardIt>::value_type reduce(par_unseq_policy p, ForwardIt first,
                                                            ForwardIt last) {
  using T = typename std::iterator_traits<ForwardIt>::value_type;
  return reduce(p, first, last, T{}, std::plus<T>{});
}

// ------------------------------------------------------------------ scans

template <class InputIt, class OutputIt, class BinaryOp>
OutputIt inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first, BinaryOp op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::inclusive_scan(rt.q, scratch, first, last, d_first, op);
  rt.q.wait();
  return result;
}

template <class InputIt, class OutputIt, class BinaryOp, class T>
OutputIt inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first, BinaryOp op,
                        T init) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::inclusive_scan(rt.q, scratch, first, last, d_first, op, init);
  rt.q.wait();
  return result;
}

template <class InputIt, class OutputIt>
OutputIt inclusive_scan(par_unseq_policy p, InputIt first, InputIt last, OutputIt d_first) {
  return inclusive_scan(p, first, last, d_first, std::plus<>{});
}

template <class InputIt, class OutputIt, class T, class BinaryOp>
OutputIt exclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first, T init,
                        BinaryOp op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::exclusive_scan(rt.q, scratch, first, last, d_first, init, op);
  rt.q.wait();
  return result;
}

template <class InputIt, class OutputIt, class T>
OutputIt exclusive_scan(par_unseq_policy p, InputIt first, InputIt last, OutputIt d_first, T init) {
  return exclusive_scan(p, first, last, d_first, init, std::plus<>{});
}

// pstl-impl/numeric.hpp:471-580.
template <class InputIt, class OutputIt, class BinaryOp, class UnaryOp>
OutputIt transform_inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first,
                                  BinaryOp binary_op, UnaryOp unary_op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::transform_inclusive_scan(rt.q, scratch, first, last, d_first, binary_op, unary_op);
  rt.q.wait();
  return result;
}

template <class InputIt, class OutputIt, class BinaryOp, class UnaryOp, class T>
OutputIt transform_inclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first,
                                  BinaryOp binary_op, UnaryOp unary_op, T init) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::transform_inclusive_scan(rt.q, scratch, first, last, d_first, binary_op, unary_op, init);
  rt.q.wait();
  return result;
}

template <class InputIt, class OutputIt, class T, class BinaryOp, class UnaryOp>
OutputIt transform_exclusive_scan(par_unseq_policy, InputIt first, InputIt last, OutputIt d_first,
                                  T init, BinaryOp binary_op, UnaryOp unary_op) {
  OutputIt result = d_first;
  std::advance(result, std::distance(first, last));
  if (first == last) return result;
  auto &rt = detail::tls_runtime::get();
  algorithms::util::allocation_group scratch{&rt.device_cache, rt.q};
  algorithms::transform_exclusive_scan(rt.q, scratch, first, last, d_first, init, binary_op, unary_op);
  rt.q.wait();
  return result;
}

}  // namespace stdpar
}  // namespace acpp
