// C-ABI entry points: transform_inclusive_scan / transform_exclusive_scan
// (include/hipSYCL/algorithms/numeric.hpp:299-327; scan/scan.hpp:85-132).
#include "scan_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::scan_detail;

extern "C" {

#define ACPP_DEFINE_TRANSFORM_SCANS(tag, T)                                                                \
  acpp_status acpp_transform_inclusive_scan_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,   \
                                                  const T *first, uint64_t n, T *d_first,                 \
                                                  acpp_algorithm_op op, acpp_unary_op transform,          \
                                                  const T *init, const acpp_launch_config *config) {      \
    return transform_scan_entry<T>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op,       \
                                   transform, false, init, T(0), config,                                   \
                                   "acpp_transform_inclusive_scan_" #tag);                                 \
  }                                                                                                        \
  acpp_status acpp_transform_exclusive_scan_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,   \
                                                  const T *first, uint64_t n, T *d_first, T init,         \
                                                  acpp_algorithm_op op, acpp_unary_op transform,          \
                                                  const acpp_launch_config *config) {                     \
    return transform_scan_entry<T>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op,       \
                                   transform, true, nullptr, init, config,                                 \
                                   "acpp_transform_exclusive_scan_" #tag);                                 \
  }
ACPP_DEFINE_TRANSFORM_SCANS(i32, int32_t)
ACPP_DEFINE_TRANSFORM_SCANS(i64, int64_t)
ACPP_DEFINE_TRANSFORM_SCANS(u32, uint32_t)
ACPP_DEFINE_TRANSFORM_SCANS(u64, uint64_t)

}  // extern "C"
