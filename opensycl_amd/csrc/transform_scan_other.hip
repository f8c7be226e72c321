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
ACPP_DEFINE_TRANSFORM_SCANS(i8, int8_t)
ACPP_DEFINE_TRANSFORM_SCANS(i16, int16_t)
ACPP_DEFINE_TRANSFORM_SCANS(u8, uint8_t)
ACPP_DEFINE_TRANSFORM_SCANS(u16, uint16_t)
ACPP_DEFINE_TRANSFORM_SCANS(f32, float)
ACPP_DEFINE_TRANSFORM_SCANS(f64, double)

}  // extern "C"
