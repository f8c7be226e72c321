// C-ABI entry points: inclusive_scan / exclusive_scan for 32-bit integers.
#include "scan_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::scan_detail;

extern "C" {

acpp_status acpp_inclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, acpp_algorithm_op op, const int32_t *init, const acpp_launch_config *config) {
  return scan_entry<int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, int32_t(0), config, "acpp_inclusive_scan_i32");
}
acpp_status acpp_exclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_i32");
}

acpp_status acpp_inclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, acpp_algorithm_op op, const uint32_t *init, const acpp_launch_config *config) {
  return scan_entry<uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, uint32_t(0), config, "acpp_inclusive_scan_u32");
}
acpp_status acpp_exclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, uint32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_u32");
}

}  // extern "C"
