// C-ABI entry points: inclusive_scan / exclusive_scan for 64-bit integers.
#include "scan_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::scan_detail;

extern "C" {

acpp_status acpp_inclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, acpp_algorithm_op op, const int64_t *init, const acpp_launch_config *config) {
  return scan_entry<int64_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, int64_t(0), config, "acpp_inclusive_scan_i64");
}
acpp_status acpp_exclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, int64_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<int64_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_i64");
}

acpp_status acpp_inclusive_scan_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, acpp_algorithm_op op, const uint64_t *init, const acpp_launch_config *config) {
  return scan_entry<uint64_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, uint64_t(0), config, "acpp_inclusive_scan_u64");
}
acpp_status acpp_exclusive_scan_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, uint64_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<uint64_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_u64");
}

}  // extern "C"
