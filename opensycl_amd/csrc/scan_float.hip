// C-ABI entry points: inclusive_scan / exclusive_scan for floating point.
#include "scan_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::scan_detail;

extern "C" {

acpp_status acpp_inclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, acpp_algorithm_op op, const float *init, const acpp_launch_config *config) {
  return scan_entry<float>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, float(0), config, "acpp_inclusive_scan_f32");
}
acpp_status acpp_exclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, float init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<float>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_f32");
}

acpp_status acpp_inclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, acpp_algorithm_op op, const double *init, const acpp_launch_config *config) {
  return scan_entry<double>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, double(0), config, "acpp_inclusive_scan_f64");
}
acpp_status acpp_exclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, double init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<double>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_f64");
}

}  // extern "C"
