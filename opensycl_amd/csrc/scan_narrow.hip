// C-ABI entry points: inclusive_scan / exclusive_scan for 8- and 16-bit integers.
#include "scan_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::scan_detail;

extern "C" {

acpp_status acpp_inclusive_scan_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, acpp_algorithm_op op, const int8_t *init, const acpp_launch_config *config) {
  return scan_entry<int8_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, int8_t(0), config, "acpp_inclusive_scan_i8");
}
acpp_status acpp_exclusive_scan_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, int8_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<int8_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_i8");
}

acpp_status acpp_inclusive_scan_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, acpp_algorithm_op op, const int16_t *init, const acpp_launch_config *config) {
  return scan_entry<int16_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, int16_t(0), config, "acpp_inclusive_scan_i16");
}
acpp_status acpp_exclusive_scan_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, int16_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<int16_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_i16");
}

acpp_status acpp_inclusive_scan_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, acpp_algorithm_op op, const uint8_t *init, const acpp_launch_config *config) {
  return scan_entry<uint8_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, uint8_t(0), config, "acpp_inclusive_scan_u8");
}
acpp_status acpp_exclusive_scan_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, uint8_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<uint8_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_u8");
}

acpp_status acpp_inclusive_scan_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, acpp_algorithm_op op, const uint16_t *init, const acpp_launch_config *config) {
  return scan_entry<uint16_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, uint16_t(0), config, "acpp_inclusive_scan_u16");
}
acpp_status acpp_exclusive_scan_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, uint16_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<uint16_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_u16");
}

}  // extern "C"
