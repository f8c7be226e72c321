// C-ABI entry points: reduce / sycl::reduction / dot for unsigned integers.
// Generated declarations live in include/acpp_mi355x.h.
#include "reduce_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::reduce_detail;

extern "C" {

acpp_status acpp_reduce_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, uint8_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<uint8_t, uint8_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_u8");
}
acpp_status acpp_reduction_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint8_t, uint8_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, uint8_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_u8");
}

acpp_status acpp_reduce_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, uint16_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<uint16_t, uint16_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_u16");
}
acpp_status acpp_reduction_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint16_t, uint16_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, uint16_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_u16");
}

acpp_status acpp_reduce_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, uint32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<uint32_t, uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_u32");
}
acpp_status acpp_reduction_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint32_t, uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, uint32_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_u32");
}

acpp_status acpp_reduce_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, uint64_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<uint64_t, uint64_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_u64");
}
acpp_status acpp_reduction_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint64_t, uint64_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, uint64_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_u64");
}

acpp_status acpp_transform_reduce_dot_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first1, uint64_t n, const uint32_t *first2, uint32_t *out, uint32_t init, const acpp_launch_config *config) {
  return dot_entry<uint32_t, uint32_t>(scratch, static_cast<hipStream_t>(stream), first1, n, first2, out, init, config, "acpp_transform_reduce_dot_u32");
}

acpp_status acpp_transform_reduce_dot_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first1, uint64_t n, const uint64_t *first2, uint64_t *out, uint64_t init, const acpp_launch_config *config) {
  return dot_entry<uint64_t, uint64_t>(scratch, static_cast<hipStream_t>(stream), first1, n, first2, out, init, config, "acpp_transform_reduce_dot_u64");
}

acpp_status acpp_reduction_identity_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, uint8_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint8_t, uint8_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_u8");
}
acpp_status acpp_reduction_identity_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, uint16_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint16_t, uint16_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_u16");
}
acpp_status acpp_reduction_identity_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, uint32_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint32_t, uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_u32");
}
acpp_status acpp_reduction_identity_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, uint64_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<uint64_t, uint64_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_u64");
}

}  // extern "C"
