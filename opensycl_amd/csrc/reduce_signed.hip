// C-ABI entry points: reduce / sycl::reduction / dot for signed integers.
// Generated declarations live in include/acpp_mi355x.h.
#include "reduce_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::reduce_detail;

extern "C" {

acpp_status acpp_reduce_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, int8_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<int8_t, int8_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_i8");
}
acpp_status acpp_reduction_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int8_t, int8_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, int8_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_i8");
}

acpp_status acpp_reduce_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, int16_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<int16_t, int16_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_i16");
}
acpp_status acpp_reduction_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int16_t, int16_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, int16_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_i16");
}

acpp_status acpp_reduce_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<int32_t, int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_i32");
}
acpp_status acpp_reduction_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int32_t, int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, int32_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_i32");
}

acpp_status acpp_reduce_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, int64_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<int64_t, int64_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_i64");
}
acpp_status acpp_reduction_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int64_t, int64_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, int64_t(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_i64");
}

acpp_status acpp_transform_reduce_dot_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first1, uint64_t n, const int32_t *first2, int32_t *out, int32_t init, const acpp_launch_config *config) {
  return dot_entry<int32_t, int32_t>(scratch, static_cast<hipStream_t>(stream), first1, n, first2, out, init, config, "acpp_transform_reduce_dot_i32");
}

acpp_status acpp_transform_reduce_dot_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first1, uint64_t n, const int64_t *first2, int64_t *out, int64_t init, const acpp_launch_config *config) {
  return dot_entry<int64_t, int64_t>(scratch, static_cast<hipStream_t>(stream), first1, n, first2, out, init, config, "acpp_transform_reduce_dot_i64");
}

acpp_status acpp_reduction_identity_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, int8_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int8_t, int8_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_i8");
}
acpp_status acpp_reduction_identity_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, int16_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int16_t, int16_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_i16");
}
acpp_status acpp_reduction_identity_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, int32_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int32_t, int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_i32");
}
acpp_status acpp_reduction_identity_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, int64_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<int64_t, int64_t>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_i64");
}

}  // extern "C"
