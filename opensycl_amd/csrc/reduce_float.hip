// C-ABI entry points: reduce / sycl::reduction / dot for floating point, bf16 and the compensated forms.
// Generated declarations live in include/acpp_mi355x.h.
#include "reduce_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::reduce_detail;

extern "C" {

acpp_status acpp_reduce_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<float, float>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_f32");
}
acpp_status acpp_reduction_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<float, float>(scratch, static_cast<hipStream_t>(stream), first, n, out, float(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_f32");
}

acpp_status acpp_transform_reduce_dot_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first1, uint64_t n, const float *first2, float *out, float init, const acpp_launch_config *config) {
  return dot_entry<float, float>(scratch, static_cast<hipStream_t>(stream), first1, n, first2, out, init, config, "acpp_transform_reduce_dot_f32");
}

acpp_status acpp_reduce_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, double init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<double, double>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op, final_with_init, config, "acpp_reduce_f64");
}
acpp_status acpp_reduction_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<double, double>(scratch, static_cast<hipStream_t>(stream), first, n, out, double(0), op, initialize_to_identity ? final_identity_init : final_accumulate, config, "acpp_reduction_f64");
}

acpp_status acpp_transform_reduce_dot_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first1, uint64_t n, const double *first2, double *out, double init, const acpp_launch_config *config) {
  return dot_entry<double, double>(scratch, static_cast<hipStream_t>(stream), first1, n, first2, out, init, config, "acpp_transform_reduce_dot_f64");
}

acpp_status acpp_reduce_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, float *out, float init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return reduce_entry<bf16_t, float>(scratch, static_cast<hipStream_t>(stream), reinterpret_cast<const bf16_t *>(first), n, out, init, op, final_with_init, config, "acpp_reduce_bf16");
}
acpp_status acpp_transform_reduce_dot_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first1, uint64_t n, const uint16_t *first2, float *out, float init, const acpp_launch_config *config) {
  return dot_entry<bf16_t, float>(scratch, static_cast<hipStream_t>(stream), reinterpret_cast<const bf16_t *>(first1), n, reinterpret_cast<const bf16_t *>(first2), out, init, config, "acpp_transform_reduce_dot_bf16");
}
acpp_status acpp_reduce_compensated_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float init, const acpp_launch_config *config) {
  return compensated_entry<float>(scratch, static_cast<hipStream_t>(stream), first, nullptr, false, n, out, init, config, "acpp_reduce_compensated_f32");
}
acpp_status acpp_reduce_compensated_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, float *out, float init, const acpp_launch_config *config) {
  return compensated_entry<bf16_t>(scratch, static_cast<hipStream_t>(stream), reinterpret_cast<const bf16_t *>(first), nullptr, false, n, out, init, config, "acpp_reduce_compensated_bf16");
}
acpp_status acpp_transform_reduce_dot_compensated_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first1, uint64_t n, const float *first2, float *out, float init, const acpp_launch_config *config) {
  return compensated_entry<float>(scratch, static_cast<hipStream_t>(stream), first1, first2, true, n, out, init, config, "acpp_transform_reduce_dot_compensated_f32");
}
acpp_status acpp_transform_reduce_dot_compensated_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first1, uint64_t n, const uint16_t *first2, float *out, float init, const acpp_launch_config *config) {
  return compensated_entry<bf16_t>(scratch, static_cast<hipStream_t>(stream), reinterpret_cast<const bf16_t *>(first1), reinterpret_cast<const bf16_t *>(first2), true, n, out, init, config, "acpp_transform_reduce_dot_compensated_bf16");
}

acpp_status acpp_reduction_identity_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<float, float>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_f32");
}
acpp_status acpp_reduction_identity_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, double identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config) {
  return reduce_entry<double, double>(scratch, static_cast<hipStream_t>(stream), first, n, out, identity, op, initialize_to_identity ? final_given_identity : final_accumulate, config, "acpp_reduction_identity_f64");
}

}  // extern "C"
