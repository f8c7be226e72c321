// C-ABI entry points: unary transform_reduce
// (include/hipSYCL/algorithms/numeric.hpp:193-212) for every element type,
// and binary transform_reduce with a pair transform (numeric.hpp:170-191).
#include "reduce_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::reduce_detail;

extern "C" {

#define ACPP_DEFINE_TRANSFORM_REDUCE(tag, T)                                                              \
  acpp_status acpp_transform_reduce_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,          \
                                          const T *first, uint64_t n, T *out, T init,                    \
                                          acpp_algorithm_op op, acpp_unary_op transform,                 \
                                          const acpp_launch_config *config) {                            \
    return transform_reduce_entry<T>(scratch, static_cast<hipStream_t>(stream), first, n, out, init, op,  \
                                     transform, config, "acpp_transform_reduce_" #tag);                  \
  }
ACPP_DEFINE_TRANSFORM_REDUCE(i8, int8_t)
ACPP_DEFINE_TRANSFORM_REDUCE(i16, int16_t)
ACPP_DEFINE_TRANSFORM_REDUCE(i32, int32_t)
ACPP_DEFINE_TRANSFORM_REDUCE(i64, int64_t)
ACPP_DEFINE_TRANSFORM_REDUCE(u8, uint8_t)
ACPP_DEFINE_TRANSFORM_REDUCE(u16, uint16_t)
ACPP_DEFINE_TRANSFORM_REDUCE(u32, uint32_t)
ACPP_DEFINE_TRANSFORM_REDUCE(u64, uint64_t)
ACPP_DEFINE_TRANSFORM_REDUCE(f32, float)
ACPP_DEFINE_TRANSFORM_REDUCE(f64, double)

#define ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(tag, T)                                                      \
  acpp_status acpp_transform_reduce_binary_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,     \
                                                 const T *first1, uint64_t n, const T *first2, T *out,    \
                                                 T init, acpp_algorithm_op op, acpp_binary_op transform,  \
                                                 const acpp_launch_config *config) {                      \
    return binary_transform_reduce_entry<T>(scratch, static_cast<hipStream_t>(stream), first1, n, first2,   \
                                            out, init, op, transform, config,                              \
                                            "acpp_transform_reduce_binary_" #tag);                         \
  }
ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(i32, int32_t)
ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(i64, int64_t)
ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(u32, uint32_t)
ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(u64, uint64_t)
ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(f32, float)
ACPP_DEFINE_BINARY_TRANSFORM_REDUCE(f64, double)

}  // extern "C"
