/*
 * acpp_mi355x.h -- C-ABI of the MI355X-native device-wide primitives.
 *
 * This is the drop-in boundary for AdaptiveCpp's data-parallel hot path:
 * the acpp::algorithms reduce / transform_reduce / inclusive_scan /
 * exclusive_scan entry points, sycl::reduction, and the atomic fetch
 * operations that stdpar offloading lowers onto.  The C++ facade in
 * include/acpp/algorithms/numeric.hpp re-declares the reference's templates
 * on top of these functions; ctypes / cgo / JNI bindings call them directly.
 *
 * Conventions shared by every entry point:
 *   - Pointers to data are device pointers (hipMalloc / hipMallocManaged
 *     memory); `n` is an element count.
 *   - `stream` is a hipStream_t passed as an opaque pointer (NULL = the null
 *     stream).  Every compute entry point is asynchronous with respect to the
 *     host: it enqueues work on `stream` and returns, like the reference,
 *     whose calls return a sycl::event (doc/algorithms.md:35-37).
 *   - `scratch` is an allocation group (allocation_cache.hpp:142-181).  The
 *     memory it hands out stays owned by the group until
 *     acpp_allocation_group_release / _destroy; the caller must keep the
 *     group alive until the enqueued work has finished, or order all users
 *     of one cache on one stream (allocation_cache.hpp:129-141).
 *   - Nothing throws.  Every function returns an acpp_status; on failure
 *     acpp_last_error_message() describes the problem (thread-local).
 *   - n == 0: reduce, transform_reduce and the scans enqueue nothing and
 *     leave `*out` untouched (numeric.hpp:165-169, decoupled_lookback_scan.hpp:565-566).
 *   - `config` may be NULL; zero fields mean "library's choice".
 */
#ifndef ACPP_MI355X_H
#define ACPP_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
typedef int acpp_status;
#define ACPP_SUCCESS 0
#define ACPP_ERROR_INVALID_ARGUMENT 1 /* null pointer, bad enum, size limit */
#define ACPP_ERROR_UNSUPPORTED 2      /* op not defined for this element type */
#define ACPP_ERROR_OUT_OF_MEMORY 3
#define ACPP_ERROR_HIP 4              /* HIP runtime reported an error */
#define ACPP_ERROR_WATCHDOG 5         /* a kernel gave up waiting: see      */
                                      /* acpp_check_asynchronous_errors     */

/* Message for the last failing call on this thread ("" if none). */
const char *acpp_last_error_message(void);

/* Library version as "major.minor.patch". */
const char *acpp_version(void);

/* ------------------------------------------------------------------------ */
/* Operators.  Values equal __acpp_sscp_algorithm_op                        */
/* (include/hipSYCL/sycl/libkernel/sscp/builtins/builtin_config.hpp:27-37). */
/* Integer types accept all nine; floating-point types accept plus,         */
/* multiply, min and max.                                                   */
/* ------------------------------------------------------------------------ */
typedef int32_t acpp_algorithm_op;
#define ACPP_OP_PLUS 0
#define ACPP_OP_MULTIPLY 1
#define ACPP_OP_MIN 2
#define ACPP_OP_MAX 3
#define ACPP_OP_BIT_AND 4
#define ACPP_OP_BIT_OR 5
#define ACPP_OP_BIT_XOR 6
#define ACPP_OP_LOGICAL_AND 7
#define ACPP_OP_LOGICAL_OR 8

/* ------------------------------------------------------------------------ */
/* Element transforms of transform_reduce (unary form) and of the transform */
/* scans.  The reference compiles an arbitrary unary callable into the      */
/* kernel (numeric.hpp:193-212, 299-327); across a C-ABI the transform is   */
/* one of these codes.  The result has the element type: integer results    */
/* wrap modulo 2^bits (abs of the most negative value is itself).           */
/* ACPP_UNARY_BIT_NOT exists only for integer element types.                */
/* ------------------------------------------------------------------------ */
typedef int32_t acpp_unary_op;
#define ACPP_UNARY_IDENTITY 0   /* x */
#define ACPP_UNARY_NEGATE 1     /* -x (std::negate) */
#define ACPP_UNARY_ABS 2        /* |x| */
#define ACPP_UNARY_SQUARE 3     /* x * x */
#define ACPP_UNARY_TWICE 4      /* x + x */
#define ACPP_UNARY_BIT_NOT 5    /* ~x (std::bit_not) */
#define ACPP_UNARY_IS_ZERO 6    /* x == 0 ? 1 : 0 */
#define ACPP_UNARY_IS_NONZERO 7 /* x != 0 ? 1 : 0 */

/* Pair transforms of the binary transform_reduce (numeric.hpp:170-191:
   *out = reduce(init, transform(x[i], y[i]) ...)), in the element type;
   integer results wrap.  The bitwise codes exist only for integer types. */
typedef int32_t acpp_binary_op;
#define ACPP_BINARY_MULTIPLIES 0 /* x * y (with plus: the dot product) */
#define ACPP_BINARY_PLUS 1       /* x + y */
#define ACPP_BINARY_MINUS 2      /* x - y */
#define ACPP_BINARY_ABS_DIFF 3   /* |x - y| */
#define ACPP_BINARY_SQ_DIFF 4    /* (x - y) * (x - y) */
#define ACPP_BINARY_MIN 5        /* x < y ? x : y */
#define ACPP_BINARY_MAX 6        /* x > y ? x : y */
#define ACPP_BINARY_BIT_AND 7    /* x & y */
#define ACPP_BINARY_BIT_OR 8     /* x | y */
#define ACPP_BINARY_BIT_XOR 9    /* x ^ y */

typedef void *acpp_stream_t; /* hipStream_t */

/* ------------------------------------------------------------------------ */
/* Launch configuration.  0 in a field = library's choice (autotuned table  */
/* or learned cost model).  Replaces the reference's hard-coded local size  */
/* 128 and groups = 4 x compute units (numeric.hpp:144-156,                  */
/* memory_streaming.hpp:33-46) and its 16 items per work-item scan tiles    */
/* (decoupled_lookback_scan.hpp:481-488, scan.hpp:26-31).                   */
/* ------------------------------------------------------------------------ */
typedef struct acpp_launch_config {
  uint32_t work_group_size;  /* 64..1024, multiple of 64 */
  uint32_t items_per_thread; /* reduce: 16-byte loads in flight per thread
                                (2, 4, 8, or 16 except for dot products);
                                scans: 16-byte rows per thread (2, 4 or 8
                                for 4- and 8-byte plus; 8 otherwise, except
                                1-byte elements: 8 or 4, and 4 for the
                                logical operators) */
  uint32_t groups_per_cu;    /* reduce and segmented: resident work-groups
                                per CU (cap on the grid) */
  uint32_t flags;            /* variant bits, 0 = default:
                                ACPP_VARIANT_BLOCK_LOOKBACK (scans) looks back
                                with the whole work-group (window = group
                                size) instead of one wave (window 64);
                                ACPP_VARIANT_NONTEMPORAL (reduce) streams the
                                input with non-temporal loads, (scans of
                                4- and 8-byte elements with operator plus)
                                full tiles with non-temporal loads and
                                stores;
                                ACPP_VARIANT_TILE_PER_GROUP (segmented) one
                                work-group per tile instead of a persistent
                                grid */
} acpp_launch_config;

#define ACPP_VARIANT_BLOCK_LOOKBACK 1u
#define ACPP_VARIANT_NONTEMPORAL 2u
#define ACPP_VARIANT_TILE_PER_GROUP 4u
/* For tests of the error path only: every wait for another work-group
   (look-backs, a reduction's wait for partials) gives up at once, as if
   its watchdog had expired.  Results are wrong and the give-up is reported
   as ACPP_ERROR_WATCHDOG; every kernel still ends normally. */
#define ACPP_VARIANT_FORCE_GIVE_UP 8u

/* ------------------------------------------------------------------------ */
/* Scratch memory: allocation_cache / allocation_group                      */
/* (include/hipSYCL/algorithms/util/allocation_cache.hpp:42-181).           */
/* ------------------------------------------------------------------------ */
typedef struct acpp_allocation_cache acpp_allocation_cache;
typedef struct acpp_allocation_group acpp_allocation_group;

#define ACPP_ALLOCATION_DEVICE 0 /* hipMalloc */
#define ACPP_ALLOCATION_SHARED 1 /* hipMallocManaged */
#define ACPP_ALLOCATION_HOST 2   /* hipHostMalloc */

/* allocation_cache(allocation_type) -- allocation_cache.hpp:44-45 */
acpp_status acpp_allocation_cache_create(int allocation_type,
                                         acpp_allocation_cache **out);
/* allocation_cache::purge -- allocation_cache.hpp:51-62.  Frees every
   allocation currently returned to the cache. */
acpp_status acpp_allocation_cache_purge(acpp_allocation_cache *cache);
/* ~allocation_cache -- allocation_cache.hpp:47-49 */
acpp_status acpp_allocation_cache_destroy(acpp_allocation_cache *cache);
/* allocation_group(cache, device) -- allocation_cache.hpp:144-146 */
acpp_status acpp_allocation_group_create(acpp_allocation_cache *cache,
                                         int device,
                                         acpp_allocation_group **out);
/* allocation_group::obtain<T>(count) -- allocation_cache.hpp:169-175 */
acpp_status acpp_allocation_group_obtain(acpp_allocation_group *group,
                                         size_t bytes, size_t alignment,
                                         void **out);
/* allocation_group::release -- allocation_cache.hpp:162-167 */
acpp_status acpp_allocation_group_release(acpp_allocation_group *group);
/* ~allocation_group -- allocation_cache.hpp:158-160 */
acpp_status acpp_allocation_group_destroy(acpp_allocation_group *group);
/* No counterpart in the reference (its look-back spins without bound).
   Look-backs of the scans, copy_if and segmented scans, and the wait of a
   reduction's last work-group for the other groups' partials, give up
   after ~2.7 s instead of hanging the device; such a give-up makes that
   call's results wrong.  *trips = the number of give-ups recorded in the
   scratch regions the group holds (since each region was last reset).
   Blocking; call once the group's work has completed.  Expected to be 0. */
acpp_status acpp_allocation_group_watchdog_trips(acpp_allocation_group *group, uint64_t *trips);
/* Asynchronous errors of `device`, the counterpart of the reference's
   queue::throw_asynchronous / wait_and_throw (include/hipSYCL/sycl/
   queue.hpp:378-383; src/runtime/error.cpp:77-90).  Every give-up above
   also raises a per-device flag in pinned host memory; this call reads and
   clears it without touching the device, and returns ACPP_ERROR_WATCHDOG
   if it was raised (ACPP_SUCCESS otherwise).  Call it after synchronising
   with the work in question (the C++ facade's queue/event wait, the stdpar
   calls and the Python Queue.wait do).  acpp_allocation_group_release and
   _destroy report a raised flag of the group's device the same way (after
   releasing) but leave it raised: only this call clears it. */
acpp_status acpp_check_asynchronous_errors(int device);

/* ------------------------------------------------------------------------ */
/* Standard-parallelism offload decision: whether a std::reduce /            */
/* transform_reduce / *_scan(par_unseq) call runs on the GPU or on the host. */
/* Replaces include/hipSYCL/std/stdpar/detail/offload.hpp:384-594 and the    */
/* sampled-time interpolation of offload_heuristic_db.hpp:247-318: memory    */
/* only the device can read forces the GPU, memory only the host can read    */
/* forces the host; pinned host and managed memory are decided by the        */
/* learned cost model (the MLP of the acpp_mlp_* kernels, trained on run     */
/* times measured on MI355X), evaluated once per device into a table of      */
/* crossover sizes.                                                          */
/* ------------------------------------------------------------------------ */
#define ACPP_STDPAR_OP_REDUCE 0
#define ACPP_STDPAR_OP_TRANSFORM_REDUCE 1 /* unary transform */
#define ACPP_STDPAR_OP_BINARY_TRANSFORM_REDUCE 2
#define ACPP_STDPAR_OP_INCLUSIVE_SCAN 3   /* also transform_inclusive_scan */
#define ACPP_STDPAR_OP_EXCLUSIVE_SCAN 4   /* also transform_exclusive_scan */
#define ACPP_STDPAR_OFFLOAD_AUTO 0   /* the decision above (default) */
#define ACPP_STDPAR_OFFLOAD_ALWAYS 1 /* every call that may offload does */
#define ACPP_STDPAR_OFFLOAD_NEVER 2  /* every call that may stay on the host does */
/* *offload = 1 (GPU) or 0 (host) for a call of operation `op` over n
   elements of `elem_bytes` bytes touching the `num_pointers` given
   pointers.  Invalid argument if the pointers mix device-only and
   host-only memory. */
acpp_status acpp_stdpar_should_offload(int device, int op, int elem_bytes, int is_float, uint64_t n,
                                       const void *const *pointers, uint32_t num_pointers, int *offload);
/* Process-wide override of the learned choice (ACPP_STDPAR_OFFLOAD_*). */
acpp_status acpp_stdpar_set_offload_policy(int policy);
/* *min_n = the smallest n the model offloads (for memory both sides can
   read); UINT64_MAX when it never does, 0 when no model is compiled in. */
acpp_status acpp_stdpar_offload_threshold(int device, int op, int elem_bytes, int is_float, uint64_t *min_n);

/* Number of compute units of `device` (the value the reference reads as
   info::device::max_compute_units, numeric.hpp:154-156). */
acpp_status acpp_device_compute_units(int device, int *out);

/* ------------------------------------------------------------------------ */
/* reduce: *out = op(init, x[0] op x[1] op ... op x[n-1])                    */
/* Replaces algorithms::reduce (numeric.hpp:226-249), i.e.                   */
/* transform_reduce with an identity transform (numeric.hpp:188-202).       */
/* One launch: grid-stride 128-bit loads, DPP wave reduction, LDS combine,  */
/* and the last work-group to finish folds the per-group partials.          */
/* ------------------------------------------------------------------------ */
acpp_status acpp_reduce_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, int8_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, int16_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, int64_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, uint8_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, uint16_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, uint32_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, uint64_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_reduce_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, double init, acpp_algorithm_op op, const acpp_launch_config *config);
/* bf16 input (raw uint16 bit patterns), fp32 accumulation and result.  No
   counterpart in the reference (it has no bf16 type); same contract. */
acpp_status acpp_reduce_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, float *out, float init, acpp_algorithm_op op, const acpp_launch_config *config);

/* Compensated sums (op is always plus): fp32 accumulation with TwoSum error
   terms carried per lane and per wave, FP64 only in the final cross-group
   combine.  Deterministic: fixed reduction tree, no floating-point atomics.
   Precision extension of algorithms::reduce (numeric.hpp:226-249). */
acpp_status acpp_reduce_compensated_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float init, const acpp_launch_config *config);
acpp_status acpp_reduce_compensated_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, float *out, float init, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* transform_reduce, binary form with the default operators                 */
/* (numeric.hpp:204-223): *out = init + sum_i first1[i] * first2[i].        */
/* ------------------------------------------------------------------------ */
acpp_status acpp_transform_reduce_dot_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first1, uint64_t n, const int32_t *first2, int32_t *out, int32_t init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first1, uint64_t n, const int64_t *first2, int64_t *out, int64_t init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first1, uint64_t n, const uint32_t *first2, uint32_t *out, uint32_t init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first1, uint64_t n, const uint64_t *first2, uint64_t *out, uint64_t init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first1, uint64_t n, const float *first2, float *out, float init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first1, uint64_t n, const double *first2, double *out, double init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first1, uint64_t n, const uint16_t *first2, float *out, float init, const acpp_launch_config *config);
/* Compensated dot products (Dot2: TwoProduct + TwoSum per element, FP64 in
   the final cross-group combine). */
acpp_status acpp_transform_reduce_dot_compensated_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first1, uint64_t n, const float *first2, float *out, float init, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_dot_compensated_bf16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first1, uint64_t n, const uint16_t *first2, float *out, float init, const acpp_launch_config *config);

/* Unary transform_reduce (numeric.hpp:193-212):                           */
/*   *out = op(init, op over transform(x[i])); n == 0 leaves *out as is.    */
acpp_status acpp_transform_reduce_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, int8_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, int16_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, int32_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, int64_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, uint8_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, uint16_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, uint32_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, uint64_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, double init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
/* Binary transform_reduce (numeric.hpp:170-191) with any reduction of
   plus, multiplies, minimum, maximum and a pair transform code:
   *out = op(init, op over transform(first1[i], first2[i])); an empty range
   leaves *out untouched (numeric.hpp:165-169).  Other reduction operators
   give ACPP_ERROR_UNSUPPORTED. */
acpp_status acpp_transform_reduce_binary_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first1, uint64_t n, const int32_t *first2, int32_t *out, int32_t init, acpp_algorithm_op op, acpp_binary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_binary_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first1, uint64_t n, const int64_t *first2, int64_t *out, int64_t init, acpp_algorithm_op op, acpp_binary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_binary_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first1, uint64_t n, const uint32_t *first2, uint32_t *out, uint32_t init, acpp_algorithm_op op, acpp_binary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_binary_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first1, uint64_t n, const uint64_t *first2, uint64_t *out, uint64_t init, acpp_algorithm_op op, acpp_binary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_binary_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first1, uint64_t n, const float *first2, float *out, float init, acpp_algorithm_op op, acpp_binary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_reduce_binary_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first1, uint64_t n, const double *first2, double *out, double init, acpp_algorithm_op op, acpp_binary_op transform, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* sycl::reduction over a 1-D range (handler.hpp:799-975 with              */
/* reduction.hpp:165-236 and reduction_descriptor.hpp:25-54):                */
/*   initialize_to_identity != 0 : *out = result                            */
/*   initialize_to_identity == 0 : *out = op(*out, result)                  */
/* where result is the reduction of x[0..n) (the identity when n == 0).     */
/* ------------------------------------------------------------------------ */
acpp_status acpp_reduction_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
/* sycl::reduction(T *var, const T &identity, combiner, props)
   (include/hipSYCL/sycl/libkernel/reduction.hpp:214-236;
   algorithms/reduction/reduction_descriptor.hpp:36-53): with
   initialize_to_identity *out = op(identity, fold) -- identity for an empty
   range -- otherwise *out = op(*out, fold).  SYCL requires `identity` to be
   an identity of the operator; it is the value an empty reduction yields
   (for example +0.0 instead of the library's -0.0 for a float sum). */
acpp_status acpp_reduction_identity_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *out, int8_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *out, int16_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *out, int32_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *out, int64_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *out, uint8_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *out, uint16_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *out, uint32_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *out, uint64_t identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *out, float identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_identity_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *out, double identity, acpp_algorithm_op op, int initialize_to_identity, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* Scans: single-pass decoupled look-back (decoupled_lookback_scan.hpp:     */
/* 555-655), one launch, no status-initialisation kernel.                   */
/*   inclusive_scan (numeric.hpp:253-279):                                  */
/*     d[i] = [init op] x[0] op ... op x[i];  init == NULL means no init.   */
/*   exclusive_scan (numeric.hpp:281-297):                                  */
/*     d[0] = init, d[i] = init op x[0] op ... op x[i-1].                   */
/* d_first may equal first (in-place); other overlaps are undefined.        */
/* Integer results are exact; floating-point results depend on the order   */
/* in which the look-back finds predecessors and may differ in the last    */
/* bits between runs.                                                       */
/* ------------------------------------------------------------------------ */
acpp_status acpp_inclusive_scan_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, acpp_algorithm_op op, const int8_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, acpp_algorithm_op op, const int16_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, acpp_algorithm_op op, const int32_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, acpp_algorithm_op op, const int64_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, acpp_algorithm_op op, const uint8_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, acpp_algorithm_op op, const uint16_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, acpp_algorithm_op op, const uint32_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, acpp_algorithm_op op, const uint64_t *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, acpp_algorithm_op op, const float *init, const acpp_launch_config *config);
acpp_status acpp_inclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, acpp_algorithm_op op, const double *init, const acpp_launch_config *config);

acpp_status acpp_exclusive_scan_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, int8_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, int16_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config);
/* No counterpart in the reference.  acpp_exclusive_scan_i32 for callers
   that cannot allocate (code under hipGraph capture, tight serving loops,
   harnesses that forbid hipMalloc): no allocation group; the scan's sync
   region (256-byte header + 2 MiB of tile status lines) is a static device
   array of the library, one per GPU, set up on the first call on each
   device.  At most 16384 tiles (2^29 elements with the default 1024 x 8-row
   tiles); larger inputs are rejected with ACPP_ERROR_INVALID_ARGUMENT.
   Calls on one GPU share that array: calls whose kernels overlap on one GPU
   (different streams) race and give wrong results; calls ordered on a
   stream, or otherwise serialised, are safe, graph replays of a captured
   call included (after a capture, eager calls clear the array first).
   Calls no hipMalloc/hipFree. */
acpp_status acpp_exclusive_scan_i32_noalloc(acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, int64_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, uint8_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, uint16_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, uint32_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, uint64_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, float init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_exclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, double init, acpp_algorithm_op op, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* Transform scans (numeric.hpp:299-327; scan/scan.hpp:85-132): the scans   */
/* above over transform(x[i]) instead of x[i].                              */
/*   transform_inclusive_scan: d[i] = [init op] t(x[0]) op ... op t(x[i])   */
/*   transform_exclusive_scan: d[0] = init,                                 */
/*                             d[i] = init op t(x[0]) op ... op t(x[i-1])   */
/* ------------------------------------------------------------------------ */
acpp_status acpp_transform_inclusive_scan_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const int8_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const int16_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const int32_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const int64_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const uint8_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const uint16_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const uint32_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, acpp_algorithm_op op, acpp_unary_op transform, const uint64_t *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, acpp_algorithm_op op, acpp_unary_op transform, const float *init, const acpp_launch_config *config);
acpp_status acpp_transform_inclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, acpp_algorithm_op op, acpp_unary_op transform, const double *init, const acpp_launch_config *config);

acpp_status acpp_transform_exclusive_scan_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, int8_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, int16_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int32_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, int64_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, uint8_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, uint16_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, uint32_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, uint64_t init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, float init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);
acpp_status acpp_transform_exclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, double init, acpp_algorithm_op op, acpp_unary_op transform, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* Bulk atomics.  These replace a user kernel in which work-item i applies  */
/* atomic_ref<T, relaxed, device>(target[slot(i)]).fetch_<op>(v_i)          */
/* (include/hipSYCL/sycl/libkernel/atomic_ref.hpp:143-197, HIP lowering     */
/* generic/hiplike/atomic_builtins.hpp:438-598).  The final contents of     */
/* `target` equal those of that kernel (the operations commute); the        */
/* per-element return values are not produced.                              */
/*   strategy ACPP_ATOMICS_AUTO        library's choice                     */
/*            ACPP_ATOMICS_PRIVATIZED  per-work-group copies in LDS /       */
/*                                     registers, combined at the end       */
/*            ACPP_ATOMICS_GLOBAL      one device-scope atomic per element  */
/* ------------------------------------------------------------------------ */
#define ACPP_ATOMICS_AUTO 0
#define ACPP_ATOMICS_PRIVATIZED 1
#define ACPP_ATOMICS_GLOBAL 2

/* Histogram: bins[keys[i]] += 1 for every keys[i] < num_bins (other keys are
   ignored); the counts are added to the existing bin values.  The
   privatized strategy supports up to 65536 bins. */
acpp_status acpp_histogram_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *keys, uint64_t n, uint32_t *bins, uint32_t num_bins, int strategy, const acpp_launch_config *config);

/* Strided fetch ops: target[i % num_slots] = op(target[i % num_slots], values[i]). */
acpp_status acpp_atomic_fetch_add_strided_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *values, uint64_t n, int32_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_min_strided_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *values, uint64_t n, int32_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_max_strided_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *values, uint64_t n, int32_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_add_strided_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *values, uint64_t n, uint32_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_min_strided_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *values, uint64_t n, uint32_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_max_strided_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *values, uint64_t n, uint32_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_add_strided_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *values, uint64_t n, int64_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_min_strided_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *values, uint64_t n, int64_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_max_strided_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *values, uint64_t n, int64_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_add_strided_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *values, uint64_t n, uint64_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_min_strided_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *values, uint64_t n, uint64_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_max_strided_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *values, uint64_t n, uint64_t *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_add_strided_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *values, uint64_t n, float *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_min_strided_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *values, uint64_t n, float *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_max_strided_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *values, uint64_t n, float *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_add_strided_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *values, uint64_t n, double *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_min_strided_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *values, uint64_t n, double *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);
acpp_status acpp_atomic_fetch_max_strided_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *values, uint64_t n, double *target, uint32_t num_slots, int strategy, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* Segmented scans and reductions: independent scans / reductions over      */
/* segments [offsets[s], offsets[s+1]) of one array, in one launch.  No     */
/* counterpart in the reference's algorithms library; the semantics are one */
/* inclusive_scan (numeric.hpp:253-279, no init) or reduce                  */
/* (numeric.hpp:226-249) per segment, i.e. one joint_inclusive_scan /       */
/* joint_reduce (generic/hiplike/group_functions.hpp:214-287, 361-432) per  */
/* segment.                                                                 */
/*   offsets: num_segments + 1 device values, non-decreasing, offsets[0] = 0,*/
/*            offsets[num_segments] = n.  Malformed offsets give wrong      */
/*            results but no out-of-bounds access.                          */
/*   reduce:  out[s] = op(init, x[offsets[s]] op ... op x[offsets[s+1]-1]); */
/*            empty segments get init.                                      */
/*   op:      plus, multiply, min or max.                                   */
/* ------------------------------------------------------------------------ */
acpp_status acpp_segmented_inclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_inclusive_scan_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_inclusive_scan_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_inclusive_scan_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_reduce_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, int32_t *out, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_reduce_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, int64_t *out, int64_t init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_reduce_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, float *out, float init, acpp_algorithm_op op, const acpp_launch_config *config);
acpp_status acpp_segmented_reduce_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, double *out, double init, acpp_algorithm_op op, const acpp_launch_config *config);

/* hipGraph plans: the same work captured once (setup kernel + persistent
   kernel, scratch from `scratch`) and replayed with one graph launch.  The
   plan references `scratch`, the input, output and offsets buffers: keep
   them alive (and the group unreleased) while the plan is used. */
typedef struct acpp_segmented_plan acpp_segmented_plan;
acpp_status acpp_segmented_plan_inclusive_scan_i32(acpp_allocation_group *scratch, const int32_t *first, uint64_t n, int32_t *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_inclusive_scan_i64(acpp_allocation_group *scratch, const int64_t *first, uint64_t n, int64_t *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_inclusive_scan_f32(acpp_allocation_group *scratch, const float *first, uint64_t n, float *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_inclusive_scan_f64(acpp_allocation_group *scratch, const double *first, uint64_t n, double *d_first, const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_reduce_i32(acpp_allocation_group *scratch, const int32_t *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, int32_t *out, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_reduce_i64(acpp_allocation_group *scratch, const int64_t *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, int64_t *out, int64_t init, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_reduce_f32(acpp_allocation_group *scratch, const float *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, float *out, float init, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_reduce_f64(acpp_allocation_group *scratch, const double *first, uint64_t n, const uint64_t *offsets, uint64_t num_segments, double *out, double init, acpp_algorithm_op op, const acpp_launch_config *config, acpp_segmented_plan **plan);
acpp_status acpp_segmented_plan_launch(acpp_segmented_plan *plan, acpp_stream_t stream);
acpp_status acpp_segmented_plan_destroy(acpp_segmented_plan *plan);

/* ------------------------------------------------------------------------ */
/* Several reductions in one kernel: parallel_for with up to four           */
/* sycl::reduction objects (include/hipSYCL/sycl/handler.hpp:799-975;        */
/* tests/sycl/reduction.cpp:116-208).  outs[k] = fold of inputs[k] with      */
/* ops[k]; with initialize_to_identity the fold starts from the identity,  */
/* otherwise from *outs[k] (reduction_descriptor.hpp:36-53).  `inputs`,     */
/* `outs` and `ops` are host arrays of `count` (1..4) entries; the pointers */
/* in them are device pointers; inputs may alias.  One launch.              */
/* ------------------------------------------------------------------------ */
acpp_status acpp_reduction_multi_i8(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int8_t *const *inputs, uint64_t n, int8_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_u8(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint8_t *const *inputs, uint64_t n, uint8_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_i16(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int16_t *const *inputs, uint64_t n, int16_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_u16(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint16_t *const *inputs, uint64_t n, uint16_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_i32(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int32_t *const *inputs, uint64_t n, int32_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_u32(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint32_t *const *inputs, uint64_t n, uint32_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_i64(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int64_t *const *inputs, uint64_t n, int64_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_u64(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint64_t *const *inputs, uint64_t n, uint64_t *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_f32(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const float *const *inputs, uint64_t n, float *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_f64(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const double *const *inputs, uint64_t n, double *const *outs, const acpp_algorithm_op *ops, int initialize_to_identity, const acpp_launch_config *config);
/* The same with the properties of each sycl::reduction object of the
   parallel_for: initialize_to_identity[k] (host array of `count` flags)
   and, when `identities` is not NULL, a given identity per reduction
   (reduction.hpp:214-236): with the flag *outs[k] = op(identities[k],
   fold), which an empty range leaves as identities[k]. */
acpp_status acpp_reduction_multi_identity_i8(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int8_t *const *inputs, uint64_t n, int8_t *const *outs, const acpp_algorithm_op *ops, const int8_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_u8(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint8_t *const *inputs, uint64_t n, uint8_t *const *outs, const acpp_algorithm_op *ops, const uint8_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_i16(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int16_t *const *inputs, uint64_t n, int16_t *const *outs, const acpp_algorithm_op *ops, const int16_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_u16(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint16_t *const *inputs, uint64_t n, uint16_t *const *outs, const acpp_algorithm_op *ops, const uint16_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_i32(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int32_t *const *inputs, uint64_t n, int32_t *const *outs, const acpp_algorithm_op *ops, const int32_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_u32(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint32_t *const *inputs, uint64_t n, uint32_t *const *outs, const acpp_algorithm_op *ops, const uint32_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_i64(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const int64_t *const *inputs, uint64_t n, int64_t *const *outs, const acpp_algorithm_op *ops, const int64_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_u64(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const uint64_t *const *inputs, uint64_t n, uint64_t *const *outs, const acpp_algorithm_op *ops, const uint64_t *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_f32(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const float *const *inputs, uint64_t n, float *const *outs, const acpp_algorithm_op *ops, const float *identities, const int *initialize_to_identity, const acpp_launch_config *config);
acpp_status acpp_reduction_multi_identity_f64(acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const double *const *inputs, uint64_t n, double *const *outs, const acpp_algorithm_op *ops, const double *identities, const int *initialize_to_identity, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* copy_if: stream compaction                                               */
/* Replaces algorithms::copy_if (include/hipSYCL/algorithms/algorithm.hpp:  */
/* 178-243): the elements x[i] for which the predicate holds are written,   */
/* in order, to d_first[0..count); *num_copied = count (device pointer, may */
/* be NULL; 0 for an empty range, algorithm.hpp:184-188).  One launch:      */
/* predicate bits, DPP ranks, decoupled look-back over the tile counts.     */
/* The predicate is given as data (a C functor cannot cross this ABI):      */
/* ------------------------------------------------------------------------ */
#define ACPP_PRED_NONZERO 0 /* x != 0 */
#define ACPP_PRED_EQ 1      /* x == operand */
#define ACPP_PRED_NE 2      /* x != operand */
#define ACPP_PRED_LT 3      /* x <  operand */
#define ACPP_PRED_LE 4      /* x <= operand */
#define ACPP_PRED_GT 5      /* x >  operand */
#define ACPP_PRED_GE 6      /* x >= operand */
#define ACPP_PRED_MASK_EQ 7 /* (x & mask) == operand, integers only */
#define ACPP_PRED_STENCIL 8 /* stencil[i] != 0 (the _stencil entry points) */
acpp_status acpp_copy_if_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, int8_t *d_first, int pred, int8_t operand, int8_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, uint8_t *d_first, int pred, uint8_t operand, uint8_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, int16_t *d_first, int pred, int16_t operand, int16_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, uint16_t *d_first, int pred, uint16_t operand, uint16_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int pred, int32_t operand, int32_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, int pred, uint32_t operand, uint32_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, int64_t *d_first, int pred, int64_t operand, int64_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, uint64_t *d_first, int pred, uint64_t operand, uint64_t mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, float *d_first, int pred, float operand, float mask, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, double *d_first, int pred, double operand, double mask, uint64_t *num_copied, const acpp_launch_config *config);
/* Predicate evaluated by the caller: copies x[i] where stencil[i] != 0. */
acpp_status acpp_copy_if_stencil_i8(acpp_allocation_group *scratch, acpp_stream_t stream, const int8_t *first, uint64_t n, const uint8_t *stencil, int8_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_u8(acpp_allocation_group *scratch, acpp_stream_t stream, const uint8_t *first, uint64_t n, const uint8_t *stencil, uint8_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_i16(acpp_allocation_group *scratch, acpp_stream_t stream, const int16_t *first, uint64_t n, const uint8_t *stencil, int16_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_u16(acpp_allocation_group *scratch, acpp_stream_t stream, const uint16_t *first, uint64_t n, const uint8_t *stencil, uint16_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, const uint8_t *stencil, int32_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, const uint8_t *stencil, uint32_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_i64(acpp_allocation_group *scratch, acpp_stream_t stream, const int64_t *first, uint64_t n, const uint8_t *stencil, int64_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_u64(acpp_allocation_group *scratch, acpp_stream_t stream, const uint64_t *first, uint64_t n, const uint8_t *stencil, uint64_t *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_f32(acpp_allocation_group *scratch, acpp_stream_t stream, const float *first, uint64_t n, const uint8_t *stencil, float *d_first, uint64_t *num_copied, const acpp_launch_config *config);
acpp_status acpp_copy_if_stencil_f64(acpp_allocation_group *scratch, acpp_stream_t stream, const double *first, uint64_t n, const uint8_t *stencil, double *d_first, uint64_t *num_copied, const acpp_launch_config *config);

/* ------------------------------------------------------------------------ */
/* Learned cost model (no counterpart in the reference, whose launch        */
/* geometry is fixed: numeric.hpp:144-156, decoupled_lookback_scan.hpp:     */
/* 481-488).  A 16 -> 64 -> 64 -> 1 tanh network that predicts the          */
/* standardised log run time of a launch configuration on a problem shape. */
/* Parameters: one float array, layout W1[64][16], b1[64], W2[64][64],      */
/* b2[64], w3[64], b3.  Inputs x: [batch][16] row-major.  All pointers are  */
/* device pointers; batch in [1, 2^20].                                     */
/* ------------------------------------------------------------------------ */
#define ACPP_MLP_INPUTS 16
#define ACPP_MLP_HIDDEN 64
#define ACPP_MLP_PARAMS 5313

/* y[b] = network(x[b]) */
acpp_status acpp_mlp_forward(acpp_stream_t stream, const float *params, const float *x, uint32_t batch,
                             float *y);
/* grad = dL/dparams of L = mean_b (y[b] - target[b])^2; *loss = L if loss
   is not NULL.  Deterministic: the same inputs give the same bits. */
acpp_status acpp_mlp_gradient(acpp_allocation_group *scratch, acpp_stream_t stream, const float *params,
                              const float *x, const float *target, uint32_t batch, float *grad,
                              float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, L2 weight decay added to
   the gradient) on the full batch; `step` counts from 1 for bias
   correction; adam_m / adam_v hold the moment estimates (zero before step
   1).  *loss = L before the update if loss is not NULL. */
acpp_status acpp_mlp_adam_step(acpp_allocation_group *scratch, acpp_stream_t stream, float *params,
                               float *adam_m, float *adam_v, uint32_t step, const float *x,
                               const float *target, uint32_t batch, float lr, float weight_decay,
                               float *loss);

/* Configuration query for the segmented forms with a trained model.
   acpp_cost_model_create copies the parameters (host memory) to `device`
   with the model's input standardisation (16 means / deviations) and target
   standardisation (y_mean, y_std of the natural log of microseconds).
   acpp_cost_model_choose_segmented computes the 16 features of the shape
   (offsets in HOST memory, num_segments + 1 entries) for each of the 15
   candidate configurations (work-group 64..1024 x rows 2/4/8), evaluates
   the network with the forward kernel on `stream`, waits for it, and fills
   `config` with the candidate of lowest predicted time (*predicted_us, if
   not NULL).  opensycl_amd/data/cost_model.json holds the model trained on
   MI355X timings. */
typedef struct acpp_cost_model acpp_cost_model;
#define ACPP_ELEM_I32 0
#define ACPP_ELEM_I64 1
#define ACPP_ELEM_F32 2
#define ACPP_ELEM_F64 3
#define ACPP_SEGMENTED_SCAN 0
#define ACPP_SEGMENTED_REDUCE 1
acpp_status acpp_cost_model_create(int device, const float *params, const double *x_mean, const double *x_std,
                                   double y_mean, double y_std, acpp_cost_model **out);
acpp_status acpp_cost_model_destroy(acpp_cost_model *model);
acpp_status acpp_cost_model_choose_segmented(acpp_cost_model *model, acpp_stream_t stream, const uint64_t *offsets,
                                             uint64_t num_segments, int elem, int kind, acpp_algorithm_op op,
                                             acpp_launch_config *config, float *predicted_us);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* ACPP_MI355X_H */
