/*
 * oracle.h -- CPU restatement of the reference's algorithms for the
 * device-wide primitives.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker for the HIP kernels (tests, smoke check, CPU baseline timing) and
 * is never linked into, loaded by or called from the product library.
 *
 * Two families of functions:
 *   oracle_<op>_<t>            serial semantics that the reference's own tests
 *                              assert against (std::accumulate /
 *                              std::inclusive_scan / serial atomic folds);
 *   oracle_*_omp_order_<t>     the order in which the reference's OpenMP host
 *                              backend evaluates the same call (multi-threaded
 *                              with OpenMP); used for floating-point order
 *                              studies and as the timed CPU baseline.
 * Element type tags: i8 i16 i32 i64 u8 u16 u32 u64 f32 f64.
 * Operator codes equal __acpp_sscp_algorithm_op
 * (include/hipSYCL/sycl/libkernel/sscp/builtins/builtin_config.hpp:27-37).
 * Every function returns 0 on success and -1 on an invalid operator.
 */
#ifndef ACPP_ORACLE_H
#define ACPP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_DECLARE(t, T)                                                                      \
  int oracle_identity_##t(int op, T *out);                                                        \
  int oracle_apply_##t(int op, T a, T b, T *out);                                                 \
  int oracle_reduce_##t(const T *x, uint64_t n, T init, int op, T *out);                          \
  int oracle_reduction_##t(const T *x, uint64_t n, T *out, int op, int initialize_to_identity);   \
  int oracle_inclusive_scan_##t(const T *x, uint64_t n, T *y, int op, const T *init);             \
  int oracle_exclusive_scan_##t(const T *x, uint64_t n, T *y, T init, int op);                    \
  int oracle_dot_##t(const T *a, const T *b, uint64_t n, T init, T *out);                         \
  int oracle_reduce_omp_order_##t(const T *x, uint64_t n, T init, int op, T *out,                 \
                                  int num_threads);                                               \
  int oracle_dot_omp_order_##t(const T *a, const T *b, uint64_t n, T init, T *out,                \
                               int num_threads);                                                  \
  int oracle_scan_omp_order_##t(const T *x, uint64_t n, T *y, int exclusive, int has_init,        \
                                T init, int op, int num_threads);                                 \
  int oracle_group_reduce_##t(const T *x, uint64_t n, uint64_t local_size, int op, T *y);         \
  int oracle_group_inclusive_scan_##t(const T *x, uint64_t n, uint64_t local_size, int op, T *y); \
  int oracle_group_exclusive_scan_##t(const T *x, uint64_t n, uint64_t local_size, T init,        \
                                      int op, T *y);                                              \
  int oracle_atomic_fold_##t(int op, const T *operands, uint64_t n, T initial, T *final_value);   \
  int oracle_segmented_inclusive_scan_##t(const T *x, const uint64_t *offsets,                    \
                                          uint64_t num_segments, T *y, int op);                   \
  int oracle_segmented_reduce_##t(const T *x, const uint64_t *offsets, uint64_t num_segments,     \
                                  T init, int op, T *out);

ORACLE_DECLARE(i8, int8_t)
ORACLE_DECLARE(i16, int16_t)
ORACLE_DECLARE(i32, int32_t)
ORACLE_DECLARE(i64, int64_t)
ORACLE_DECLARE(u8, uint8_t)
ORACLE_DECLARE(u16, uint16_t)
ORACLE_DECLARE(u32, uint32_t)
ORACLE_DECLARE(u64, uint64_t)
ORACLE_DECLARE(f32, float)
ORACLE_DECLARE(f64, double)

/* FP64 oracles for floating-point sums: Neumaier-compensated FP64
   accumulation (error ~ n * 2^-106 relative to the magnitude sum).
   bf16 inputs are raw uint16 bit patterns; widening to float is exact. */
double oracle_sum_f32_fp64(const float *x, uint64_t n);
double oracle_sum_bf16_fp64(const uint16_t *x, uint64_t n);
double oracle_dot_f32_fp64(const float *a, const float *b, uint64_t n);
double oracle_dot_bf16_fp64(const uint16_t *a, const uint16_t *b, uint64_t n);
double oracle_abs_sum_f32_fp64(const float *x, uint64_t n);
int oracle_segmented_reduce_f32_fp64(const float *x, const uint64_t *offsets,
                                     uint64_t num_segments, double *out);

/* Histogram by fetch_add(1): bins[keys[i]] += 1 for keys[i] < num_bins. */
int oracle_histogram_u32(const uint32_t *keys, uint64_t n, uint32_t *bins, uint32_t num_bins,
                         int num_threads);
/* slots[i % num_slots] = max(slots[i % num_slots], values[i]). */
int oracle_fetch_max_strided_i32(const int32_t *values, uint64_t n, int32_t *slots,
                                 uint32_t num_slots, int num_threads);
/* copy_if: out[0..count) = the x[i] (in order) for which the predicate
   holds (codes: 0 x != 0, 1 ==, 2 !=, 3 <, 4 <=, 5 >, 6 >=, 7 (x & mask) ==
   operand, 8 stencil[i] != 0). */
int oracle_copy_if_i32(const int32_t *x, uint64_t n, const uint8_t *stencil, int pred, int32_t operand, int32_t mask, int32_t *out, uint64_t *count);
int oracle_copy_if_u32(const uint32_t *x, uint64_t n, const uint8_t *stencil, int pred, uint32_t operand, uint32_t mask, uint32_t *out, uint64_t *count);
int oracle_copy_if_i64(const int64_t *x, uint64_t n, const uint8_t *stencil, int pred, int64_t operand, int64_t mask, int64_t *out, uint64_t *count);
int oracle_copy_if_u64(const uint64_t *x, uint64_t n, const uint8_t *stencil, int pred, uint64_t operand, uint64_t mask, uint64_t *out, uint64_t *count);
int oracle_copy_if_f32(const float *x, uint64_t n, const uint8_t *stencil, int pred, float operand, float mask, float *out, uint64_t *count);
int oracle_copy_if_f64(const double *x, uint64_t n, const uint8_t *stencil, int pred, double operand, double mask, double *out, uint64_t *count);

/* Number of OpenMP threads a parallel region would use by default. */
int oracle_max_threads(void);
/* omp_get_num_procs(): the processors available to this process, the value
   the reference reports as the host device's max_compute_units
   (src/runtime/omp/omp_hardware_manager.cpp:124-128). */
int oracle_num_procs(void);

/* Deliberate-mistake switch for the tests that show the oracle's pinning
   checks catch plausible errors.  0 = faithful (default).  1 dropped last
   element, 2 inclusive for exclusive, 3 ignored init, 4 short group, 5 lost
   histogram key, 6 short copy_if count, 7 abs(x) = x; in the OpenMP-order
   restatements: 8 sixteen elements per work-item, 9 serial left fold, 10
   adjacent-pairs tree, 11 swapped tree operands, 12 1024 x 8 scan tiles, 13
   swapped work-item fold operands, 14 op(result, init); 15 |x-y| computed
   as x-y in the binary transform_reduce; 16 shift_group_right by param - 1. */
void oracle_set_fault(int fault);

/* Element transforms (transform codes 0..7 of acpp_unary_op); -1 for an
   unknown code or bit_not on floating point. */
int oracle_transform_i8(const int8_t *x, uint64_t n, int8_t *y, int code);
int oracle_transform_i16(const int16_t *x, uint64_t n, int16_t *y, int code);
int oracle_transform_i32(const int32_t *x, uint64_t n, int32_t *y, int code);
int oracle_transform_i64(const int64_t *x, uint64_t n, int64_t *y, int code);
int oracle_transform_u8(const uint8_t *x, uint64_t n, uint8_t *y, int code);
int oracle_transform_u16(const uint16_t *x, uint64_t n, uint16_t *y, int code);
int oracle_transform_u32(const uint32_t *x, uint64_t n, uint32_t *y, int code);
int oracle_transform_u64(const uint64_t *x, uint64_t n, uint64_t *y, int code);
int oracle_transform_f32(const float *x, uint64_t n, float *y, int code);
int oracle_transform_f64(const double *x, uint64_t n, double *y, int code);

/* Binary transform_reduce: *out = op(...op(op(init, t(a0, b0)), t(a1, b1))...)
   for pair transform codes 0..9 (acpp_binary_op); -1 for an unknown code or
   operator, or a bitwise code on floating point. */
int oracle_transform_reduce_binary_i32(const int32_t *a, const int32_t *b, uint64_t n, int32_t init, int op, int code, int32_t *out);
int oracle_transform_reduce_binary_i64(const int64_t *a, const int64_t *b, uint64_t n, int64_t init, int op, int code, int64_t *out);
int oracle_transform_reduce_binary_u32(const uint32_t *a, const uint32_t *b, uint64_t n, uint32_t init, int op, int code, uint32_t *out);
int oracle_transform_reduce_binary_u64(const uint64_t *a, const uint64_t *b, uint64_t n, uint64_t init, int op, int code, uint64_t *out);
int oracle_transform_reduce_binary_f32(const float *a, const float *b, uint64_t n, float init, int op, int code, float *out);
int oracle_transform_reduce_binary_f64(const double *a, const double *b, uint64_t n, double init, int op, int code, double *out);

/* Group data movement (broadcast, shifts, permute by xor, select, any /
   all / none) within sub-groups of 64 (level 0) or work-groups (level 1) of
   local_size elements; see oracle.c for the operation codes. */
int oracle_group_move_i32(int level, int which, int64_t param, const int32_t *x, const uint32_t *src, uint64_t n, uint64_t local_size, int32_t *out);
int oracle_group_move_i64(int level, int which, int64_t param, const int64_t *x, const uint32_t *src, uint64_t n, uint64_t local_size, int64_t *out);
int oracle_group_move_f32(int level, int which, int64_t param, const float *x, const uint32_t *src, uint64_t n, uint64_t local_size, float *out);
int oracle_group_move_f64(int level, int which, int64_t param, const double *x, const uint32_t *src, uint64_t n, uint64_t local_size, double *out);

#ifdef __cplusplus
}
#endif

#endif /* ACPP_ORACLE_H */
