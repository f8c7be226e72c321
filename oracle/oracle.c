/*
 * oracle.c -- CPU restatement of the reference's device-wide primitives.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reference citations are relative
 * to the AdaptiveCpp source tree.
 *
 * Integer arithmetic is done in unsigned types so that wrap-around is
 * defined; the reference computes in T and relies on two's-complement wrap.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

static int g_fault = 0;
void oracle_set_fault(int fault) { g_fault = fault; }

int oracle_max_threads(void) { return omp_get_max_threads(); }
int oracle_num_procs(void) { return omp_get_num_procs(); }

enum { OP_PLUS = 0, OP_MUL, OP_MIN, OP_MAX, OP_BAND, OP_BOR, OP_BXOR, OP_LAND, OP_LOR };
/* rmw operation numbering of the atomic builtin ABI
   (src/compiler/sscp/StdAtomicRemapperPass.cpp:54). */
enum { RMW_AND = 0, RMW_OR, RMW_XOR, RMW_ADD, RMW_SUB, RMW_MIN, RMW_MAX };

/* ------------------------------------------------------------------------- */
/* Operators and identities                                                  */
/* ------------------------------------------------------------------------- */
/* apply(op, a, b): sycl::plus / multiplies / minimum / maximum / bit_* /
   logical_* (include/hipSYCL/sycl/libkernel/functional.hpp).  min/max follow
   sycl::minimum (x < y ? x : y) and sycl::maximum (x > y ? x : y). */
#define DEF_INT_OPS(t, T, WU)                                                       \
  static inline int valid_op_##t(int op) { return op >= OP_PLUS && op <= OP_LOR; }  \
  static inline T sub_##t(T a, T b) { return (T)((WU)a - (WU)b); }                  \
  static inline T ap_##t(int op, T a, T b) {                                        \
    switch (op) {                                                                   \
      case OP_PLUS: return (T)((WU)a + (WU)b);                                      \
      case OP_MUL: return (T)((WU)a * (WU)b);                                       \
      case OP_MIN: return (a < b) ? a : b;                                          \
      case OP_MAX: return (a > b) ? a : b;                                          \
      case OP_BAND: return (T)(a & b);                                              \
      case OP_BOR: return (T)(a | b);                                               \
      case OP_BXOR: return (T)(a ^ b);                                              \
      case OP_LAND: return (T)(a && b);                                             \
      default: return (T)(a || b);                                                  \
    }                                                                               \
  }

#define DEF_FP_OPS(t, T)                                                             \
  static inline int valid_op_##t(int op) { return op >= OP_PLUS && op <= OP_MAX; }   \
  static inline T sub_##t(T a, T b) { return a - b; }                                \
  static inline T ap_##t(int op, T a, T b) {                                         \
    switch (op) {                                                                    \
      case OP_PLUS: return a + b;                                                    \
      case OP_MUL: return a * b;                                                     \
      case OP_MIN: return (a < b) ? a : b;                                           \
      default: return (a > b) ? a : b;                                               \
    }                                                                                \
  }

DEF_INT_OPS(i8, int8_t, uint32_t)
DEF_INT_OPS(i16, int16_t, uint32_t)
DEF_INT_OPS(i32, int32_t, uint32_t)
DEF_INT_OPS(i64, int64_t, uint64_t)
DEF_INT_OPS(u8, uint8_t, uint32_t)
DEF_INT_OPS(u16, uint16_t, uint32_t)
DEF_INT_OPS(u32, uint32_t, uint32_t)
DEF_INT_OPS(u64, uint64_t, uint64_t)
DEF_FP_OPS(f32, float)
DEF_FP_OPS(f64, double)

/* True identities (x op id == x for every x): the semantics the serial
   oracle and the GPU kernels use. */
#define DEF_INT_ID(t, T, TMIN, TMAX)                                  \
  static inline T id_##t(int op) {                                    \
    switch (op) {                                                     \
      case OP_MUL: return (T)1;                                       \
      case OP_MIN: return (T)(TMAX);                                  \
      case OP_MAX: return (T)(TMIN);                                  \
      case OP_BAND: return (T)~(T)0;                                  \
      case OP_LAND: return (T)1;                                      \
      default: return (T)0;                                           \
    }                                                                 \
  }                                                                   \
  /* Identity the reference's algorithms layer pads with            \
     (numeric.hpp:35-53): for integers maximum -> numeric_limits<T>  \
     ::min() is the lowest value, so it coincides with id_##t. */     \
  static inline T refid_##t(int op) { return id_##t(op); }

DEF_INT_ID(i8, int8_t, INT8_MIN, INT8_MAX)
DEF_INT_ID(i16, int16_t, INT16_MIN, INT16_MAX)
DEF_INT_ID(i32, int32_t, INT32_MIN, INT32_MAX)
DEF_INT_ID(i64, int64_t, INT64_MIN, INT64_MAX)
DEF_INT_ID(u8, uint8_t, 0, UINT8_MAX)
DEF_INT_ID(u16, uint16_t, 0, UINT16_MAX)
DEF_INT_ID(u32, uint32_t, 0, UINT32_MAX)
DEF_INT_ID(u64, uint64_t, 0, UINT64_MAX)

static inline float id_f32(int op) {
  switch (op) {
    case OP_PLUS: return -0.0f;
    case OP_MUL: return 1.0f;
    case OP_MIN: return INFINITY;
    default: return -INFINITY;
  }
}
static inline double id_f64(int op) {
  switch (op) {
    case OP_PLUS: return -0.0;
    case OP_MUL: return 1.0;
    case OP_MIN: return INFINITY;
    default: return -INFINITY;
  }
}
/* The reference's known identities for floating point (numeric.hpp:44-53):
   plus -> T{} (+0.0), multiplies -> 1, minimum -> numeric_limits<T>::max(),
   maximum -> numeric_limits<T>::min() (the smallest positive normal). */
static inline float refid_f32(int op) {
  switch (op) {
    case OP_PLUS: return 0.0f;
    case OP_MUL: return 1.0f;
    case OP_MIN: return 3.40282347e+38f;
    default: return 1.17549435e-38f;
  }
}
static inline double refid_f64(int op) {
  switch (op) {
    case OP_PLUS: return 0.0;
    case OP_MUL: return 1.0;
    case OP_MIN: return 1.7976931348623157e+308;
    default: return 2.2250738585072014e-308;
  }
}

/* ------------------------------------------------------------------------- */
/* Per-type functions                                                        */
/* ------------------------------------------------------------------------- */
#define ORACLE_DEFINE(t, T, IS_INT)                                                                 \
  int oracle_identity_##t(int op, T *out) {                                                         \
    if (!valid_op_##t(op)) return -1;                                                               \
    *out = id_##t(op);                                                                              \
    return 0;                                                                                       \
  }                                                                                                 \
  int oracle_apply_##t(int op, T a, T b, T *out) {                                                  \
    if (!valid_op_##t(op)) return -1;                                                               \
    *out = ap_##t(op, a, b);                                                                        \
    return 0;                                                                                       \
  }                                                                                                 \
  /* std::accumulate(first, last, init, op): tests/sycl/reduction.cpp:35-36,                    \
     tests/pstl/reduce.cpp:29-30. */                                                                \
  int oracle_reduce_##t(const T *x, uint64_t n, T init, int op, T *out) {                           \
    if (!valid_op_##t(op)) return -1;                                                               \
    T acc = (g_fault == 3) ? id_##t(op) : init;                                                     \
    for (uint64_t i = 0; i < n; ++i) acc = ap_##t(op, acc, x[i]);                                   \
    *out = acc;                                                                                     \
    return 0;                                                                                       \
  }                                                                                                 \
  /* sycl::reduction: with initialize_to_identity the old value is discarded,                     \
     otherwise *out = op(*out, result) (reduction_descriptor.hpp:46-53,                           \
     reduction.hpp:210-236; tests/sycl/reduction.cpp:323-344). */                                   \
  int oracle_reduction_##t(const T *x, uint64_t n, T *out, int op, int initialize_to_identity) {    \
    if (!valid_op_##t(op)) return -1;                                                               \
    T acc = id_##t(op);                                                                             \
    for (uint64_t i = 0; i < n; ++i) acc = ap_##t(op, acc, x[i]);                                   \
    *out = initialize_to_identity ? acc : ap_##t(op, *out, acc);                                    \
    return 0;                                                                                       \
  }                                                                                                 \
  /* std::inclusive_scan with optional init (tests/pstl/inclusive_scan.cpp:77-88);                \
     with init, element 0 is op(init, x0) (decoupled_lookback_scan.hpp:264-271). */                 \
  int oracle_inclusive_scan_##t(const T *x, uint64_t n, T *y, int op, const T *init) {              \
    if (!valid_op_##t(op)) return -1;                                                               \
    if (n == 0) return 0;                                                                           \
    T acc = init ? ap_##t(op, *init, x[0]) : x[0];                                                  \
    y[0] = acc;                                                                                     \
    uint64_t end = (g_fault == 1) ? n - 1 : n;                                                      \
    for (uint64_t i = 1; i < end; ++i) {                                                            \
      acc = ap_##t(op, acc, x[i]);                                                                  \
      y[i] = acc;                                                                                   \
    }                                                                                               \
    if (g_fault == 1 && n > 1) y[n - 1] = acc;                                                      \
    return 0;                                                                                       \
  }                                                                                                 \
  /* std::exclusive_scan (tests/pstl/exclusive_scan.cpp:77-102):                                  \
     y[0] = init, y[i] = init op x[0] op ... op x[i-1]                                             \
     (decoupled_lookback_scan.hpp:272-276). */                                                      \
  int oracle_exclusive_scan_##t(const T *x, uint64_t n, T *y, T init, int op) {                     \
    if (!valid_op_##t(op)) return -1;                                                               \
    T acc = init;                                                                                   \
    for (uint64_t i = 0; i < n; ++i) {                                                              \
      T xi = x[i];                                                                                  \
      if (g_fault == 2) acc = ap_##t(op, acc, xi);                                                  \
      y[i] = acc;                                                                                   \
      if (g_fault != 2) acc = ap_##t(op, acc, xi);                                                  \
    }                                                                                               \
    return 0;                                                                                       \
  }                                                                                                 \
  /* transform_reduce with the default plus / multiplies (numeric.hpp:204-223). */                  \
  int oracle_dot_##t(const T *a, const T *b, uint64_t n, T init, T *out) {                          \
    T acc = init;                                                                                   \
    for (uint64_t i = 0; i < n; ++i) acc = ap_##t(OP_PLUS, acc, ap_##t(OP_MUL, a[i], b[i]));        \
    *out = acc;                                                                                     \
    return 0;                                                                                       \
  }                                                                                                 \
  /* The reference's reduction on its OpenMP backend:                                              \
     - transform_reduce_impl / wg_model_reduction use local size 128                               \
       (numeric.hpp:139-161);                                                                       \
     - data_streamer on a host device: ceil(ceil(n/128)/8) groups, each                            \
       work-item folds 8 consecutive elements into a reducer that starts at the                    \
       identity (memory_streaming.hpp:33-46, 94-113; wi_reducer.hpp:49-70);                         \
     - per group an LDS tree l[i] = op(l[i], l[i+s]), s = 64..1                                     \
       (group_reduction_algorithms.hpp:139-152);                                                    \
     - further stages of 128-wide groups over the partials until one value                         \
       remains (reduction_engine.hpp:123-142, 164-179, 248-363);                                    \
     - *out = op(init, result) (reduction_descriptor.hpp:36-38).                                    \
     For integers this equals oracle_reduce_##t; for floating point it is the                       \
     reference's evaluation order, including its padding identities. */                            \
  /* Faults 10 (adjacent pairs combined first) and 11 (operands swapped)                           \
     restate the tree wrongly; the crafted-input tests must catch both. */                          \
  static void tree128_##t(int op, T *l) {                                                           \
    if (g_fault == 10) {                                                                            \
      for (int w = 1; w < 128; w *= 2)                                                              \
        for (int i = 0; i + w < 128; i += 2 * w) l[i] = ap_##t(op, l[i], l[i + w]);                 \
      return;                                                                                       \
    }                                                                                               \
    for (int s = 64; s > 0; s /= 2)                                                                 \
      for (int i = 0; i < s; ++i)                                                                   \
        l[i] = (g_fault == 11) ? ap_##t(op, l[i + s], l[i]) : ap_##t(op, l[i], l[i + s]);           \
  }                                                                                                 \
  static void omp_stage_reduce_##t(int op, const T *in, uint64_t items, T *partials,                \
                                   uint64_t groups, int num_threads) {                              \
    _Pragma("omp parallel for num_threads(num_threads) schedule(static)")                          \
    for (uint64_t g = 0; g < groups; ++g) {                                                         \
      T l[128];                                                                                     \
      for (int lid = 0; lid < 128; ++lid) {                                                         \
        uint64_t id = g * 128 + lid;                                                                \
        T acc = refid_##t(op);                                                                      \
        if (id < items) acc = ap_##t(op, acc, in[id]);                                              \
        l[lid] = acc;                                                                               \
      }                                                                                             \
      tree128_##t(op, l);                                                                           \
      partials[g] = l[0];                                                                           \
    }                                                                                               \
  }                                                                                                 \
  static int omp_order_impl_##t(const T *a, const T *b, uint64_t n, T init, int op, T *out,        \
                                int num_threads) {                                                  \
    if (n == 0) return 0; /* numeric.hpp:165-169 */                                                 \
    if (num_threads <= 0) num_threads = omp_get_max_threads();                                      \
    const int red_op = b ? OP_PLUS : op;                                                            \
    if (g_fault == 9) { /* planted: a serial left fold instead of the reference's order */          \
      T acc = init;                                                                                 \
      for (uint64_t i = 0; i < n; ++i)                                                              \
        acc = ap_##t(red_op, acc, b ? ap_##t(OP_MUL, a[i], b[i]) : a[i]);                           \
      *out = acc;                                                                                   \
      return 0;                                                                                     \
    }                                                                                               \
    /* Fault 8 plants 16 elements per work-item in place of 8. */                                   \
    const uint64_t per_item = (g_fault == 8) ? 16 : 8;                                              \
    const uint64_t default_groups = (n + 127) / 128;                                                \
    const uint64_t groups = (default_groups + per_item - 1) / per_item;                             \
    T *stage = (T *)malloc(sizeof(T) * groups);                                                     \
    if (!stage) return -1;                                                                          \
    _Pragma("omp parallel for num_threads(num_threads) schedule(static)")                          \
    for (uint64_t g = 0; g < groups; ++g) {                                                         \
      T l[128];                                                                                     \
      for (int lid = 0; lid < 128; ++lid) {                                                         \
        const uint64_t gid = g * 128 + lid;                                                         \
        T acc = refid_##t(red_op);                                                                  \
        for (uint64_t i = 0; i < per_item; ++i) {                                                   \
          const uint64_t pos = per_item * gid + i;                                                  \
          if (pos < n) {                                                                            \
            T v = b ? ap_##t(OP_MUL, a[pos], b[pos]) : a[pos];                                      \
            /* fault 13 plants swapped operands in the work-item fold */                            \
            acc = (g_fault == 13) ? ap_##t(red_op, v, acc) : ap_##t(red_op, acc, v);                \
          }                                                                                         \
        }                                                                                           \
        l[lid] = acc;                                                                               \
      }                                                                                             \
      tree128_##t(red_op, l);                                                                       \
      stage[g] = l[0];                                                                              \
    }                                                                                               \
    uint64_t items = groups;                                                                        \
    T *next = (T *)malloc(sizeof(T) * ((items + 127) / 128 + 1));                                   \
    if (!next) {                                                                                    \
      free(stage);                                                                                  \
      return -1;                                                                                    \
    }                                                                                               \
    while (items > 1) {                                                                             \
      const uint64_t ng = (items + 127) / 128;                                                      \
      omp_stage_reduce_##t(red_op, stage, items, next, ng, num_threads);                            \
      memcpy(stage, next, sizeof(T) * ng);                                                          \
      items = ng;                                                                                   \
    }                                                                                               \
    /* fault 14 plants op(result, init) in place of op(init, result) */                             \
    *out = (g_fault == 14) ? ap_##t(red_op, stage[0], init) : ap_##t(red_op, init, stage[0]);       \
    free(next);                                                                                     \
    free(stage);                                                                                    \
    return 0;                                                                                       \
  }                                                                                                 \
  int oracle_reduce_omp_order_##t(const T *x, uint64_t n, T init, int op, T *out,                   \
                                  int num_threads) {                                                \
    if (!valid_op_##t(op)) return -1;                                                               \
    return omp_order_impl_##t(x, NULL, n, init, op, out, num_threads);                              \
  }                                                                                                 \
  int oracle_dot_omp_order_##t(const T *a, const T *b, uint64_t n, T init, T *out,                  \
                               int num_threads) {                                                   \
    return omp_order_impl_##t(a, b, n, init, OP_PLUS, out, num_threads);                            \
  }                                                                                                 \
  /* The reference's scan on its OpenMP backend: work-groups of 1024 x 16                         \
     elements (scan.hpp:26-31, decoupled_lookback_scan.hpp:481-488), dynamic                       \
     group ids (379-391); lane 0 scans the group serially                                          \
     (iterate_host_and_inclusive_group_scan, 139-177), publishes the aggregate,                    \
     looks back over predecessors combining op(earlier, later) until a                             \
     published prefix (219-257), publishes its prefix and adds the exclusive                       \
     prefix to its elements.  Exclusive scans read element i-1 and put init                        \
     at position 0 (load_data_element, 259-277).  In-place calls are not                           \
     supported by this restatement. */                                                             \
  int oracle_scan_omp_order_##t(const T *x, uint64_t n, T *y, int exclusive, int has_init,         \
                                T init, int op, int num_threads) {                                  \
    if (!valid_op_##t(op)) return -1;                                                               \
    if (n == 0) return 0;                                                                           \
    if (num_threads <= 0) num_threads = omp_get_max_threads();                                      \
    /* fault 12 plants 1024 x 8 tiles in place of 1024 x 16 */                                      \
    const uint64_t per_group = (g_fault == 12) ? 1024 * 8 : 1024 * 16;                              \
    const uint64_t groups = (n + per_group - 1) / per_group;                                        \
    T *aggregate = (T *)malloc(sizeof(T) * groups);                                                 \
    T *prefix = (T *)malloc(sizeof(T) * groups);                                                    \
    _Atomic uint32_t *status = (_Atomic uint32_t *)malloc(sizeof(_Atomic uint32_t) * groups);      \
    if (!aggregate || !prefix || !status) {                                                         \
      free(aggregate);                                                                              \
      free(prefix);                                                                                 \
      free((void *)status);                                                                         \
      return -1;                                                                                    \
    }                                                                                               \
    for (uint64_t g = 0; g < groups; ++g) atomic_init(&status[g], 0u);                              \
    _Atomic uint64_t counter;                                                                       \
    atomic_init(&counter, 0);                                                                       \
    _Pragma("omp parallel num_threads(num_threads)")                                               \
    {                                                                                               \
      for (;;) {                                                                                    \
        const uint64_t g = atomic_fetch_add_explicit(&counter, 1, memory_order_relaxed);            \
        if (g >= groups) break;                                                                     \
        const uint64_t begin = g * per_group;                                                       \
        const uint64_t end = (begin + per_group < n) ? begin + per_group : n;                       \
        T acc = x[0];                                                                               \
        for (uint64_t i = begin; i < end; ++i) {                                                    \
          T e;                                                                                      \
          if (exclusive)                                                                            \
            e = (i == 0) ? init : x[i - 1];                                                         \
          else                                                                                      \
            e = (i == 0 && has_init) ? ap_##t(op, init, x[0]) : x[i];                               \
          acc = (i == begin) ? e : ap_##t(op, acc, e);                                              \
          y[i] = acc;                                                                               \
        }                                                                                           \
        if (g == 0) {                                                                               \
          aggregate[0] = acc;                                                                       \
          prefix[0] = acc;                                                                          \
          atomic_store_explicit(&status[0], 2u, memory_order_release);                              \
          continue;                                                                                 \
        }                                                                                           \
        aggregate[g] = acc;                                                                         \
        atomic_store_explicit(&status[g], 1u, memory_order_release);                                \
        T excl = acc;                                                                               \
        int have = 0;                                                                               \
        for (int64_t lb = (int64_t)g - 1; lb >= 0; --lb) {                                          \
          uint32_t s;                                                                               \
          while ((s = atomic_load_explicit(&status[lb], memory_order_acquire)) == 0u) {             \
          }                                                                                         \
          const T v = (s == 2u) ? prefix[lb] : aggregate[lb];                                       \
          excl = have ? ap_##t(op, v, excl) : v;                                                    \
          have = 1;                                                                                 \
          if (s == 2u) break;                                                                       \
        }                                                                                           \
        prefix[g] = ap_##t(op, excl, acc);                                                          \
        atomic_store_explicit(&status[g], 2u, memory_order_release);                                \
        for (uint64_t i = begin; i < end; ++i) y[i] = ap_##t(op, excl, y[i]);                       \
      }                                                                                             \
    }                                                                                               \
    free(aggregate);                                                                                \
    free(prefix);                                                                                   \
    free((void *)status);                                                                           \
    return 0;                                                                                       \
  }                                                                                                 \
  /* Work-group collectives on the host backend: the group leader folds the                       \
     group serially (include/hipSYCL/sycl/libkernel/host/group_functions.hpp:                       \
     52-75 reduce, 495-515 inclusive scan, 407-428 exclusive scan).  Every                         \
     work-item receives the result; the last group may be partial. */                              \
  int oracle_group_reduce_##t(const T *x, uint64_t n, uint64_t local_size, int op, T *y) {          \
    if (!valid_op_##t(op) || local_size == 0) return -1;                                            \
    for (uint64_t b = 0; b < n; b += local_size) {                                                  \
      uint64_t e = (b + local_size < n) ? b + local_size : n;                                       \
      if (g_fault == 4 && e - b > 1) e -= 1;                                                        \
      T r = x[b];                                                                                   \
      for (uint64_t i = b + 1; i < e; ++i) r = ap_##t(op, r, x[i]);                                 \
      uint64_t ee = (b + local_size < n) ? b + local_size : n;                                      \
      for (uint64_t i = b; i < ee; ++i) y[i] = r;                                                   \
    }                                                                                               \
    return 0;                                                                                       \
  }                                                                                                 \
  int oracle_group_inclusive_scan_##t(const T *x, uint64_t n, uint64_t local_size, int op, T *y) {  \
    if (!valid_op_##t(op) || local_size == 0) return -1;                                            \
    for (uint64_t b = 0; b < n; b += local_size) {                                                  \
      const uint64_t e = (b + local_size < n) ? b + local_size : n;                                 \
      y[b] = x[b];                                                                                  \
      for (uint64_t i = b + 1; i < e; ++i) y[i] = ap_##t(op, x[i], y[i - 1]);                       \
    }                                                                                               \
    return 0;                                                                                       \
  }                                                                                                 \
  int oracle_group_exclusive_scan_##t(const T *x, uint64_t n, uint64_t local_size, T init,         \
                                      int op, T *y) {                                               \
    if (!valid_op_##t(op) || local_size == 0) return -1;                                            \
    for (uint64_t b = 0; b < n; b += local_size) {                                                  \
      const uint64_t e = (b + local_size < n) ? b + local_size : n;                                 \
      y[b] = init;                                                                                  \
      for (uint64_t i = b + 1; i < e; ++i) y[i] = ap_##t(op, x[i - 1], y[i - 1]);                   \
    }                                                                                               \
    return 0;                                                                                       \
  }                                                                                                 \
  /* Serial fold of atomic read-modify-writes on one location                                     \
     (tests/sycl/atomic.cpp:112-123; host/atomic_builtins.hpp:235-312). */                          \
  int oracle_atomic_fold_##t(int op, const T *operands, uint64_t n, T initial, T *final_value) {    \
    T v = initial;                                                                                  \
    for (uint64_t i = 0; i < n; ++i) {                                                              \
      const T x = operands[i];                                                                      \
      switch (op) {                                                                                 \
        case RMW_ADD: v = ap_##t(OP_PLUS, v, x); break;                                             \
        case RMW_SUB: v = sub_##t(v, x); break;                                                     \
        case RMW_MIN: v = (x < v) ? x : v; break;                                                   \
        case RMW_MAX: v = (x > v) ? x : v; break;                                                   \
        case RMW_AND:                                                                               \
        case RMW_OR:                                                                                \
        case RMW_XOR:                                                                               \
          if (!IS_INT) return -1;                                                                   \
          v = ap_##t(op == RMW_AND ? OP_BAND : (op == RMW_OR ? OP_BOR : OP_BXOR), v, x);            \
          break;                                                                                    \
        default: return -1;                                                                         \
      }                                                                                             \
    }                                                                                               \
    *final_value = v;                                                                               \
    return 0;                                                                                       \
  }                                                                                                 \
  /* Segmented forms: an independent scan / reduction per segment                                 \
     [offsets[s], offsets[s+1]), the semantics of one joint_* call per segment                     \
     (include/hipSYCL/sycl/libkernel/generic/hiplike/group_functions.hpp:214-287,                  \
     361-432). */                                                                                   \
  int oracle_segmented_inclusive_scan_##t(const T *x, const uint64_t *offsets,                      \
                                          uint64_t num_segments, T *y, int op) {                    \
    if (!valid_op_##t(op)) return -1;                                                               \
    for (uint64_t s = 0; s < num_segments; ++s) {                                                   \
      const uint64_t b = offsets[s], e = offsets[s + 1];                                            \
      if (b >= e) continue;                                                                         \
      T acc = x[b];                                                                                 \
      y[b] = acc;                                                                                   \
      for (uint64_t i = b + 1; i < e; ++i) {                                                        \
        acc = ap_##t(op, acc, x[i]);                                                                \
        y[i] = acc;                                                                                 \
      }                                                                                             \
    }                                                                                               \
    return 0;                                                                                       \
  }                                                                                                 \
  int oracle_segmented_reduce_##t(const T *x, const uint64_t *offsets, uint64_t num_segments,       \
                                  T init, int op, T *out) {                                         \
    if (!valid_op_##t(op)) return -1;                                                               \
    for (uint64_t s = 0; s < num_segments; ++s) {                                                   \
      T acc = init;                                                                                 \
      for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i) acc = ap_##t(op, acc, x[i]);           \
      out[s] = acc;                                                                                 \
    }                                                                                               \
    return 0;                                                                                       \
  }

ORACLE_DEFINE(i8, int8_t, 1)
ORACLE_DEFINE(i16, int16_t, 1)
ORACLE_DEFINE(i32, int32_t, 1)
ORACLE_DEFINE(i64, int64_t, 1)
ORACLE_DEFINE(u8, uint8_t, 1)
ORACLE_DEFINE(u16, uint16_t, 1)
ORACLE_DEFINE(u32, uint32_t, 1)
ORACLE_DEFINE(u64, uint64_t, 1)
ORACLE_DEFINE(f32, float, 0)
ORACLE_DEFINE(f64, double, 0)

/* ------------------------------------------------------------------------- */
/* FP64 oracles                                                              */
/* ------------------------------------------------------------------------- */
/* Neumaier summation in FP64 over 64 fixed chunks (deterministic for any
   thread count), chunks combined by a final Neumaier pass. */
typedef struct {
  double s, c;
} neum_t;

static inline void neum_add(neum_t *a, double x) {
  const double t = a->s + x;
  if (fabs(a->s) >= fabs(x))
    a->c += (a->s - t) + x;
  else
    a->c += (x - t) + a->s;
  a->s = t;
}

#define NCHUNK 64
#define DEF_FP64_SUM(name, ARGS, VALUE)                                         \
  double name ARGS {                                                            \
    neum_t parts[NCHUNK];                                                       \
    _Pragma("omp parallel for schedule(static)")                                \
    for (int c = 0; c < NCHUNK; ++c) {                                          \
      const uint64_t lo = (n * (uint64_t)c) / NCHUNK;                           \
      const uint64_t hi = (n * (uint64_t)(c + 1)) / NCHUNK;                     \
      neum_t acc = {0.0, 0.0};                                                  \
      for (uint64_t i = lo; i < hi; ++i) neum_add(&acc, (VALUE));               \
      parts[c] = acc;                                                           \
    }                                                                           \
    neum_t tot = {0.0, 0.0};                                                    \
    for (int c = 0; c < NCHUNK; ++c) {                                          \
      neum_add(&tot, parts[c].s);                                               \
      neum_add(&tot, parts[c].c);                                               \
    }                                                                           \
    return tot.s + tot.c;                                                       \
  }

static inline double bf16_to_double(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}

DEF_FP64_SUM(oracle_sum_f32_fp64, (const float *x, uint64_t n), (double)x[i])
DEF_FP64_SUM(oracle_abs_sum_f32_fp64, (const float *x, uint64_t n), fabs((double)x[i]))
DEF_FP64_SUM(oracle_sum_bf16_fp64, (const uint16_t *x, uint64_t n), bf16_to_double(x[i]))
/* Products of two floats are exact in FP64 (48-bit significands). */
DEF_FP64_SUM(oracle_dot_f32_fp64, (const float *a, const float *b, uint64_t n),
             (double)a[i] * (double)b[i])
DEF_FP64_SUM(oracle_dot_bf16_fp64, (const uint16_t *a, const uint16_t *b, uint64_t n),
             bf16_to_double(a[i]) * bf16_to_double(b[i]))

int oracle_segmented_reduce_f32_fp64(const float *x, const uint64_t *offsets,
                                     uint64_t num_segments, double *out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (uint64_t s = 0; s < num_segments; ++s) {
    neum_t acc = {0.0, 0.0};
    for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i) neum_add(&acc, (double)x[i]);
    out[s] = acc.s + acc.c;
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Atomics workloads                                                         */
/* ------------------------------------------------------------------------- */
/* Per-thread private counts, summed afterwards: exact, order-independent. */
int oracle_histogram_u32(const uint32_t *keys, uint64_t n, uint32_t *bins, uint32_t num_bins,
                         int num_threads) {
  if (num_threads <= 0) num_threads = omp_get_max_threads();
  uint32_t *priv = (uint32_t *)calloc((size_t)num_threads * num_bins, sizeof(uint32_t));
  if (!priv) return -1;
#pragma omp parallel num_threads(num_threads)
  {
    uint32_t *mine = priv + (size_t)omp_get_thread_num() * num_bins;
#pragma omp for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t k = keys[i];
      if (k < num_bins && !(g_fault == 5 && k == 0)) mine[k] += 1u;
    }
  }
  for (int t = 0; t < num_threads; ++t)
    for (uint32_t b = 0; b < num_bins; ++b) bins[b] += priv[(size_t)t * num_bins + b];
  free(priv);
  return 0;
}

int oracle_fetch_max_strided_i32(const int32_t *values, uint64_t n, int32_t *slots,
                                 uint32_t num_slots, int num_threads) {
  if (num_threads <= 0) num_threads = omp_get_max_threads();
  if (num_slots == 0) return -1;
  int32_t *priv = (int32_t *)malloc((size_t)num_threads * num_slots * sizeof(int32_t));
  if (!priv) return -1;
  for (size_t i = 0; i < (size_t)num_threads * num_slots; ++i) priv[i] = INT32_MIN;
#pragma omp parallel num_threads(num_threads)
  {
    int32_t *mine = priv + (size_t)omp_get_thread_num() * num_slots;
#pragma omp for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t s = (uint32_t)(i % num_slots);
      if (values[i] > mine[s]) mine[s] = values[i];
    }
  }
  for (int t = 0; t < num_threads; ++t)
    for (uint32_t s = 0; s < num_slots; ++s) {
      const int32_t v = priv[(size_t)t * num_slots + s];
      if (v > slots[s]) slots[s] = v;
    }
  free(priv);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* copy_if (include/hipSYCL/algorithms/algorithm.hpp:178-243)                */
/* ------------------------------------------------------------------------- */
/* The reference's structure: flags[i] = pred(x[i]) ? 1 : 0 (generator,
   algorithm.hpp:194-205), an exclusive plus-scan of the flags (ScanT =
   size_t, algorithm.hpp:192 and 237-242), then every selected x[i] is
   written to d_first[scan[i]] and the last element reports
   scan[n-1] + flag[n-1] as the number copied (algorithm.hpp:207-234).
   Predicate codes: 0 x != 0, 1 ==, 2 !=, 3 <, 4 <=, 5 >, 6 >=, 7 (x & mask)
   == operand, 8 stencil[i] != 0.  Fault 6 drops the last selected element
   from the count. */
#define DEF_COPY_IF(t, T, IS_INT)                                                             \
  static int pred_##t(int pred, T x, T operand, T mask, const uint8_t *stencil, uint64_t i) { \
    (void)mask;                                                                             \
    switch (pred) {                                                                         \
      case 0: return x != (T)0;                                                             \
      case 1: return x == operand;                                                          \
      case 2: return x != operand;                                                          \
      case 3: return x < operand;                                                           \
      case 4: return x <= operand;                                                          \
      case 5: return x > operand;                                                           \
      case 6: return x >= operand;                                                          \
      case 7: return IS_INT ? (COPY_IF_MASK(T, x, mask) == operand) : 0;                     \
      default: return stencil[i] != 0;                                                      \
    }                                                                                       \
  }                                                                                         \
  int oracle_copy_if_##t(const T *x, uint64_t n, const uint8_t *stencil, int pred, T operand, \
                         T mask, T *out, uint64_t *count) {                                 \
    if (n == 0) {                                                                           \
      *count = 0;                                                                           \
      return 0;                                                                             \
    }                                                                                       \
    uint64_t *scan = (uint64_t *)malloc(n * sizeof(uint64_t));                              \
    if (!scan) return -1;                                                                   \
    uint64_t acc = 0;                                                                       \
    for (uint64_t i = 0; i < n; ++i) {                                                      \
      scan[i] = acc;                                                                        \
      acc += pred_##t(pred, x[i], operand, mask, stencil, i) ? 1u : 0u;                      \
    }                                                                                       \
    int last = 0;                                                                           \
    for (uint64_t i = 0; i < n; ++i) {                                                      \
      last = pred_##t(pred, x[i], operand, mask, stencil, i);                               \
      if (last) out[scan[i]] = x[i];                                                        \
    }                                                                                       \
    *count = scan[n - 1] + (uint64_t)last;                                                  \
    if (g_fault == 6 && *count > 0) *count -= 1;                                            \
    free(scan);                                                                             \
    return 0;                                                                               \
  }

#define COPY_IF_MASK(T, x, mask) COPY_IF_MASK_##T(x, mask)
#define COPY_IF_MASK_int8_t(x, mask) ((int8_t)((x) & (mask)))
#define COPY_IF_MASK_uint8_t(x, mask) ((uint8_t)((x) & (mask)))
#define COPY_IF_MASK_int16_t(x, mask) ((int16_t)((x) & (mask)))
#define COPY_IF_MASK_uint16_t(x, mask) ((uint16_t)((x) & (mask)))
#define COPY_IF_MASK_int32_t(x, mask) ((int32_t)((x) & (mask)))
#define COPY_IF_MASK_uint32_t(x, mask) ((uint32_t)((x) & (mask)))
#define COPY_IF_MASK_int64_t(x, mask) ((int64_t)((x) & (mask)))
#define COPY_IF_MASK_uint64_t(x, mask) ((uint64_t)((x) & (mask)))
#define COPY_IF_MASK_float(x, mask) ((float)0)
#define COPY_IF_MASK_double(x, mask) ((double)0)
DEF_COPY_IF(i8, int8_t, 1)
DEF_COPY_IF(u8, uint8_t, 1)
DEF_COPY_IF(i16, int16_t, 1)
DEF_COPY_IF(u16, uint16_t, 1)
DEF_COPY_IF(i32, int32_t, 1)
DEF_COPY_IF(u32, uint32_t, 1)
DEF_COPY_IF(i64, int64_t, 1)
DEF_COPY_IF(u64, uint64_t, 1)
DEF_COPY_IF(f32, float, 0)
DEF_COPY_IF(f64, double, 0)

/* ------------------------------------------------------------------------- */
/* Element transforms of transform_reduce (numeric.hpp:193-212) and of the   */
/* transform scans (numeric.hpp:299-327; scan/scan.hpp:85-132, whose          */
/* generator yields unary_op(*it) and then scans as usual).  y[i] =           */
/* t(x[i]) for the transform codes 0 identity, 1 -x, 2 |x|, 3 x*x, 4 x+x,     */
/* 5 ~x (integers), 6 x == 0, 7 x != 0.  Integer results wrap modulo          */
/* 2^bits (computed in the unsigned type WU).  Fault 7 turns |x| into x.      */
/* ------------------------------------------------------------------------- */
#define DEF_TRANSFORM_INT(t, T, WU, SIGNED)                                          \
  int oracle_transform_##t(const T *x, uint64_t n, T *y, int code) {                 \
    if (code < 0 || code > 7) return -1;                                             \
    for (uint64_t i = 0; i < n; ++i) {                                               \
      const T v = x[i];                                                              \
      const WU u = (WU)v;                                                            \
      T r = v;                                                                       \
      switch (code) {                                                                \
        case 1: r = (T)((WU)0 - u); break;                                           \
        case 2: r = (SIGNED && ((u >> (8 * sizeof(T) - 1)) & 1u) && g_fault != 7)    \
                        ? (T)((WU)0 - u) : v;                                        \
                break;                                                               \
        case 3: r = (T)(u * u); break;                                               \
        case 4: r = (T)(u + u); break;                                               \
        case 5: r = (T)(~u); break;                                                  \
        case 6: r = (T)(v == (T)0); break;                                           \
        case 7: r = (T)(v != (T)0); break;                                           \
        default: break;                                                              \
      }                                                                              \
      y[i] = r;                                                                      \
    }                                                                                \
    return 0;                                                                        \
  }
#define DEF_TRANSFORM_FP(t, T, FABS)                                                 \
  int oracle_transform_##t(const T *x, uint64_t n, T *y, int code) {                 \
    if (code < 0 || code > 7 || code == 5) return -1;                                \
    for (uint64_t i = 0; i < n; ++i) {                                               \
      const T v = x[i];                                                              \
      T r = v;                                                                       \
      switch (code) {                                                                \
        case 1: r = -v; break;                                                       \
        case 2: r = (g_fault == 7) ? v : FABS(v); break;                             \
        case 3: r = v * v; break;                                                    \
        case 4: r = v + v; break;                                                    \
        case 6: r = (v == (T)0) ? (T)1 : (T)0; break;                                \
        case 7: r = (v != (T)0) ? (T)1 : (T)0; break;                                \
        default: break;                                                              \
      }                                                                              \
      y[i] = r;                                                                      \
    }                                                                                \
    return 0;                                                                        \
  }
DEF_TRANSFORM_INT(i8, int8_t, uint32_t, 1)
DEF_TRANSFORM_INT(i16, int16_t, uint32_t, 1)
DEF_TRANSFORM_INT(i32, int32_t, uint32_t, 1)
DEF_TRANSFORM_INT(i64, int64_t, uint64_t, 1)
DEF_TRANSFORM_INT(u8, uint8_t, uint32_t, 0)
DEF_TRANSFORM_INT(u16, uint16_t, uint32_t, 0)
DEF_TRANSFORM_INT(u32, uint32_t, uint32_t, 0)
DEF_TRANSFORM_INT(u64, uint64_t, uint64_t, 0)
DEF_TRANSFORM_FP(f32, float, fabsf)
DEF_TRANSFORM_FP(f64, double, fabs)

/* ------------------------------------------------------------------------- */
/* Binary transform_reduce (numeric.hpp:170-191): acc = init; acc =           */
/* op(acc, t(a[i], b[i])) in order (std::transform_reduce's serial order, as  */
/* tests/pstl/transform_reduce.cpp checks it).  Pair transform codes: 0 x*y,  */
/* 1 x+y, 2 x-y, 3 |x-y|, 4 (x-y)^2, 5 x<y?x:y, 6 x>y?x:y, 7 x&y, 8 x|y,       */
/* 9 x^y (integers).  Integer results wrap (computed in the unsigned WU).     */
/* Fault 15 turns |x-y| into x-y.                                             */
/* ------------------------------------------------------------------------- */
#define DEF_BINARY_INT(t, T, WU, SIGNED)                                                        \
  static int bt_##t(int code, T x, T y, T *r) {                                                 \
    const WU a = (WU)x, b = (WU)y;                                                              \
    switch (code) {                                                                             \
      case 0: *r = (T)(a * b); return 0;                                                        \
      case 1: *r = (T)(a + b); return 0;                                                        \
      case 2: *r = (T)(a - b); return 0;                                                        \
      case 3: {                                                                                 \
        const T d = (T)(a - b);                                                                 \
        const int neg = SIGNED && ((((WU)d) >> (8 * sizeof(T) - 1)) & 1u);                      \
        *r = (neg && g_fault != 15) ? (T)((WU)0 - (WU)d) : d;                                   \
        return 0;                                                                               \
      }                                                                                         \
      case 4: *r = (T)((a - b) * (a - b)); return 0;                                            \
      case 5: *r = (x < y) ? x : y; return 0;                                                   \
      case 6: *r = (x > y) ? x : y; return 0;                                                   \
      case 7: *r = (T)(a & b); return 0;                                                        \
      case 8: *r = (T)(a | b); return 0;                                                        \
      case 9: *r = (T)(a ^ b); return 0;                                                        \
      default: return -1;                                                                       \
    }                                                                                           \
  }
#define DEF_BINARY_FP(t, T, FABS)                                                               \
  static int bt_##t(int code, T x, T y, T *r) {                                                 \
    switch (code) {                                                                             \
      case 0: *r = x * y; return 0;                                                             \
      case 1: *r = x + y; return 0;                                                             \
      case 2: *r = x - y; return 0;                                                             \
      case 3: *r = (g_fault == 15) ? (x - y) : FABS(x - y); return 0;                           \
      case 4: {                                                                                 \
        const T d = x - y;                                                                      \
        *r = d * d;                                                                             \
        return 0;                                                                               \
      }                                                                                         \
      case 5: *r = (x < y) ? x : y; return 0;                                                   \
      case 6: *r = (x > y) ? x : y; return 0;                                                   \
      default: return -1;                                                                       \
    }                                                                                           \
  }
#define DEF_BINARY_REDUCE(t, T)                                                                 \
  int oracle_transform_reduce_binary_##t(const T *a, const T *b, uint64_t n, T init, int op,    \
                                         int code, T *out) {                                    \
    if (!valid_op_##t(op)) return -1;                                                           \
    T acc = init, v;                                                                            \
    if (bt_##t(code, (T)0, (T)0, &v) != 0) return -1;                                           \
    for (uint64_t i = 0; i < n; ++i) {                                                          \
      bt_##t(code, a[i], b[i], &v);                                                             \
      acc = ap_##t(op, acc, v);                                                                 \
    }                                                                                           \
    *out = acc;                                                                                 \
    return 0;                                                                                   \
  }
DEF_BINARY_INT(i32, int32_t, uint32_t, 1)
DEF_BINARY_INT(i64, int64_t, uint64_t, 1)
DEF_BINARY_INT(u32, uint32_t, uint32_t, 0)
DEF_BINARY_INT(u64, uint64_t, uint64_t, 0)
DEF_BINARY_FP(f32, float, fabsf)
DEF_BINARY_FP(f64, double, fabs)
DEF_BINARY_REDUCE(i32, int32_t)
DEF_BINARY_REDUCE(i64, int64_t)
DEF_BINARY_REDUCE(u32, uint32_t)
DEF_BINARY_REDUCE(u64, uint64_t)
DEF_BINARY_REDUCE(f32, float)
DEF_BINARY_REDUCE(f64, double)

/* ------------------------------------------------------------------------- */
/* Group data movement (include/hipSYCL/sycl/libkernel/group_functions.hpp:   */
/* 57-140 broadcast / any / all / none, 240-276 shift_group_left/right,       */
/* permute_group_by_xor, select_from_group).  The elements are split into     */
/* work-groups of local_size (the last may be partial); level 0 applies the   */
/* operation within sub-groups of 64 consecutive work-items of a work-group   */
/* (the wave64 sub-group), level 1 within the work-group.  which: 0 broadcast */
/* of member param, 1 shift left by param, 2 shift right by param, 3 permute  */
/* by xor with param, 4 select member src[i], 5 any (x > param), 6 all, 7     */
/* none, 8-10 the joint_ forms of 5-7 (over the group's elements).  A source  */
/* outside the group leaves the member's own value at sub-group level and     */
/* gives member 0's value at work-group level (the reference leaves the       */
/* result undefined there).  Fault 16 shifts right by param - 1.              */
/* ------------------------------------------------------------------------- */
#define DEF_GROUP_MOVE(t, T)                                                                          \
  int oracle_group_move_##t(int level, int which, int64_t param, const T *x, const uint32_t *src,    \
                            uint64_t n, uint64_t local_size, T *out) {                               \
    if (local_size == 0 || which < 0 || which > 10 || level < 0 || level > 1) return -1;             \
    for (uint64_t b0 = 0; b0 < n; b0 += local_size) {                                                \
      const uint64_t wg = (b0 + local_size < n) ? local_size : n - b0;                               \
      const uint64_t span = level == 0 ? 64 : wg;                                                     \
      for (uint64_t g0 = 0; g0 < wg; g0 += span) {                                                    \
        const uint64_t size = (g0 + span < wg) ? span : wg - g0;                                      \
        const T *xs = x + b0 + g0;                                                                    \
        int any = 0, all = 1;                                                                         \
        for (uint64_t l = 0; l < size; ++l) {                                                         \
          const int p = (double)xs[l] > (double)param;                                                \
          any |= p;                                                                                   \
          all &= p;                                                                                   \
        }                                                                                             \
        for (uint64_t l = 0; l < size; ++l) {                                                         \
          const uint64_t i = b0 + g0 + l;                                                             \
          const T own = level == 0 ? xs[l] : xs[0];                                                   \
          int64_t s = -1;                                                                             \
          switch (which) {                                                                            \
            case 0: s = param; break;                                                                 \
            case 1: s = (int64_t)l + param; break;                                                    \
            case 2: s = (int64_t)l - (g_fault == 16 ? param - 1 : param); break;                      \
            case 3: s = (int64_t)(l ^ (uint64_t)param); break;                                        \
            case 4: s = (int64_t)src[i]; break;                                                       \
            default: break;                                                                           \
          }                                                                                           \
          if (which <= 4)                                                                             \
            out[i] = (s >= 0 && (uint64_t)s < size) ? xs[s] : own;                                   \
          else if (which == 5 || which == 8)                                                          \
            out[i] = (T)(any ? 1 : 0);                                                                \
          else if (which == 6 || which == 9)                                                          \
            out[i] = (T)(all ? 1 : 0);                                                                \
          else                                                                                        \
            out[i] = (T)(any ? 0 : 1);                                                                \
        }                                                                                             \
      }                                                                                               \
    }                                                                                                 \
    return 0;                                                                                         \
  }
DEF_GROUP_MOVE(i32, int32_t)
DEF_GROUP_MOVE(i64, int64_t)
DEF_GROUP_MOVE(f32, float)
DEF_GROUP_MOVE(f64, double)
