// Test kernels for the device headers (group collectives and atomic_ref).
// Built into tests/lib/libacpp_test_kernels.so; the GPU tests call these
// launchers through ctypes and compare with the oracle.  The kernels mirror
// the reference's own tests:
//   tests/sycl/group_functions/group_functions_{reduce,scan}.cpp
//   tests/compiler/cbs/{reduce_for,babelstream_dot}.cpp
//   tests/sycl/atomic.cpp, tests/pstl/std_atomic.cpp
//   tests/sycl/reduction.cpp (reducers inside a kernel)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "acpp/device/atomic_ref.hpp"
#include "acpp/device/group_functions.hpp"
#include "acpp/device/reducer.hpp"

namespace {

enum kind_t : int {
  k_wg_reduce = 0,
  k_wg_inclusive = 1,
  k_wg_exclusive = 2,  // with init
  k_sg_reduce = 3,
  k_sg_inclusive = 4,
  k_sg_exclusive = 5,  // with init
  k_wg_reduce_init = 6,
  k_wg_inclusive_init = 7,
  k_wg_exclusive_identity = 8,  // exclusive scan without init (known identity)
  // the same through the extern "C" builtin ABI
  k_abi_wg_reduce = 10,
  k_abi_wg_inclusive = 11,
  k_abi_wg_exclusive = 12,
  k_abi_sg_reduce = 13,
  k_abi_sg_inclusive = 14,
  k_abi_sg_exclusive = 15,
};

template <class T, class F>
__device__ T run_kind(int kind, T x, T init, F op) {
  acpp::group g = acpp::this_group();
  acpp::sub_group sg = acpp::this_sub_group();
  switch (kind) {
    case k_wg_reduce: return acpp::reduce_over_group(g, x, op);
    case k_wg_inclusive: return acpp::inclusive_scan_over_group(g, x, op);
    case k_wg_exclusive: return acpp::exclusive_scan_over_group(g, x, init, op);
    case k_sg_reduce: return acpp::reduce_over_group(sg, x, op);
    case k_sg_inclusive: return acpp::inclusive_scan_over_group(sg, x, op);
    case k_sg_exclusive: return acpp::exclusive_scan_over_group(sg, x, init, op);
    case k_wg_reduce_init: return acpp::reduce_over_group(g, x, init, op);
    case k_wg_inclusive_init: return acpp::inclusive_scan_over_group(g, x, init, op);
    default: break;
  }
  if constexpr (acpp::known_identity<F, T>::value) {
    if (kind == k_wg_exclusive_identity) return acpp::exclusive_scan_over_group(g, x, op);
  }
  return x;
}

template <class T>
__device__ T run_op(int kind, int op, T x, T init) {
  switch (op) {
    case 0: return run_kind(kind, x, init, acpp::plus<T>{});
    case 1: return run_kind(kind, x, init, acpp::multiplies<T>{});
    case 2: return run_kind(kind, x, init, acpp::minimum<T>{});
    case 3: return run_kind(kind, x, init, acpp::maximum<T>{});
    default: break;
  }
  if constexpr (std::is_integral_v<T>) {
    switch (op) {
      case 4: return run_kind(kind, x, init, acpp::bit_and<T>{});
      case 5: return run_kind(kind, x, init, acpp::bit_or<T>{});
      case 6: return run_kind(kind, x, init, acpp::bit_xor<T>{});
      case 7: return run_kind(kind, x, init, acpp::logical_and<T>{});
      case 8: return run_kind(kind, x, init, acpp::logical_or<T>{});
      default: break;
    }
  }
  return x;
}

#define ABI_CALL(tag, T)                                                                    \
  __device__ T abi_##tag(int kind, int op, T x, T init) {                                   \
    switch (kind) {                                                                         \
      case k_abi_wg_reduce: return __acpp_sscp_work_group_reduce_##tag(op, x);              \
      case k_abi_wg_inclusive: return __acpp_sscp_work_group_inclusive_scan_##tag(op, x);   \
      case k_abi_wg_exclusive: return __acpp_sscp_work_group_exclusive_scan_##tag(op, x, init); \
      case k_abi_sg_reduce: return __acpp_sscp_sub_group_reduce_##tag(op, x);               \
      case k_abi_sg_inclusive: return __acpp_sscp_sub_group_inclusive_scan_##tag(op, x);    \
      default: return __acpp_sscp_sub_group_exclusive_scan_##tag(op, x, init);              \
    }                                                                                       \
  }
ABI_CALL(i8, int8_t)
ABI_CALL(i32, int32_t)
ABI_CALL(i64, int64_t)
ABI_CALL(u32, uint32_t)
ABI_CALL(f16, _Float16)
ABI_CALL(f32, float)
ABI_CALL(f64, double)

template <class T>
__device__ T abi_dispatch(int kind, int op, T x, T init) {
  if constexpr (std::is_same_v<T, int8_t>) return abi_i8(kind, op, x, init);
  else if constexpr (std::is_same_v<T, int32_t>) return abi_i32(kind, op, x, init);
  else if constexpr (std::is_same_v<T, int64_t>) return abi_i64(kind, op, x, init);
  else if constexpr (std::is_same_v<T, uint32_t>) return abi_u32(kind, op, x, init);
  else if constexpr (std::is_same_v<T, _Float16>) return abi_f16(kind, op, x, init);
  else if constexpr (std::is_same_v<T, float>) return abi_f32(kind, op, x, init);
  else return abi_f64(kind, op, x, init);
}

template <class T>
__global__ void collective_kernel(int kind, int op, const T *in, T *out, T init) {
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const T x = in[gid];
  T r;
  if (kind >= k_abi_wg_reduce)
    r = abi_dispatch<T>(kind, op, x, init);
  else
    r = run_op<T>(kind, op, x, init);
  out[gid] = r;
}

template <class T>
__global__ void joint_kernel(int kind, int op, const T *in, T *out, uint64_t seg_len, T init) {
  const T *first = in + static_cast<uint64_t>(blockIdx.x) * seg_len;
  T *result = out + static_cast<uint64_t>(blockIdx.x) * seg_len;
  acpp::group g = acpp::this_group();
  // kinds: 0 joint_reduce(init) -> result[0], 1 joint_inclusive_scan,
  // 2 joint_exclusive_scan(init), 3 joint_inclusive_scan(init),
  // 4 joint_reduce without init (plus)
  if (kind == 0) {
    T r = T{};
    switch (op) {
      case 0: r = acpp::joint_reduce(g, first, first + seg_len, init, acpp::plus<T>{}); break;
      case 2: r = acpp::joint_reduce(g, first, first + seg_len, init, acpp::minimum<T>{}); break;
      default: r = acpp::joint_reduce(g, first, first + seg_len, init, acpp::maximum<T>{}); break;
    }
    if (g.leader()) result[0] = r;
  } else if (kind == 1) {
    acpp::joint_inclusive_scan(g, first, first + seg_len, result, acpp::plus<T>{});
  } else if (kind == 2) {
    acpp::joint_exclusive_scan(g, first, first + seg_len, result, init, acpp::plus<T>{});
  } else if (kind == 3) {
    acpp::joint_inclusive_scan(g, first, first + seg_len, result, acpp::plus<T>{}, init);
  } else {
    T r = acpp::joint_reduce(g, first, first + seg_len, acpp::plus<T>{});
    if (g.leader()) result[0] = r;
  }
}

// tests/compiler/cbs/babelstream_dot.cpp: 32 groups of 4 work-items, a
// grid-stride dot product, then one partial per group.
__global__ void babelstream_dot_kernel(const int32_t *a, const int32_t *b, int32_t *sums,
                                       uint64_t n) {
  const uint64_t gsize = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  int32_t acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += gsize)
    acc += a[i] * b[i];
  const int32_t r = acpp::reduce_over_group(acpp::this_group(), acc, acpp::plus<int32_t>{});
  if (threadIdx.x == 0) sums[blockIdx.x] = r;
}

// tests/sycl/atomic.cpp:81-123: work-items 1..n-1 apply an atomic op with
// operand data[i] to data[0].  rmw op numbering of the atomic builtin ABI:
// and, or, xor, add, sub, min, max.
template <class T, bool ABI>
__global__ void atomic_fold_kernel(int op, T *data, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0 || i >= n) return;
  const T x = data[i];
  using acpp::memory_order;
  using acpp::memory_scope;
  if constexpr (ABI) {
    const auto as = acpp::address_space::global_space;
    const auto o = memory_order::relaxed;
    const auto s = memory_scope::device;
    if constexpr (std::is_same_v<T, int32_t>) {
      switch (op) {
        case 0: __acpp_sscp_atomic_fetch_and_i32(as, o, s, data, x); break;
        case 1: __acpp_sscp_atomic_fetch_or_i32(as, o, s, data, x); break;
        case 2: __acpp_sscp_atomic_fetch_xor_i32(as, o, s, data, x); break;
        case 3: __acpp_sscp_atomic_fetch_add_i32(as, o, s, data, x); break;
        case 4: __acpp_sscp_atomic_fetch_sub_i32(as, o, s, data, x); break;
        case 5: __acpp_sscp_atomic_fetch_min_i32(as, o, s, data, x); break;
        default: __acpp_sscp_atomic_fetch_max_i32(as, o, s, data, x); break;
      }
    } else if constexpr (std::is_same_v<T, float>) {
      switch (op) {
        case 3: __acpp_sscp_atomic_fetch_add_f32(as, o, s, data, x); break;
        case 4: __acpp_sscp_atomic_fetch_sub_f32(as, o, s, data, x); break;
        case 5: __acpp_sscp_atomic_fetch_min_f32(as, o, s, data, x); break;
        default: __acpp_sscp_atomic_fetch_max_f32(as, o, s, data, x); break;
      }
    }
  } else {
    acpp::atomic_ref<T, memory_order::relaxed, memory_scope::device> r{data[0]};
    if constexpr (std::is_integral_v<T>) {
      switch (op) {
        case 0: r.fetch_and(x); return;
        case 1: r.fetch_or(x); return;
        case 2: r.fetch_xor(x); return;
        default: break;
      }
    }
    switch (op) {
      case 3: r.fetch_add(x); break;
      case 4: r.fetch_sub(x); break;
      case 5: r.fetch_min(x); break;
      case 6: r.fetch_max(x); break;
      default: break;
    }
  }
}

// Work-group-local atomics (tests/sycl/atomic.cpp's local-memory cases):
// every work-group folds its slice data[g*L+1 .. g*L+L-1] into an LDS value
// that starts at data[g*L], through atomic_ref<T, relaxed, work_group,
// local_space>, and work-item 0 writes the value to out[g].
template <class T>
__global__ void atomic_fold_local_kernel(int op, const T *data, T *out) {
  __shared__ T acc;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x;
  if (threadIdx.x == 0) acc = data[base];
  __syncthreads();
  if (threadIdx.x != 0) {
    const T x = data[base + threadIdx.x];
    acpp::atomic_ref<T, acpp::memory_order::relaxed, acpp::memory_scope::work_group,
                     acpp::address_space::local_space>
        r{acc};
    if constexpr (std::is_integral_v<T>) {
      switch (op) {
        case 0: r.fetch_and(x); break;
        case 1: r.fetch_or(x); break;
        case 2: r.fetch_xor(x); break;
        default: break;
      }
    }
    switch (op) {
      case 3: r.fetch_add(x); break;
      case 4: r.fetch_sub(x); break;
      case 5: r.fetch_min(x); break;
      case 6: r.fetch_max(x); break;
      default: break;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// atomic_ref on a pointer: every work-item advances *slot by one element
// (fetch_add(1) on an int*, so the address moves by sizeof(int)), and
// records the value it saw in seen[i].
__global__ void atomic_pointer_kernel(int **slot, int **seen, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  acpp::atomic_ref<int *, acpp::memory_order::relaxed, acpp::memory_scope::device,
                   acpp::address_space::global_space>
      r{*slot};
  seen[i] = r.fetch_add(1);
}

// 8- and 16-bit atomic builtins: work-items slots..n-1 apply the op with
// operand data[i] to data[i % slots].  With slots < 4 bytes' worth of
// elements, neighbouring elements of one 32-bit word are updated
// concurrently.  rmw op numbering as above.
#define TK_NARROW_SWITCH(tag)                                                      \
  switch (op) {                                                                    \
    case 0: __acpp_sscp_atomic_fetch_and_##tag(as, o, s, t, x); break;             \
    case 1: __acpp_sscp_atomic_fetch_or_##tag(as, o, s, t, x); break;              \
    case 2: __acpp_sscp_atomic_fetch_xor_##tag(as, o, s, t, x); break;             \
    case 3: __acpp_sscp_atomic_fetch_add_##tag(as, o, s, t, x); break;             \
    case 4: __acpp_sscp_atomic_fetch_sub_##tag(as, o, s, t, x); break;             \
    case 5: __acpp_sscp_atomic_fetch_min_##tag(as, o, s, t, x); break;             \
    default: __acpp_sscp_atomic_fetch_max_##tag(as, o, s, t, x); break;            \
  }
#define TK_NARROW_SWITCH_ARITH(tag)                                                \
  switch (op) {                                                                    \
    case 3: __acpp_sscp_atomic_fetch_add_##tag(as, o, s, t, x); break;             \
    case 4: __acpp_sscp_atomic_fetch_sub_##tag(as, o, s, t, x); break;             \
    case 5: __acpp_sscp_atomic_fetch_min_##tag(as, o, s, t, x); break;             \
    default: __acpp_sscp_atomic_fetch_max_##tag(as, o, s, t, x); break;            \
  }
template <class T>
__global__ void narrow_fold_kernel(int op, T *data, uint64_t n, uint32_t slots) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < slots || i >= n) return;
  const T x = data[i];
  T *t = data + (i % slots);
  const auto as = acpp::address_space::global_space;
  const auto o = acpp::memory_order::relaxed;
  const auto s = acpp::memory_scope::device;
  if constexpr (std::is_same_v<T, int8_t>) {
    TK_NARROW_SWITCH(i8)
  } else if constexpr (std::is_same_v<T, int16_t>) {
    TK_NARROW_SWITCH(i16)
  } else if constexpr (std::is_same_v<T, uint8_t>) {
    TK_NARROW_SWITCH_ARITH(u8)
  } else {
    TK_NARROW_SWITCH_ARITH(u16)
  }
}
#undef TK_NARROW_SWITCH
#undef TK_NARROW_SWITCH_ARITH

// One work-item: the tests/pstl/std_atomic.cpp sequence on data[k] of a
// 4-element array whose other elements must stay as they are: load + store
// (x + 1), exchange(1) + store(old + 1), and compare-exchange with a wrong
// and a right expected value.  prev[0..5] receive the observed values.
template <class T>
__global__ void narrow_single_kernel(T *data, int k, T operand, T *prev) {
  const auto as = acpp::address_space::global_space;
  const auto o = acpp::memory_order::seq_cst;
  const auto s = acpp::memory_scope::device;
  T *p = data + k;
  if constexpr (std::is_same_v<T, int8_t>) {
    const T a = __acpp_sscp_atomic_load_i8(as, o, s, p);
    __acpp_sscp_atomic_store_i8(as, o, s, p, static_cast<T>(a + 1));
    const T b = __acpp_sscp_atomic_exchange_i8(as, o, s, p, 1);
    __acpp_sscp_atomic_store_i8(as, o, s, p, static_cast<T>(b + 1));
    T e0 = 0, e1 = static_cast<T>(b + 1);
    const bool ok0 = __acpp_sscp_cmp_exch_strong_i8(as, o, o, s, p, &e0, operand);
    const bool ok1 = __acpp_sscp_cmp_exch_weak_i8(as, o, o, s, p, &e1, operand);
    prev[0] = a; prev[1] = b; prev[2] = e0; prev[3] = e1;
    prev[4] = static_cast<T>(ok0); prev[5] = static_cast<T>(ok1);
  } else {
    const T a = __acpp_sscp_atomic_load_i16(as, o, s, p);
    __acpp_sscp_atomic_store_i16(as, o, s, p, static_cast<T>(a + 1));
    const T b = __acpp_sscp_atomic_exchange_i16(as, o, s, p, 1);
    __acpp_sscp_atomic_store_i16(as, o, s, p, static_cast<T>(b + 1));
    T e0 = 0, e1 = static_cast<T>(b + 1);
    const bool ok0 = __acpp_sscp_cmp_exch_strong_i16(as, o, o, s, p, &e0, operand);
    const bool ok1 = __acpp_sscp_cmp_exch_weak_i16(as, o, o, s, p, &e1, operand);
    prev[0] = a; prev[1] = b; prev[2] = e0; prev[3] = e1;
    prev[4] = static_cast<T>(ok0); prev[5] = static_cast<T>(ok1);
  }
}

// tests/pstl/std_atomic.cpp:91-136: one work-item applies one operation and
// records the previous value; also load/store/exchange/compare_exchange.
template <class T>
__global__ void atomic_single_kernel(int op, T *data, T operand, T *prev) {
  acpp::atomic_ref<T, acpp::memory_order::relaxed, acpp::memory_scope::device> r{data[0]};
  T p = T{};
  switch (op) {
    case 0: p = r.fetch_and(operand); break;
    case 1: p = r.fetch_or(operand); break;
    case 2: p = r.fetch_xor(operand); break;
    case 3: p = r.fetch_add(operand); break;
    case 4: p = r.fetch_sub(operand); break;
    case 5: p = r.fetch_min(operand); break;
    case 6: p = r.fetch_max(operand); break;
    case 7: {  // load + store: x = x + 1
      p = r.load();
      r.store(p + 1);
      break;
    }
    case 8: {  // exchange(1), then store(old + current)
      p = r.exchange(1);
      r.store(p + r.load());
      break;
    }
    default: {  // compare_exchange: expected 0 fails, expected 42 succeeds
      T e0 = 0;
      bool ok0 = r.compare_exchange_strong(e0, operand);
      T e1 = 42;
      bool ok1 = r.compare_exchange_strong(e1, operand);
      prev[1] = e0;
      prev[2] = e1;
      prev[3] = static_cast<T>(ok0 ? 1 : 0);
      prev[4] = static_cast<T>(ok1 ? 1 : 0);
      p = e1;
      break;
    }
  }
  prev[0] = p;
}

template <class T>
int launch_collective(int kind, const void *in, void *out, uint64_t n, uint32_t local_size, int op,
                      const void *init, hipStream_t s) {
  if (local_size == 0 || local_size > 1024 || n % local_size) return 1;
  T iv = init ? *static_cast<const T *>(init) : T{};
  hipLaunchKernelGGL(collective_kernel<T>, dim3(static_cast<uint32_t>(n / local_size)), dim3(local_size),
                     0, s, kind, op, static_cast<const T *>(in), static_cast<T *>(out), iv);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

template <class T>
int launch_joint(int kind, const void *in, void *out, uint64_t n, uint64_t seg_len,
                 uint32_t local_size, int op, const void *init, hipStream_t s) {
  if (seg_len == 0 || n % seg_len || local_size == 0 || local_size > 1024) return 1;
  T iv = init ? *static_cast<const T *>(init) : T{};
  hipLaunchKernelGGL(joint_kernel<T>, dim3(static_cast<uint32_t>(n / seg_len)), dim3(local_size), 0, s,
                     kind, op, static_cast<const T *>(in), static_cast<T *>(out), seg_len, iv);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// Broadcast / predicates / data movement.  level 0: sub-group (wave), 1:
// work-group.  which: 0 broadcast(param), 1 shift_left(param),
// 2 shift_right(param), 3 permute_by_xor(param), 4 select_from_group(src[i]),
// 5 any_of(x > param), 6 all_of(x > param), 7 none_of(x > param),
// 8 joint_any_of / 9 joint_all_of / 10 joint_none_of over the work-group's
// slice of `in` (x > param).
template <class T>
__global__ void movement_kernel(int level, int which, uint32_t param, const T *in, const uint32_t *src,
                                T *out) {
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const T x = in[gid];
  const T thr = static_cast<T>(static_cast<int32_t>(param));  // thresholds may be negative
  auto gt = [thr](T v) { return v > thr; };
  T r{};
  if (level == 0) {
    acpp::sub_group g = acpp::this_sub_group();
    switch (which) {
      case 0: r = acpp::group_broadcast(g, x, param); break;
      case 1: r = acpp::shift_group_left(g, x, param); break;
      case 2: r = acpp::shift_group_right(g, x, param); break;
      case 3: r = acpp::permute_group_by_xor(g, x, param); break;
      case 4: r = acpp::select_from_group(g, x, src[gid]); break;
      case 5: r = acpp::any_of_group(g, x, gt) ? T(1) : T(0); break;
      case 6: r = acpp::all_of_group(g, x, gt) ? T(1) : T(0); break;
      case 7: r = acpp::none_of_group(g, x, gt) ? T(1) : T(0); break;
      default: {
        const T *first = in + static_cast<uint64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u);
        const T *last = first + g.get_local_linear_range();
        bool b = which == 8 ? acpp::joint_any_of(g, first, last, gt)
                            : which == 9 ? acpp::joint_all_of(g, first, last, gt)
                                         : acpp::joint_none_of(g, first, last, gt);
        r = b ? T(1) : T(0);
      }
    }
  } else {
    acpp::group g = acpp::this_group();
    switch (which) {
      case 0: r = acpp::group_broadcast(g, x, param); break;
      case 1: r = acpp::shift_group_left(g, x, param); break;
      case 2: r = acpp::shift_group_right(g, x, param); break;
      case 3: r = acpp::permute_group_by_xor(g, x, param); break;
      case 4: r = acpp::select_from_group(g, x, src[gid]); break;
      case 5: r = acpp::any_of_group(g, x, gt) ? T(1) : T(0); break;
      case 6: r = acpp::all_of_group(g, x, gt) ? T(1) : T(0); break;
      case 7: r = acpp::none_of_group(g, x, gt) ? T(1) : T(0); break;
      default: {
        const T *first = in + static_cast<uint64_t>(blockIdx.x) * blockDim.x;
        const T *last = first + blockDim.x;
        bool b = which == 8 ? acpp::joint_any_of(g, first, last, gt)
                            : which == 9 ? acpp::joint_all_of(g, first, last, gt)
                                         : acpp::joint_none_of(g, first, last, gt);
        r = b ? T(1) : T(0);
      }
    }
  }
  out[gid] = r;
}

template <class T>
int launch_movement(int level, int which, uint32_t param, const void *in, const void *src, void *out,
                    uint64_t n, uint32_t local_size, hipStream_t s) {
  if (local_size == 0 || local_size > 1024 || n % local_size != 0) return 1;
  hipLaunchKernelGGL(movement_kernel<T>, dim3(static_cast<uint32_t>(n / local_size)), dim3(local_size), 0, s,
                     level, which, param, static_cast<const T *>(in), static_cast<const uint32_t *>(src),
                     static_cast<T *>(out));
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // namespace

// A user kernel with a reduction: grid-stride per-work-item reducers, then
// combine_into.  The reducer's named operator (+=, *=, &=, |=, ^=) is used
// where it has one; `user_identity` starts the reducers at an explicit
// identity instead of the known one.
template <class T, class Op>
__device__ void reducer_step(acpp::reducer<T, Op> &r, const T &v) {
  if constexpr (std::is_same_v<Op, acpp::plus<T>>)
    r += v;
  else if constexpr (std::is_same_v<Op, acpp::multiplies<T>>)
    r *= v;
  else if constexpr (std::is_integral_v<T> && std::is_same_v<Op, acpp::bit_and<T>>)
    r &= v;
  else if constexpr (std::is_integral_v<T> && std::is_same_v<Op, acpp::bit_or<T>>)
    r |= v;
  else if constexpr (std::is_integral_v<T> && std::is_same_v<Op, acpp::bit_xor<T>>)
    r ^= v;
  else
    r.combine(v);
}

template <class T, class Op>
__global__ void reducer_kernel(const T *x, uint64_t n, T *out, int user_identity) {
  acpp::reducer<T, Op> r = user_identity ? acpp::reducer<T, Op>(acpp::known_identity<Op, T>::get())
                                         : acpp::reducer<T, Op>();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    reducer_step(r, x[i]);
  acpp::combine_into(acpp::this_group(), r, out);
}

template <class T>
int launch_reducer(int op, const void *x, uint64_t n, void *out, uint32_t groups, uint32_t local,
                   int user_identity, hipStream_t s) {
  if (local == 0 || local > 1024 || local % 64 != 0 || groups == 0) return 1;
  const T *xp = static_cast<const T *>(x);
  T *op_ = static_cast<T *>(out);
#define ACPP_TK_REDUCER(OP)                                                                        \
  hipLaunchKernelGGL((reducer_kernel<T, OP<T>>), dim3(groups), dim3(local), 0, s, xp, n, op_, user_identity); \
  break;
  switch (op) {
    case 0: ACPP_TK_REDUCER(acpp::plus)
    case 1: ACPP_TK_REDUCER(acpp::multiplies)
    case 2: ACPP_TK_REDUCER(acpp::minimum)
    case 3: ACPP_TK_REDUCER(acpp::maximum)
    case 7: ACPP_TK_REDUCER(acpp::logical_and)
    case 8: ACPP_TK_REDUCER(acpp::logical_or)
    default:
      if constexpr (std::is_integral_v<T>) {
        switch (op) {
          case 4: ACPP_TK_REDUCER(acpp::bit_and)
          case 5: ACPP_TK_REDUCER(acpp::bit_or)
          case 6: ACPP_TK_REDUCER(acpp::bit_xor)
          default: return 1;
        }
        break;
      } else {
        return 1;
      }
  }
#undef ACPP_TK_REDUCER
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

extern "C" {

// tag: 0 i8, 1 i32, 2 i64, 3 u32, 4 f32, 5 f64, 6 f16
int tk_group_collective(int tag, int kind, const void *in, void *out, uint64_t n,
                        uint32_t local_size, int op, const void *init, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (tag) {
    case 0: return launch_collective<int8_t>(kind, in, out, n, local_size, op, init, s);
    case 1: return launch_collective<int32_t>(kind, in, out, n, local_size, op, init, s);
    case 2: return launch_collective<int64_t>(kind, in, out, n, local_size, op, init, s);
    case 3: return launch_collective<uint32_t>(kind, in, out, n, local_size, op, init, s);
    case 4: return launch_collective<float>(kind, in, out, n, local_size, op, init, s);
    case 5: return launch_collective<double>(kind, in, out, n, local_size, op, init, s);
    case 6: return launch_collective<_Float16>(kind, in, out, n, local_size, op, init, s);
    default: return 1;
  }
}

int tk_joint(int tag, int kind, const void *in, void *out, uint64_t n, uint64_t seg_len,
             uint32_t local_size, int op, const void *init, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (tag) {
    case 1: return launch_joint<int32_t>(kind, in, out, n, seg_len, local_size, op, init, s);
    case 2: return launch_joint<int64_t>(kind, in, out, n, seg_len, local_size, op, init, s);
    case 4: return launch_joint<float>(kind, in, out, n, seg_len, local_size, op, init, s);
    default: return 1;
  }
}

// tag: 1 i32, 2 i64, 4 f32, 5 f64; n a multiple of local_size
int tk_group_movement(int tag, int level, int which, uint32_t param, const void *in, const void *src,
                      void *out, uint64_t n, uint32_t local_size, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (tag) {
    case 1: return launch_movement<int32_t>(level, which, param, in, src, out, n, local_size, s);
    case 2: return launch_movement<int64_t>(level, which, param, in, src, out, n, local_size, s);
    case 4: return launch_movement<float>(level, which, param, in, src, out, n, local_size, s);
    case 5: return launch_movement<double>(level, which, param, in, src, out, n, local_size, s);
    default: return 1;
  }
}

int tk_babelstream_dot(const int32_t *a, const int32_t *b, int32_t *sums, uint64_t n,
                       uint32_t num_groups, uint32_t local_size, void *stream) {
  if (local_size == 0 || local_size > 1024 || num_groups == 0) return 1;
  hipLaunchKernelGGL(babelstream_dot_kernel, dim3(num_groups), dim3(local_size), 0,
                     static_cast<hipStream_t>(stream), a, b, sums, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// tag: 1 i32, 2 i64, 3 u32, 4 f32, 5 f64, 6 u64; abi: use the extern "C"
// builtins (i32 and f32 only)
int tk_atomic_fold(int tag, int op, void *data, uint64_t n, int abi, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<uint32_t>((n + 255) / 256)), block(256);
  if (abi) {
    if (tag == 1) hipLaunchKernelGGL((atomic_fold_kernel<int32_t, true>), grid, block, 0, s, op, static_cast<int32_t *>(data), n);
    else if (tag == 4) hipLaunchKernelGGL((atomic_fold_kernel<float, true>), grid, block, 0, s, op, static_cast<float *>(data), n);
    else return 1;
  } else {
    switch (tag) {
      case 1: hipLaunchKernelGGL((atomic_fold_kernel<int, false>), grid, block, 0, s, op, static_cast<int *>(data), n); break;
      case 2: hipLaunchKernelGGL((atomic_fold_kernel<long long, false>), grid, block, 0, s, op, static_cast<long long *>(data), n); break;
      case 3: hipLaunchKernelGGL((atomic_fold_kernel<unsigned int, false>), grid, block, 0, s, op, static_cast<unsigned int *>(data), n); break;
      case 4: hipLaunchKernelGGL((atomic_fold_kernel<float, false>), grid, block, 0, s, op, static_cast<float *>(data), n); break;
      case 5: hipLaunchKernelGGL((atomic_fold_kernel<double, false>), grid, block, 0, s, op, static_cast<double *>(data), n); break;
      case 6: hipLaunchKernelGGL((atomic_fold_kernel<unsigned long long, false>), grid, block, 0, s, op, static_cast<unsigned long long *>(data), n); break;
      default: return 1;
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int tk_atomic_single(int tag, int op, void *data, const void *operand, void *prev, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (tag) {
    case 1: hipLaunchKernelGGL(atomic_single_kernel<int>, dim3(1), dim3(1), 0, s, op, static_cast<int *>(data), *static_cast<const int *>(operand), static_cast<int *>(prev)); break;
    case 2: hipLaunchKernelGGL(atomic_single_kernel<long long>, dim3(1), dim3(1), 0, s, op, static_cast<long long *>(data), *static_cast<const long long *>(operand), static_cast<long long *>(prev)); break;
    case 3: hipLaunchKernelGGL(atomic_single_kernel<unsigned int>, dim3(1), dim3(1), 0, s, op, static_cast<unsigned int *>(data), *static_cast<const unsigned int *>(operand), static_cast<unsigned int *>(prev)); break;
    case 6: hipLaunchKernelGGL(atomic_single_kernel<unsigned long long>, dim3(1), dim3(1), 0, s, op, static_cast<unsigned long long *>(data), *static_cast<const unsigned long long *>(operand), static_cast<unsigned long long *>(prev)); break;
    default: return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// Local-memory atomic_ref folds.  tag: 1 i32, 2 i64, 3 u32, 4 f32, 5 f64, 6 u64;
// groups x local elements, one result per group.
int tk_atomic_fold_local(int tag, int op, const void *data, void *out, uint32_t groups, uint32_t local,
                         void *stream) {
  if (local == 0 || local > 1024 || groups == 0) return 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
#define ACPP_TK_LOCAL(T)                                                                                   \
  hipLaunchKernelGGL(atomic_fold_local_kernel<T>, dim3(groups), dim3(local), 0, s, op,                     \
                     static_cast<const T *>(data), static_cast<T *>(out));                                 \
  break;
  switch (tag) {
    case 1: ACPP_TK_LOCAL(int32_t)
    case 2: ACPP_TK_LOCAL(int64_t)
    case 3: ACPP_TK_LOCAL(uint32_t)
    case 4: ACPP_TK_LOCAL(float)
    case 5: ACPP_TK_LOCAL(double)
    case 6: ACPP_TK_LOCAL(uint64_t)
    default: return 1;
  }
#undef ACPP_TK_LOCAL
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int tk_atomic_pointer(void *slot, void *seen, uint64_t n, void *stream) {
  const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
  hipLaunchKernelGGL(atomic_pointer_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<int **>(slot), static_cast<int **>(seen), n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

// acpp::reducer + combine_into.  tag: 1 i32, 2 i64, 3 u32, 4 f32, 5 f64, 6 u64.
int tk_reducer(int tag, int op, const void *x, uint64_t n, void *out, uint32_t groups, uint32_t local,
               int user_identity, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (tag) {
    case 1: return launch_reducer<int32_t>(op, x, n, out, groups, local, user_identity, s);
    case 2: return launch_reducer<int64_t>(op, x, n, out, groups, local, user_identity, s);
    case 3: return launch_reducer<uint32_t>(op, x, n, out, groups, local, user_identity, s);
    case 4: return launch_reducer<float>(op, x, n, out, groups, local, user_identity, s);
    case 5: return launch_reducer<double>(op, x, n, out, groups, local, user_identity, s);
    case 6: return launch_reducer<uint64_t>(op, x, n, out, groups, local, user_identity, s);
    default: return 1;
  }
}

// 8/16-bit builtins.  tag: 7 i8, 8 i16, 9 u8, 10 u16.
int tk_atomic_fold_narrow(int tag, int op, void *data, uint64_t n, uint32_t slots, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 block(256), grid(static_cast<uint32_t>((n + 255) / 256));
  switch (tag) {
    case 7: hipLaunchKernelGGL(narrow_fold_kernel<int8_t>, grid, block, 0, s, op, static_cast<int8_t *>(data), n, slots); break;
    case 8: hipLaunchKernelGGL(narrow_fold_kernel<int16_t>, grid, block, 0, s, op, static_cast<int16_t *>(data), n, slots); break;
    case 9: hipLaunchKernelGGL(narrow_fold_kernel<uint8_t>, grid, block, 0, s, op, static_cast<uint8_t *>(data), n, slots); break;
    case 10: hipLaunchKernelGGL(narrow_fold_kernel<uint16_t>, grid, block, 0, s, op, static_cast<uint16_t *>(data), n, slots); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int tk_atomic_single_narrow(int tag, void *data, int k, int operand, void *prev, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (k < 0 || k > 3) return -1;
  if (tag == 7)
    hipLaunchKernelGGL(narrow_single_kernel<int8_t>, dim3(1), dim3(1), 0, s, static_cast<int8_t *>(data), k,
                       static_cast<int8_t>(operand), static_cast<int8_t *>(prev));
  else if (tag == 8)
    hipLaunchKernelGGL(narrow_single_kernel<int16_t>, dim3(1), dim3(1), 0, s, static_cast<int16_t *>(data), k,
                       static_cast<int16_t>(operand), static_cast<int16_t *>(prev));
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
