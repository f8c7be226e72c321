// atomic_ref for HIP kernels on gfx950, and the extern "C" atomic builtin
// ABI that stdpar's std::atomic lowering targets.
//
// Replaces, for user kernels:
//   include/hipSYCL/sycl/libkernel/atomic_ref.hpp:45-303 (sycl::atomic_ref)
//   include/hipSYCL/sycl/libkernel/generic/hiplike/atomic_builtins.hpp:438-598
//     (HIP path: atomicAdd with order/scope ignored, AGENT-scope min/max,
//      CAS loops for fp min/max/sub)
//   include/hipSYCL/sycl/libkernel/sscp/builtins/atomic.hpp:16-327 and
//   src/libkernel/sscp/amdgpu/atomic.cpp:51-64, 335-547 (the SSCP ABI
//     __acpp_sscp_atomic_<op>_<t>(address_space, order, scope, ptr, x)).
//
// Every operation maps onto a clang __hip_atomic_* builtin with the
// requested memory order and a real memory scope (work-item ... system), so
// the compiler emits the gfx950 cache-coherence bits for that scope (for
// example, device scope is coherent across the XCDs' L2 caches).  Integer
// and floating-point add, min and max are native read-modify-writes where
// gfx950 has them.
//
// Enumerations use the numbering of the reference's builtin ABI
// (src/compiler/sscp/StdAtomicRemapperPass.cpp:27-52).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

namespace acpp {

enum class memory_order : int { relaxed = 0, acquire = 1, release = 2, acq_rel = 3, seq_cst = 4 };
enum class memory_scope : int { work_item = 0, sub_group = 1, work_group = 2, device = 3, system = 4 };
enum class address_space : int {
  global_space = 0,
  local_space = 1,
  constant_space = 2,
  private_space = 3,
  generic_space = 4
};

namespace detail {

// The clang builtins take the order and scope as constants: expand the
// (order, scope) pair into a switch that the optimiser folds away when the
// arguments are compile-time constants, as they are for atomic_ref's
// defaults.
#define ACPP_ATOMIC_SCOPE_SWITCH(scope, ORDER, EXPR)                          \
  switch (scope) {                                                           \
    case memory_scope::work_item: {                                          \
      constexpr int S = __HIP_MEMORY_SCOPE_SINGLETHREAD;                     \
      return EXPR(ORDER, S);                                                 \
    }                                                                        \
    case memory_scope::sub_group: {                                          \
      constexpr int S = __HIP_MEMORY_SCOPE_WAVEFRONT;                        \
      return EXPR(ORDER, S);                                                 \
    }                                                                        \
    case memory_scope::work_group: {                                         \
      constexpr int S = __HIP_MEMORY_SCOPE_WORKGROUP;                        \
      return EXPR(ORDER, S);                                                 \
    }                                                                        \
    case memory_scope::device: {                                             \
      constexpr int S = __HIP_MEMORY_SCOPE_AGENT;                            \
      return EXPR(ORDER, S);                                                 \
    }                                                                        \
    default: {                                                               \
      constexpr int S = __HIP_MEMORY_SCOPE_SYSTEM;                           \
      return EXPR(ORDER, S);                                                 \
    }                                                                        \
  }

#define ACPP_ATOMIC_ORDER_SCOPE_SWITCH(order, scope, EXPR)                   \
  switch (order) {                                                           \
    case memory_order::relaxed: ACPP_ATOMIC_SCOPE_SWITCH(scope, __ATOMIC_RELAXED, EXPR) \
    case memory_order::acquire: ACPP_ATOMIC_SCOPE_SWITCH(scope, __ATOMIC_ACQUIRE, EXPR) \
    case memory_order::release: ACPP_ATOMIC_SCOPE_SWITCH(scope, __ATOMIC_RELEASE, EXPR) \
    case memory_order::acq_rel: ACPP_ATOMIC_SCOPE_SWITCH(scope, __ATOMIC_ACQ_REL, EXPR) \
    default: ACPP_ATOMIC_SCOPE_SWITCH(scope, __ATOMIC_SEQ_CST, EXPR)         \
  }

// Loads may not use release orders; stores may not use acquire orders.
__host__ __device__ constexpr memory_order load_order(memory_order o) {
  return (o == memory_order::release) ? memory_order::relaxed
         : (o == memory_order::acq_rel) ? memory_order::acquire
                                        : o;
}
__host__ __device__ constexpr memory_order store_order(memory_order o) {
  return (o == memory_order::acquire) ? memory_order::relaxed
         : (o == memory_order::acq_rel) ? memory_order::release
                                        : o;
}

template <class T>
__device__ __forceinline__ T atomic_load(T *p, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_load(p, O, S)
  switch (load_order(o)) {
    case memory_order::relaxed: ACPP_ATOMIC_SCOPE_SWITCH(s, __ATOMIC_RELAXED, ACPP_E)
    case memory_order::acquire: ACPP_ATOMIC_SCOPE_SWITCH(s, __ATOMIC_ACQUIRE, ACPP_E)
    default: ACPP_ATOMIC_SCOPE_SWITCH(s, __ATOMIC_SEQ_CST, ACPP_E)
  }
#undef ACPP_E
}

template <class T>
__device__ __forceinline__ void atomic_store(T *p, T v, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_store(p, v, O, S)
  switch (store_order(o)) {
    case memory_order::relaxed: ACPP_ATOMIC_SCOPE_SWITCH(s, __ATOMIC_RELAXED, ACPP_E)
    case memory_order::release: ACPP_ATOMIC_SCOPE_SWITCH(s, __ATOMIC_RELEASE, ACPP_E)
    default: ACPP_ATOMIC_SCOPE_SWITCH(s, __ATOMIC_SEQ_CST, ACPP_E)
  }
#undef ACPP_E
}

template <class T>
__device__ __forceinline__ T atomic_exchange(T *p, T v, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_exchange(p, v, O, S)
  ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
}

template <class T>
__device__ __forceinline__ bool atomic_cas(T *p, T &expected, T desired, memory_order success,
                                           memory_order failure, memory_scope s, bool weak) {
  // Failure order: no stronger than success and never a release order.
  (void)failure;
  if (weak) {
#define ACPP_E(O, S) __hip_atomic_compare_exchange_weak(p, &expected, desired, O, __ATOMIC_RELAXED, S)
    ACPP_ATOMIC_ORDER_SCOPE_SWITCH(success, s, ACPP_E)
#undef ACPP_E
  } else {
#define ACPP_E(O, S) __hip_atomic_compare_exchange_strong(p, &expected, desired, O, __ATOMIC_RELAXED, S)
    ACPP_ATOMIC_ORDER_SCOPE_SWITCH(success, s, ACPP_E)
#undef ACPP_E
  }
}

template <class T>
__device__ __forceinline__ T atomic_fetch_add(T *p, T v, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_fetch_add(p, v, O, S)
  ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
}

template <class T>
__device__ __forceinline__ T atomic_fetch_sub(T *p, T v, memory_order o, memory_scope s) {
  if constexpr (std::is_floating_point_v<T>) {
    return atomic_fetch_add(p, -v, o, s);
  } else {
#define ACPP_E(O, S) __hip_atomic_fetch_sub(p, v, O, S)
    ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
  }
}

template <class T>
__device__ __forceinline__ T atomic_fetch_and(T *p, T v, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_fetch_and(p, v, O, S)
  ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
}

template <class T>
__device__ __forceinline__ T atomic_fetch_or(T *p, T v, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_fetch_or(p, v, O, S)
  ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
}

template <class T>
__device__ __forceinline__ T atomic_fetch_xor(T *p, T v, memory_order o, memory_scope s) {
#define ACPP_E(O, S) __hip_atomic_fetch_xor(p, v, O, S)
  ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
}

// min/max return the previous value (the reference's host backend returns
// the operand instead when it stored it, host/atomic_builtins.hpp:291-312;
// the HIP path and the C++ semantics return the previous value).
template <class T>
__device__ __forceinline__ T atomic_fetch_min(T *p, T v, memory_order o, memory_scope s) {
  if constexpr (std::is_floating_point_v<T>) {
    // Compare-and-swap loop with the (old < x ? keep : store) rule of
    // sycl::minimum, so NaN handling matches the reference.
    T old = atomic_load(p, memory_order::relaxed, s);
    while (!(old < v) && old != v) {
      if (atomic_cas(p, old, v, o, memory_order::relaxed, s, false)) break;
    }
    return old;
  } else {
#define ACPP_E(O, S) __hip_atomic_fetch_min(p, v, O, S)
    ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
  }
}

template <class T>
__device__ __forceinline__ T atomic_fetch_max(T *p, T v, memory_order o, memory_scope s) {
  if constexpr (std::is_floating_point_v<T>) {
    T old = atomic_load(p, memory_order::relaxed, s);
    while (!(old > v) && old != v) {
      if (atomic_cas(p, old, v, o, memory_order::relaxed, s, false)) break;
    }
    return old;
  } else {
#define ACPP_E(O, S) __hip_atomic_fetch_max(p, v, O, S)
    ACPP_ATOMIC_ORDER_SCOPE_SWITCH(o, s, ACPP_E)
#undef ACPP_E
  }
}

}  // namespace detail

// sycl::atomic_ref (atomic_ref.hpp:45-303).
template <class T, memory_order DefaultOrder, memory_scope DefaultScope,
          address_space Space = address_space::generic_space>
class atomic_ref {
  static_assert(std::is_same_v<T, int> || std::is_same_v<T, unsigned int> ||
                    std::is_same_v<T, long> || std::is_same_v<T, unsigned long> ||
                    std::is_same_v<T, long long> || std::is_same_v<T, unsigned long long> ||
                    std::is_same_v<T, float> || std::is_same_v<T, double> ||
                    std::is_pointer_v<T>,
                "Invalid data type for atomic_ref");
  static_assert(Space == address_space::generic_space || Space == address_space::global_space ||
                    Space == address_space::local_space,
                "Invalid address space for atomic_ref");

  using storage_t = std::conditional_t<std::is_pointer_v<T>, uintptr_t, T>;

 public:
  using value_type = T;
  using difference_type = std::conditional_t<std::is_pointer_v<T>, ptrdiff_t, T>;
  static constexpr size_t required_alignment = alignof(T);
  static constexpr bool is_always_lock_free = true;
  static constexpr memory_order default_read_order = detail::load_order(DefaultOrder);
  static constexpr memory_order default_write_order = detail::store_order(DefaultOrder);
  static constexpr memory_order default_read_modify_write_order = DefaultOrder;
  static constexpr memory_scope default_scope = DefaultScope;

  __device__ explicit atomic_ref(T &ref) : ptr_(reinterpret_cast<storage_t *>(&ref)) {}
  __device__ atomic_ref(const atomic_ref &) = default;

  __device__ bool is_lock_free() const { return true; }

  __device__ void store(T v, memory_order o = default_write_order,
                        memory_scope s = default_scope) const {
    detail::atomic_store(ptr_, to_storage(v), o, s);
  }
  __device__ T operator=(T v) const {
    store(v);
    return v;
  }
  __device__ T load(memory_order o = default_read_order, memory_scope s = default_scope) const {
    return from_storage(detail::atomic_load(ptr_, o, s));
  }
  __device__ operator T() const { return load(); }
  __device__ T exchange(T v, memory_order o = default_read_modify_write_order,
                        memory_scope s = default_scope) const {
    return from_storage(detail::atomic_exchange(ptr_, to_storage(v), o, s));
  }
  __device__ bool compare_exchange_weak(T &expected, T desired, memory_order success,
                                        memory_order failure,
                                        memory_scope s = default_scope) const {
    storage_t e = to_storage(expected);
    bool ok = detail::atomic_cas(ptr_, e, to_storage(desired), success, failure, s, true);
    expected = from_storage(e);
    return ok;
  }
  __device__ bool compare_exchange_weak(T &expected, T desired,
                                        memory_order o = default_read_modify_write_order,
                                        memory_scope s = default_scope) const {
    return compare_exchange_weak(expected, desired, o, o, s);
  }
  __device__ bool compare_exchange_strong(T &expected, T desired, memory_order success,
                                          memory_order failure,
                                          memory_scope s = default_scope) const {
    storage_t e = to_storage(expected);
    bool ok = detail::atomic_cas(ptr_, e, to_storage(desired), success, failure, s, false);
    expected = from_storage(e);
    return ok;
  }
  __device__ bool compare_exchange_strong(T &expected, T desired,
                                          memory_order o = default_read_modify_write_order,
                                          memory_scope s = default_scope) const {
    return compare_exchange_strong(expected, desired, o, o, s);
  }

  // Arithmetic (integers, floating point; pointers by difference_type).
  __device__ T fetch_add(difference_type v, memory_order o = default_read_modify_write_order,
                         memory_scope s = default_scope) const {
    if constexpr (std::is_pointer_v<T>) {
      return from_storage(detail::atomic_fetch_add(
          ptr_, static_cast<storage_t>(v * static_cast<ptrdiff_t>(sizeof(*T{}))), o, s));
    } else {
      return detail::atomic_fetch_add(ptr_, v, o, s);
    }
  }
  __device__ T fetch_sub(difference_type v, memory_order o = default_read_modify_write_order,
                         memory_scope s = default_scope) const {
    if constexpr (std::is_pointer_v<T>) {
      return from_storage(detail::atomic_fetch_sub(
          ptr_, static_cast<storage_t>(v * static_cast<ptrdiff_t>(sizeof(*T{}))), o, s));
    } else {
      return detail::atomic_fetch_sub(ptr_, v, o, s);
    }
  }
  template <class U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
  __device__ T fetch_and(T v, memory_order o = default_read_modify_write_order,
                         memory_scope s = default_scope) const {
    return detail::atomic_fetch_and(ptr_, v, o, s);
  }
  template <class U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
  __device__ T fetch_or(T v, memory_order o = default_read_modify_write_order,
                        memory_scope s = default_scope) const {
    return detail::atomic_fetch_or(ptr_, v, o, s);
  }
  template <class U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
  __device__ T fetch_xor(T v, memory_order o = default_read_modify_write_order,
                         memory_scope s = default_scope) const {
    return detail::atomic_fetch_xor(ptr_, v, o, s);
  }
  template <class U = T, std::enable_if_t<std::is_arithmetic_v<U>, int> = 0>
  __device__ T fetch_min(T v, memory_order o = default_read_modify_write_order,
                         memory_scope s = default_scope) const {
    return detail::atomic_fetch_min(ptr_, v, o, s);
  }
  template <class U = T, std::enable_if_t<std::is_arithmetic_v<U>, int> = 0>
  __device__ T fetch_max(T v, memory_order o = default_read_modify_write_order,
                         memory_scope s = default_scope) const {
    return detail::atomic_fetch_max(ptr_, v, o, s);
  }

  __device__ T operator++(int) const { return fetch_add(difference_type(1)); }
  __device__ T operator--(int) const { return fetch_sub(difference_type(1)); }
  __device__ T operator++() const { return fetch_add(difference_type(1)) + difference_type(1); }
  __device__ T operator--() const { return fetch_sub(difference_type(1)) - difference_type(1); }
  __device__ T operator+=(difference_type v) const { return fetch_add(v) + v; }
  __device__ T operator-=(difference_type v) const { return fetch_sub(v) - v; }
  template <class U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
  __device__ T operator&=(T v) const { return fetch_and(v) & v; }
  template <class U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
  __device__ T operator|=(T v) const { return fetch_or(v) | v; }
  template <class U = T, std::enable_if_t<std::is_integral_v<U>, int> = 0>
  __device__ T operator^=(T v) const { return fetch_xor(v) ^ v; }

 private:
  __device__ static storage_t to_storage(T v) {
    if constexpr (std::is_pointer_v<T>)
      return reinterpret_cast<storage_t>(v);
    else
      return v;
  }
  __device__ static T from_storage(storage_t v) {
    if constexpr (std::is_pointer_v<T>)
      return reinterpret_cast<T>(v);
    else
      return v;
  }
  storage_t *ptr_;
};

}  // namespace acpp

// ---------------------------------------------------------------------------
// extern "C" atomic builtin ABI (sscp/builtins/atomic.hpp:16-327), the
// target of stdpar's std::atomic remapping
// (src/compiler/sscp/StdAtomicRemapperPass.cpp:149-228).  The address-space
// argument does not change the code generated for gfx950 (flat addressing
// covers global and LDS) and is accepted for ABI compatibility.
// ---------------------------------------------------------------------------
typedef acpp::address_space __acpp_sscp_address_space;
typedef acpp::memory_order __acpp_sscp_memory_order;
typedef acpp::memory_scope __acpp_sscp_memory_scope;

#define ACPP_SSCP_ATOMIC_RMW(opname, tag, T)                                                      \
  extern "C" __device__ __attribute__((always_inline)) inline T __acpp_sscp_atomic_##opname##_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr, T x) {                                                                              \
    return acpp::detail::atomic_##opname(ptr, x, order, scope);                                   \
  }

#define ACPP_SSCP_ATOMIC_COMMON(tag, T)                                                           \
  extern "C" __device__ __attribute__((always_inline)) inline void __acpp_sscp_atomic_store_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr, T x) {                                                                              \
    acpp::detail::atomic_store(ptr, x, order, scope);                                             \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline T __acpp_sscp_atomic_load_##tag(    \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr) {                                                                                   \
    return acpp::detail::atomic_load(ptr, order, scope);                                          \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline T __acpp_sscp_atomic_exchange_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr, T x) {                                                                              \
    return acpp::detail::atomic_exchange(ptr, x, order, scope);                                   \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline bool __acpp_sscp_cmp_exch_strong_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order success,                                \
      __acpp_sscp_memory_order failure, __acpp_sscp_memory_scope scope, T *ptr, T *expected,      \
      T desired) {                                                                                \
    return acpp::detail::atomic_cas(ptr, *expected, desired, success, failure, scope, false);     \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline bool __acpp_sscp_cmp_exch_weak_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order success,                                \
      __acpp_sscp_memory_order failure, __acpp_sscp_memory_scope scope, T *ptr, T *expected,      \
      T desired) {                                                                                \
    return acpp::detail::atomic_cas(ptr, *expected, desired, success, failure, scope, true);      \
  }                                                                                               \
  ACPP_SSCP_ATOMIC_RMW(fetch_add, tag, T)                                                         \
  ACPP_SSCP_ATOMIC_RMW(fetch_sub, tag, T)                                                         \
  ACPP_SSCP_ATOMIC_RMW(fetch_min, tag, T)                                                         \
  ACPP_SSCP_ATOMIC_RMW(fetch_max, tag, T)

#define ACPP_SSCP_ATOMIC_INT(tag, T)      \
  ACPP_SSCP_ATOMIC_COMMON(tag, T)         \
  ACPP_SSCP_ATOMIC_RMW(fetch_and, tag, T) \
  ACPP_SSCP_ATOMIC_RMW(fetch_or, tag, T)  \
  ACPP_SSCP_ATOMIC_RMW(fetch_xor, tag, T)

ACPP_SSCP_ATOMIC_INT(i32, int32_t)
ACPP_SSCP_ATOMIC_INT(i64, int64_t)
ACPP_SSCP_ATOMIC_INT(u32, uint32_t)
ACPP_SSCP_ATOMIC_INT(u64, uint64_t)
ACPP_SSCP_ATOMIC_COMMON(f32, float)
ACPP_SSCP_ATOMIC_COMMON(f64, double)

// ---------------------------------------------------------------------------
// 8- and 16-bit entry points (sscp/builtins/atomic.hpp:16-311: store, load,
// exchange, compare-exchange and fetch and/or/xor/add/sub/min/max for i8 and
// i16; fetch add/sub/min/max for u8 and u16).  gfx950 has no 8- or 16-bit
// atomic read-modify-write, so each is a compare-and-swap loop on the
// aligned 32-bit word that holds the element: the other bytes of the word
// are carried over unchanged, so neighbouring elements may be updated
// concurrently.  Loads read the whole word.
// ---------------------------------------------------------------------------
namespace acpp {
namespace detail {

template <class T>
struct subword {
  static_assert(sizeof(T) == 1 || sizeof(T) == 2, "8- or 16-bit elements");
  using U = std::make_unsigned_t<T>;
  uint32_t *word;
  uint32_t shift;
  uint32_t mask;
  __device__ explicit subword(T *p) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    word = reinterpret_cast<uint32_t *>(a & ~static_cast<uintptr_t>(3));
    shift = static_cast<uint32_t>(a & 3) * 8;
    mask = ((sizeof(T) == 1) ? 0xffu : 0xffffu) << shift;
  }
  __device__ T get(uint32_t w) const { return static_cast<T>(static_cast<U>((w & mask) >> shift)); }
  __device__ uint32_t put(uint32_t w, T v) const {
    return (w & ~mask) | ((static_cast<uint32_t>(static_cast<U>(v)) << shift) & mask);
  }
};

// Replaces the element by f(old) atomically; returns old.
template <class T, class F>
__device__ __forceinline__ T subword_rmw(T *p, F f, memory_order o, memory_scope s) {
  const subword<T> sw(p);
  uint32_t w = atomic_load(sw.word, memory_order::relaxed, s);
  for (;;) {
    const T old = sw.get(w);
    if (atomic_cas(sw.word, w, sw.put(w, f(old)), o, memory_order::relaxed, s, false)) return old;
  }
}

template <class T>
__device__ __forceinline__ T subword_load(T *p, memory_order o, memory_scope s) {
  const subword<T> sw(p);
  return sw.get(atomic_load(sw.word, o, s));
}

// Compare-exchange of one element.  Fails (and updates `expected`) only if
// the element differs from `expected`; a change to the word's other bytes
// makes it retry, so the weak form behaves like the strong one.
template <class T>
__device__ __forceinline__ bool subword_cas(T *p, T &expected, T desired, memory_order o, memory_scope s) {
  const subword<T> sw(p);
  uint32_t w = atomic_load(sw.word, memory_order::relaxed, s);
  for (;;) {
    const T cur = sw.get(w);
    if (cur != expected) {
      expected = cur;
      return false;
    }
    if (atomic_cas(sw.word, w, sw.put(w, desired), o, memory_order::relaxed, s, false)) return true;
  }
}

}  // namespace detail
}  // namespace acpp

#define ACPP_SSCP_SUBWORD_RMW(opname, tag, T, EXPR)                                               \
  extern "C" __device__ __attribute__((always_inline)) inline T __acpp_sscp_atomic_##opname##_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr, T x) {                                                                              \
    return acpp::detail::subword_rmw(ptr, [x](T a) -> T { return EXPR; }, order, scope);        \
  }

#define ACPP_SSCP_SUBWORD_ARITH(tag, T)                                                          \
  ACPP_SSCP_SUBWORD_RMW(fetch_add, tag, T, static_cast<T>(a + x))                                 \
  ACPP_SSCP_SUBWORD_RMW(fetch_sub, tag, T, static_cast<T>(a - x))                                 \
  ACPP_SSCP_SUBWORD_RMW(fetch_min, tag, T, (x < a) ? x : a)                                       \
  ACPP_SSCP_SUBWORD_RMW(fetch_max, tag, T, (x > a) ? x : a)

#define ACPP_SSCP_SUBWORD_FULL(tag, T)                                                           \
  ACPP_SSCP_SUBWORD_ARITH(tag, T)                                                                 \
  ACPP_SSCP_SUBWORD_RMW(fetch_and, tag, T, static_cast<T>(a & x))                                 \
  ACPP_SSCP_SUBWORD_RMW(fetch_or, tag, T, static_cast<T>(a | x))                                  \
  ACPP_SSCP_SUBWORD_RMW(fetch_xor, tag, T, static_cast<T>(a ^ x))                                 \
  ACPP_SSCP_SUBWORD_RMW(exchange, tag, T, ((void)a, x))                                           \
  extern "C" __device__ __attribute__((always_inline)) inline void __acpp_sscp_atomic_store_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr, T x) {                                                                              \
    (void)acpp::detail::subword_rmw(ptr, [x](T) -> T { return x; }, acpp::detail::store_order(order), scope); \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline T __acpp_sscp_atomic_load_##tag(    \
      __acpp_sscp_address_space, __acpp_sscp_memory_order order, __acpp_sscp_memory_scope scope,  \
      T *ptr) {                                                                                   \
    return acpp::detail::subword_load(ptr, order, scope);                                         \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline bool __acpp_sscp_cmp_exch_strong_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order success, __acpp_sscp_memory_order,     \
      __acpp_sscp_memory_scope scope, T *ptr, T *expected, T desired) {                           \
    return acpp::detail::subword_cas(ptr, *expected, desired, success, scope);                    \
  }                                                                                               \
  extern "C" __device__ __attribute__((always_inline)) inline bool __acpp_sscp_cmp_exch_weak_##tag( \
      __acpp_sscp_address_space, __acpp_sscp_memory_order success, __acpp_sscp_memory_order,     \
      __acpp_sscp_memory_scope scope, T *ptr, T *expected, T desired) {                           \
    return acpp::detail::subword_cas(ptr, *expected, desired, success, scope);                    \
  }

ACPP_SSCP_SUBWORD_FULL(i8, int8_t)
ACPP_SSCP_SUBWORD_FULL(i16, int16_t)
ACPP_SSCP_SUBWORD_ARITH(u8, uint8_t)
ACPP_SSCP_SUBWORD_ARITH(u16, uint16_t)

#undef ACPP_SSCP_SUBWORD_FULL
#undef ACPP_SSCP_SUBWORD_ARITH
#undef ACPP_SSCP_SUBWORD_RMW
#undef ACPP_SSCP_ATOMIC_INT
#undef ACPP_SSCP_ATOMIC_COMMON
#undef ACPP_SSCP_ATOMIC_RMW
#undef ACPP_ATOMIC_ORDER_SCOPE_SWITCH
#undef ACPP_ATOMIC_SCOPE_SWITCH
