// Work-group and sub-group collectives for HIP kernels on gfx950.
//
// Replaces, for user kernels, the reference's group algorithm library:
//   include/hipSYCL/sycl/libkernel/group_functions.hpp:142-239, 305-311
//     reduce_over_group, inclusive/exclusive_scan_over_group, joint_reduce,
//     joint_inclusive/exclusive_scan;
//   include/hipSYCL/sycl/libkernel/generic/hiplike/group_functions.hpp:35-71,
//     214-351, 455-525 and hip/group_functions.hpp:75-133 (the HIP
//     implementation: warp shuffles + per-warp partials in LDS);
//   include/hipSYCL/sycl/libkernel/sscp/builtins/{reduction,scan_inclusive,
//     scan_exclusive}.hpp (the extern "C" builtin ABI, provided at the end).
//
// Design: a sub-group is one wave64.  Inside a wave the scan runs on DPP
// (row_shr 1/2/4/8, row_bcast 15/31) with lane predicates instead of an
// identity element, so any associative operator works -- including
// user functors without a known identity -- and the first element of an
// inclusive scan is passed through untouched, as in a sequential scan.
// Across waves one LDS exchange of the wave totals (at most 16) finishes the
// job.  Partially filled last waves are supported: their active lanes are a
// prefix of the wave, and each lane only reads lower lanes.
//
// All work-items of the group must call a work-group collective (as in
// SYCL).  Each collective may be called repeatedly in a kernel; it begins
// with a barrier that protects its LDS slots from the previous call.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <type_traits>

#include "acpp/device/ops.hpp"
#include "acpp/functional.hpp"
#include "acpp/device/wave.hpp"

namespace acpp {

// Operators: acpp::plus ... acpp::logical_or (acpp/functional.hpp).

// Known identities (sycl::known_identity, functional.hpp:123-244).
template <class Op, class T>
struct known_identity {
  static constexpr bool value = false;
};
#define ACPP_KNOWN_IDENTITY(OP, OPSFN)                                            \
  template <class U, class T>                                                     \
  struct known_identity<OP<U>, T> {                                               \
    static constexpr bool value = std::is_integral_v<T> || ops::is_floating_v<T>;  \
    __host__ __device__ static T get() { return ops::OPSFN<T, T>::identity(); }   \
  };
ACPP_KNOWN_IDENTITY(plus, plus)
ACPP_KNOWN_IDENTITY(multiplies, multiply)
ACPP_KNOWN_IDENTITY(minimum, minimum)
ACPP_KNOWN_IDENTITY(maximum, maximum)
ACPP_KNOWN_IDENTITY(bit_and, bit_and)
ACPP_KNOWN_IDENTITY(bit_or, bit_or)
ACPP_KNOWN_IDENTITY(bit_xor, bit_xor)
ACPP_KNOWN_IDENTITY(logical_and, logical_and)
ACPP_KNOWN_IDENTITY(logical_or, logical_or)
ACPP_KNOWN_IDENTITY(std::plus, plus)
ACPP_KNOWN_IDENTITY(std::multiplies, multiply)
ACPP_KNOWN_IDENTITY(std::bit_and, bit_and)
ACPP_KNOWN_IDENTITY(std::bit_or, bit_or)
ACPP_KNOWN_IDENTITY(std::bit_xor, bit_xor)
#undef ACPP_KNOWN_IDENTITY

// ---------------------------------------------------------------------------
// Groups
// ---------------------------------------------------------------------------
// The work-group of the calling work-item (the HIP block).
struct group {
  __device__ uint32_t get_local_linear_id() const {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  }
  __device__ uint32_t get_local_linear_range() const {
    return blockDim.x * blockDim.y * blockDim.z;
  }
  __device__ uint32_t get_group_linear_id() const {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  }
  __device__ bool leader() const { return get_local_linear_id() == 0; }
};

// The sub-group (wave64) of the calling work-item.
struct sub_group {
  __device__ uint32_t get_local_linear_id() const { return group{}.get_local_linear_id() & 63u; }
  __device__ uint32_t get_group_linear_id() const { return group{}.get_local_linear_id() >> 6; }
  // Work-items in this sub-group (64 except in a partial last wave).
  __device__ uint32_t get_local_linear_range() const {
    const uint32_t base = group{}.get_local_linear_id() & ~63u;
    const uint32_t r = group{}.get_local_linear_range() - base;
    return r < 64u ? r : 64u;
  }
  __device__ static constexpr uint32_t get_max_local_range() { return 64u; }
  __device__ bool leader() const { return get_local_linear_id() == 0; }
};

__device__ inline group this_group() { return group{}; }
__device__ inline sub_group this_sub_group() { return sub_group{}; }

namespace detail {

// Move T through DPP as 32-bit pieces; a lane whose source is invalid keeps
// its own value (the caller masks it out).
template <int Ctrl, int RowMask, class T>
__device__ __forceinline__ T dpp_keep(const T &x) {
  static_assert(std::is_trivially_copyable_v<T>, "group collectives need trivially copyable T");
  constexpr int words = (sizeof(T) + 3) / 4;
  uint32_t w[words] = {};
  __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
  for (int i = 0; i < words; ++i)
    w[i] = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(w[i]), static_cast<int>(w[i]),
                                                             Ctrl, RowMask, 0xf, false));
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

template <class T>
__device__ __forceinline__ T read_lane_any(const T &x, int l) {
  constexpr int words = (sizeof(T) + 3) / 4;
  uint32_t w[words] = {};
  __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
  for (int i = 0; i < words; ++i)
    w[i] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(w[i]), l));
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Inclusive scan over the active lanes of the wave, element order = lane
// order; op(earlier, later).  No identity is needed.
template <class T, class F>
__device__ __forceinline__ T wave_inclusive(T x, F op) {
  const int lane = wave::lane_id();
  const int r = lane & 15;
  T t;
  t = dpp_keep<wave::row_shr1, 0xf>(x);
  x = (r >= 1) ? op(t, x) : x;
  t = dpp_keep<wave::row_shr2, 0xf>(x);
  x = (r >= 2) ? op(t, x) : x;
  t = dpp_keep<wave::row_shr4, 0xf>(x);
  x = (r >= 4) ? op(t, x) : x;
  t = dpp_keep<wave::row_shr8, 0xf>(x);
  x = (r >= 8) ? op(t, x) : x;
  t = dpp_keep<wave::row_bcast15, 0xa>(x);
  x = (lane & 16) ? op(t, x) : x;
  t = dpp_keep<wave::row_bcast31, 0xc>(x);
  x = (lane >= 32) ? op(t, x) : x;
  return x;
}

// Scan value of the previous lane (meaningless in lane 0).
template <class T>
__device__ __forceinline__ T wave_previous(const T &incl) {
  return dpp_keep<wave::wave_shr1, 0xf>(incl);
}

template <class T>
struct alignas(alignof(T) < 4 ? 4 : alignof(T)) raw_slot {
  unsigned char bytes[sizeof(T)];
};

template <class T>
__device__ __forceinline__ T load_slot(const raw_slot<T> &s) {
  T v;
  __builtin_memcpy(&v, s.bytes, sizeof(T));
  return v;
}
template <class T>
__device__ __forceinline__ void store_slot(raw_slot<T> &s, const T &v) {
  __builtin_memcpy(s.bytes, &v, sizeof(T));
}

constexpr int max_waves = 16;

// Publish each wave's total, return the op-combination of the totals of the
// waves before the calling wave in `before` (valid when wave > 0) and of all
// waves in `all`.
template <class T, class F>
__device__ __forceinline__ void combine_waves(const T &wave_total, F op, T &before, T &all) {
  __shared__ raw_slot<T> slots[max_waves];
  const uint32_t lid = group{}.get_local_linear_id();
  const uint32_t w = lid >> 6;
  const uint32_t nw = (group{}.get_local_linear_range() + 63u) >> 6;
  __syncthreads();  // slots may still be read by a previous collective
  if ((lid & 63u) == 0) store_slot(slots[w], wave_total);
  __syncthreads();
  T acc = load_slot(slots[0]);
  if (w == 1) before = acc;
  for (uint32_t i = 1; i < nw; ++i) {
    acc = op(acc, load_slot(slots[i]));
    if (i + 1 == w) before = acc;
  }
  all = acc;
}

}  // namespace detail

// ---------------------------------------------------------------------------
// Sub-group (wave) collectives
// ---------------------------------------------------------------------------
template <class T, class F>
__device__ T reduce_over_group(sub_group, T x, F op) {
  const T incl = detail::wave_inclusive(x, op);
  return detail::read_lane_any(incl, wave::last_active_lane());
}

template <class T, class F>
__device__ T reduce_over_group(sub_group g, T x, T init, F op) {
  return op(reduce_over_group(g, x, op), init);
}

template <class T, class F>
__device__ T inclusive_scan_over_group(sub_group, T x, F op) {
  return detail::wave_inclusive(x, op);
}

template <class T, class F>
__device__ T inclusive_scan_over_group(sub_group, T x, T init, F op) {
  return op(init, detail::wave_inclusive(x, op));
}

template <class T, class F>
__device__ T exclusive_scan_over_group(sub_group, T x, T init, F op) {
  const T prev = detail::wave_previous(detail::wave_inclusive(x, op));
  return (wave::lane_id() == 0) ? init : op(init, prev);
}

template <class T, class F>
__device__ T exclusive_scan_over_group(sub_group g, T x, F op) {
  static_assert(known_identity<F, T>::value, "exclusive scan without init needs a known identity");
  return exclusive_scan_over_group(g, x, known_identity<F, T>::get(), op);
}

// ---------------------------------------------------------------------------
// Work-group collectives
// ---------------------------------------------------------------------------
template <class T, class F>
__device__ T reduce_over_group(group, T x, F op) {
  const T incl = detail::wave_inclusive(x, op);
  const T total = detail::read_lane_any(incl, wave::last_active_lane());
  T before = total, all = total;
  detail::combine_waves(total, op, before, all);
  return all;
}

// reduce_over_group(g, x, init, op) == op(reduce_over_group(g, x, op), init)
// (group_functions.hpp:305-311).
template <class T, class F>
__device__ T reduce_over_group(group g, T x, T init, F op) {
  return op(reduce_over_group(g, x, op), init);
}

template <class T, class F>
__device__ T inclusive_scan_over_group(group, T x, F op) {
  T incl = detail::wave_inclusive(x, op);
  const T total = detail::read_lane_any(incl, wave::last_active_lane());
  T before = total, all = total;
  detail::combine_waves(total, op, before, all);
  if ((group{}.get_local_linear_id() >> 6) > 0) incl = op(before, incl);
  return incl;
}

template <class T, class F>
__device__ T inclusive_scan_over_group(group g, T x, T init, F op) {
  return op(init, inclusive_scan_over_group(g, x, op));
}

template <class T, class F>
__device__ T exclusive_scan_over_group(group, T x, T init, F op) {
  const T incl = detail::wave_inclusive(x, op);
  const T prev = detail::wave_previous(incl);
  const T total = detail::read_lane_any(incl, wave::last_active_lane());
  T before = total, all = total;
  detail::combine_waves(total, op, before, all);
  T r = init;
  if ((group{}.get_local_linear_id() >> 6) > 0) r = op(r, before);
  if (wave::lane_id() > 0) r = op(r, prev);
  return r;
}

template <class T, class F>
__device__ T exclusive_scan_over_group(group g, T x, F op) {
  static_assert(known_identity<F, T>::value, "exclusive scan without init needs a known identity");
  return exclusive_scan_over_group(g, x, known_identity<F, T>::get(), op);
}

// ---------------------------------------------------------------------------
// Joint algorithms: the group cooperates on [first, last)
// (group_functions.hpp:142-159; hiplike/group_functions.hpp:214-287, 361-432,
// 501-525).
// ---------------------------------------------------------------------------
template <class Ptr, class T, class F>
__device__ T joint_reduce(group g, Ptr first, Ptr last, T init, F op) {
  const uint32_t lid = g.get_local_linear_id();
  const uint32_t range = g.get_local_linear_range();
  const ptrdiff_t n = last - first;
  if (n <= 0) return init;
  // Every work-item folds a strided subsequence; work-items without an
  // element contribute their first element's neighbour-free value via
  // the predicated reduction below.
  bool has = static_cast<ptrdiff_t>(lid) < n;
  T acc = has ? T(first[lid]) : T(first[0]);
  for (auto i = static_cast<ptrdiff_t>(lid) + range; i < n; i += range) acc = op(acc, T(first[i]));
  // Reduce over the work-items that hold a value (a prefix of the group).
  const uint32_t holders = static_cast<uint32_t>(n < static_cast<ptrdiff_t>(range) ? n : range);
  T incl = detail::wave_inclusive(acc, op);
  const uint32_t w = lid >> 6;
  const uint32_t wave_holders =
      holders > w * 64 ? ((holders - w * 64) < 64 ? holders - w * 64 : 64) : 0;
  const T total = wave_holders ? detail::read_lane_any(incl, static_cast<int>(wave_holders - 1)) : acc;
  // Waves without holders take no part in the combination below.
  __shared__ detail::raw_slot<T> slots[detail::max_waves];
  __syncthreads();
  if ((lid & 63u) == 0 && wave_holders) detail::store_slot(slots[w], total);
  __syncthreads();
  T r = detail::load_slot(slots[0]);
  const uint32_t nhw = (holders + 63u) >> 6;
  for (uint32_t i = 1; i < nhw; ++i) r = op(r, detail::load_slot(slots[i]));
  return op(init, r);
}

template <class Ptr, class F>
__device__ auto joint_reduce(group g, Ptr first, Ptr last, F op) {
  using T = std::decay_t<decltype(*first)>;
  static_assert(known_identity<F, T>::value, "joint_reduce without init needs a known identity");
  return joint_reduce(g, first, last, known_identity<F, T>::get(), op);
}

template <class InPtr, class OutPtr, class T, class F>
__device__ OutPtr joint_inclusive_scan(group g, InPtr first, InPtr last, OutPtr result, F op,
                                       T init) {
  const uint32_t lid = g.get_local_linear_id();
  const uint32_t range = g.get_local_linear_range();
  const ptrdiff_t n = last - first;
  T carry = init;
  for (ptrdiff_t base = 0; base < n; base += range) {
    const auto i = base + static_cast<ptrdiff_t>(lid);
    // Out-of-range work-items repeat the last element; their results are
    // not written and they sit after every in-range work-item.
    T x = T(first[i < n ? i : n - 1]);
    T s = inclusive_scan_over_group(g, x, op);
    s = op(carry, s);
    if (i < n) result[i] = s;
    const uint32_t last_in = static_cast<uint32_t>((n - base) < static_cast<ptrdiff_t>(range) ? (n - base) : range) - 1;
    // Broadcast the chunk's last value as the next carry.
    __shared__ detail::raw_slot<T> c;
    __syncthreads();
    if (lid == last_in) detail::store_slot(c, s);
    __syncthreads();
    carry = detail::load_slot(c);
  }
  return result + n;
}

template <class InPtr, class OutPtr, class F>
__device__ OutPtr joint_inclusive_scan(group g, InPtr first, InPtr last, OutPtr result, F op) {
  using T = std::decay_t<decltype(*first)>;
  const uint32_t lid = g.get_local_linear_id();
  const uint32_t range = g.get_local_linear_range();
  const ptrdiff_t n = last - first;
  if (n <= 0) return result;
  // First chunk without a carry, later chunks with the previous total.
  T x0 = T(first[static_cast<ptrdiff_t>(lid) < n ? lid : n - 1]);
  T s0 = inclusive_scan_over_group(g, x0, op);
  if (static_cast<ptrdiff_t>(lid) < n) result[lid] = s0;
  if (n <= static_cast<ptrdiff_t>(range)) return result + n;
  __shared__ detail::raw_slot<T> c;
  __syncthreads();
  if (lid == range - 1) detail::store_slot(c, s0);
  __syncthreads();
  const T carry = detail::load_slot(c);
  joint_inclusive_scan(g, first + range, last, result + range, op, carry);
  return result + n;
}

template <class InPtr, class OutPtr, class T, class F>
__device__ OutPtr joint_exclusive_scan(group g, InPtr first, InPtr last, OutPtr result, T init,
                                       F op) {
  const uint32_t lid = g.get_local_linear_id();
  const uint32_t range = g.get_local_linear_range();
  const ptrdiff_t n = last - first;
  T carry = init;
  for (ptrdiff_t base = 0; base < n; base += range) {
    const auto i = base + static_cast<ptrdiff_t>(lid);
    T x = T(first[i < n ? i : n - 1]);
    T s = exclusive_scan_over_group(g, x, carry, op);
    if (i < n) result[i] = s;
    const uint32_t last_in = static_cast<uint32_t>((n - base) < static_cast<ptrdiff_t>(range) ? (n - base) : range) - 1;
    __shared__ detail::raw_slot<T> c;
    __syncthreads();
    if (lid == last_in) detail::store_slot(c, op(s, x));
    __syncthreads();
    carry = detail::load_slot(c);
  }
  return result + n;
}

template <class InPtr, class OutPtr, class F>
__device__ OutPtr joint_exclusive_scan(group g, InPtr first, InPtr last, OutPtr result, F op) {
  using T = std::decay_t<decltype(*first)>;
  static_assert(known_identity<F, T>::value, "exclusive scan without init needs a known identity");
  return joint_exclusive_scan(g, first, last, result, known_identity<F, T>::get(), op);
}

// ---------------------------------------------------------------------------
// Broadcast, barrier, any/all/none, shifts, permutes and selects
// (include/hipSYCL/sycl/libkernel/group_functions.hpp:57-140, 240-276;
// the HIP implementation is generic/hiplike/group_functions.hpp:75-212 and
// 546-640: shuffles for sub-groups, an LDS array of group size for groups).
//
// Sub-group forms move 32-bit pieces with ds_bpermute (any source lane) or
// readlane (a uniform source lane) and take the ballot for any/all/none.
// Work-group forms use __syncthreads_or / _and for the predicates and one
// LDS slot per work-item for the data movement.
//
// A source outside the group gives the caller's own value for a sub-group
// (as the HIP shuffles do) and work-item 0's value for a work-group (the
// reference reads slot 0; it also reads slot `size` when the target is
// exactly the group size, which is outside the written range -- here that
// case reads slot 0 as well).
// ---------------------------------------------------------------------------
namespace detail {

template <class T>
__device__ __forceinline__ T bpermute(const T &x, uint32_t src_lane) {
  constexpr int words = (sizeof(T) + 3) / 4;
  uint32_t w[words] = {};
  __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
  for (int i = 0; i < words; ++i)
    w[i] = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(static_cast<int>(src_lane << 2),
                                                              static_cast<int>(w[i])));
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Read the value of `src` in the sub-group, or keep x when src is outside.
template <class T>
__device__ __forceinline__ T sub_group_fetch(sub_group g, const T &x, uint64_t src) {
  const bool ok = src < g.get_local_linear_range();
  const T y = bpermute(x, static_cast<uint32_t>(ok ? src : g.get_local_linear_id()));
  return ok ? y : x;
}

// Work-group data movement through one LDS slot per work-item.
template <class T>
__device__ __forceinline__ T group_fetch(group g, const T &x, uint64_t src) {
  __shared__ raw_slot<T> slots[1024];
  const uint32_t lid = g.get_local_linear_id();
  __syncthreads();  // slots may still be read by a previous collective
  store_slot(slots[lid], x);
  __syncthreads();
  return load_slot(slots[src < g.get_local_linear_range() ? src : 0]);
}

}  // namespace detail

template <class T>
__device__ T group_broadcast(sub_group, T x, uint32_t local_linear_id = 0) {
  return detail::read_lane_any(x, static_cast<int>(local_linear_id));
}
template <class T>
__device__ T group_broadcast(group g, T x, uint32_t local_linear_id = 0) {
  __shared__ detail::raw_slot<T> slot;
  __syncthreads();
  if (g.get_local_linear_id() == local_linear_id) detail::store_slot(slot, x);
  __syncthreads();
  return detail::load_slot(slot);
}

__device__ inline void group_barrier(group) { __syncthreads(); }
__device__ inline void group_barrier(sub_group) { __builtin_amdgcn_wave_barrier(); }

__device__ inline bool any_of_group(sub_group, bool pred) { return wave::ballot(pred) != 0; }
__device__ inline bool all_of_group(sub_group, bool pred) {
  return wave::ballot(!pred) == 0;  // inactive lanes do not vote
}
__device__ inline bool none_of_group(sub_group g, bool pred) { return !any_of_group(g, pred); }
__device__ inline bool any_of_group(group, bool pred) { return __syncthreads_or(pred) != 0; }
__device__ inline bool all_of_group(group, bool pred) { return __syncthreads_and(pred) != 0; }
__device__ inline bool none_of_group(group g, bool pred) { return !any_of_group(g, pred); }

template <class G, class T, class P>
__device__ bool any_of_group(G g, T x, P pred) { return any_of_group(g, static_cast<bool>(pred(x))); }
template <class G, class T, class P>
__device__ bool all_of_group(G g, T x, P pred) { return all_of_group(g, static_cast<bool>(pred(x))); }
template <class G, class T, class P>
__device__ bool none_of_group(G g, T x, P pred) { return none_of_group(g, static_cast<bool>(pred(x))); }

// joint_*_of: the group checks [first, last) together, strided by the
// group size (HGF __acpp_joint_any_of and friends).
template <class G, class Ptr, class P>
__device__ bool joint_any_of(G g, Ptr first, Ptr last, P pred) {
  bool found = false;
  for (Ptr p = first + g.get_local_linear_id(); p < last; p += g.get_local_linear_range())
    found = found || static_cast<bool>(pred(*p));
  return any_of_group(g, found);
}
template <class G, class Ptr, class P>
__device__ bool joint_all_of(G g, Ptr first, Ptr last, P pred) {
  bool all = true;
  for (Ptr p = first + g.get_local_linear_id(); p < last; p += g.get_local_linear_range())
    all = all && static_cast<bool>(pred(*p));
  return all_of_group(g, all);
}
template <class G, class Ptr, class P>
__device__ bool joint_none_of(G g, Ptr first, Ptr last, P pred) {
  return !joint_any_of(g, first, last, pred);
}

template <class T>
__device__ T shift_group_left(sub_group g, T x, uint32_t delta = 1) {
  return detail::sub_group_fetch(g, x, static_cast<uint64_t>(g.get_local_linear_id()) + delta);
}
template <class T>
__device__ T shift_group_right(sub_group g, T x, uint32_t delta = 1) {
  const uint32_t l = g.get_local_linear_id();
  return detail::sub_group_fetch(g, x, l >= delta ? static_cast<uint64_t>(l - delta) : ~0ull);
}
template <class T>
__device__ T permute_group_by_xor(sub_group g, T x, uint32_t mask) {
  return detail::sub_group_fetch(g, x, static_cast<uint64_t>(g.get_local_linear_id() ^ mask));
}
template <class T>
__device__ T select_from_group(sub_group g, T x, uint32_t remote_local_id) {
  return detail::sub_group_fetch(g, x, remote_local_id);
}

template <class T>
__device__ T shift_group_left(group g, T x, uint32_t delta = 1) {
  return detail::group_fetch(g, x, static_cast<uint64_t>(g.get_local_linear_id()) + delta);
}
template <class T>
__device__ T shift_group_right(group g, T x, uint32_t delta = 1) {
  const uint32_t l = g.get_local_linear_id();
  return detail::group_fetch(g, x, l >= delta ? static_cast<uint64_t>(l - delta) : ~0ull);
}
template <class T>
__device__ T permute_group_by_xor(group g, T x, uint32_t mask) {
  return detail::group_fetch(g, x, static_cast<uint64_t>(g.get_local_linear_id() ^ mask));
}
template <class T>
__device__ T select_from_group(group g, T x, uint32_t remote_local_id) {
  return detail::group_fetch(g, x, remote_local_id);
}

}  // namespace acpp

// ---------------------------------------------------------------------------
// extern "C" builtin ABI of the reference's SSCP compilation flow
// (include/hipSYCL/sycl/libkernel/sscp/builtins/reduction.hpp:16-80,
// scan_inclusive.hpp:16-86, scan_exclusive.hpp:16-102).  The operator is
// the __acpp_sscp_algorithm_op value (builtin_config.hpp:27-37).
// Floating-point variants (f16 as _Float16, f32, f64) accept plus,
// multiply, min and max and return T{} for the other operators, like the reference's AMDGPU builtins
// (src/libkernel/sscp/amdgpu/reduction.cpp:15-31).
// ---------------------------------------------------------------------------
namespace acpp {
namespace detail {
template <class T, class G, class F>
__device__ __forceinline__ T dispatch_collective(int kind, G g, T x, T init, F op) {
  if (kind == 0) return reduce_over_group(g, x, op);
  if (kind == 1) return inclusive_scan_over_group(g, x, op);
  return exclusive_scan_over_group(g, x, init, op);
}

template <class T, class G>
__device__ __forceinline__ T sscp_collective(int kind, G g, int op, T x, T init) {
  switch (op) {
    case ops::op_plus: return dispatch_collective(kind, g, x, init, plus<T>{});
    case ops::op_multiply: return dispatch_collective(kind, g, x, init, multiplies<T>{});
    case ops::op_min: return dispatch_collective(kind, g, x, init, minimum<T>{});
    case ops::op_max: return dispatch_collective(kind, g, x, init, maximum<T>{});
    default: break;
  }
  if constexpr (std::is_integral_v<T>) {
    switch (op) {
      case ops::op_bit_and: return dispatch_collective(kind, g, x, init, bit_and<T>{});
      case ops::op_bit_or: return dispatch_collective(kind, g, x, init, bit_or<T>{});
      case ops::op_bit_xor: return dispatch_collective(kind, g, x, init, bit_xor<T>{});
      case ops::op_logical_and: return dispatch_collective(kind, g, x, init, logical_and<T>{});
      case ops::op_logical_or: return dispatch_collective(kind, g, x, init, logical_or<T>{});
      default: break;
    }
  }
  return T{};
}
}  // namespace detail
}  // namespace acpp

typedef int32_t __acpp_sscp_algorithm_op;

#define ACPP_SSCP_COLLECTIVES(tag, T)                                                                    \
  extern "C" __device__ __attribute__((always_inline, convergent)) inline T                              \
      __acpp_sscp_work_group_reduce_##tag(__acpp_sscp_algorithm_op op, T x) {                            \
    return acpp::detail::sscp_collective<T>(0, acpp::group{}, op, x, T{});                                \
  }                                                                                                      \
  extern "C" __device__ __attribute__((always_inline, convergent)) inline T                              \
      __acpp_sscp_work_group_inclusive_scan_##tag(__acpp_sscp_algorithm_op op, T x) {                    \
    return acpp::detail::sscp_collective<T>(1, acpp::group{}, op, x, T{});                                \
  }                                                                                                      \
  extern "C" __device__ __attribute__((always_inline, convergent)) inline T                              \
      __acpp_sscp_work_group_exclusive_scan_##tag(__acpp_sscp_algorithm_op op, T x, T init) {            \
    return acpp::detail::sscp_collective<T>(2, acpp::group{}, op, x, init);                               \
  }                                                                                                      \
  extern "C" __device__ __attribute__((always_inline, convergent)) inline T                              \
      __acpp_sscp_sub_group_reduce_##tag(__acpp_sscp_algorithm_op op, T x) {                             \
    return acpp::detail::sscp_collective<T>(0, acpp::sub_group{}, op, x, T{});                            \
  }                                                                                                      \
  extern "C" __device__ __attribute__((always_inline, convergent)) inline T                              \
      __acpp_sscp_sub_group_inclusive_scan_##tag(__acpp_sscp_algorithm_op op, T x) {                     \
    return acpp::detail::sscp_collective<T>(1, acpp::sub_group{}, op, x, T{});                            \
  }                                                                                                      \
  extern "C" __device__ __attribute__((always_inline, convergent)) inline T                              \
      __acpp_sscp_sub_group_exclusive_scan_##tag(__acpp_sscp_algorithm_op op, T x, T init) {             \
    return acpp::detail::sscp_collective<T>(2, acpp::sub_group{}, op, x, init);                           \
  }

ACPP_SSCP_COLLECTIVES(i8, int8_t)
ACPP_SSCP_COLLECTIVES(i16, int16_t)
ACPP_SSCP_COLLECTIVES(i32, int32_t)
ACPP_SSCP_COLLECTIVES(i64, int64_t)
ACPP_SSCP_COLLECTIVES(u8, uint8_t)
ACPP_SSCP_COLLECTIVES(u16, uint16_t)
ACPP_SSCP_COLLECTIVES(u32, uint32_t)
ACPP_SSCP_COLLECTIVES(u64, uint64_t)
ACPP_SSCP_COLLECTIVES(f16, _Float16)
ACPP_SSCP_COLLECTIVES(f32, float)
ACPP_SSCP_COLLECTIVES(f64, double)
#undef ACPP_SSCP_COLLECTIVES
