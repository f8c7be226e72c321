// Bulk atomic workloads: histogram by fetch_add(1) and strided fetch_min /
// fetch_max / fetch_add into a slot array.
//
// Reference behaviour being replaced: a user kernel in which work-item i
// applies sycl::atomic_ref<T, relaxed, device>(target[slot(i)]).fetch_<op>(v_i)
// (include/hipSYCL/sycl/libkernel/atomic_ref.hpp:143-197), which the HIP
// backend lowers to one global atomic per element
// (generic/hiplike/atomic_builtins.hpp:438-598).  Under skewed keys (Zipf)
// the hot bins serialise at the L2 / memory-side atomic units.
//
// MI355X design:
//   - strategy "privatized": every work-group keeps private counters in LDS
//     and only the group totals reach global memory.  Up to 32768 bins the
//     counters are 32-bit; up to 65536 bins two 16-bit counters share one
//     32-bit LDS word (65536 bins = 128 KiB of the 160 KiB LDS).  A 16-bit
//     counter that reaches 0x8000 is drained by the one work-item that saw
//     it cross: it subtracts 0x8000 in LDS and adds 0x8000 to the global
//     bin, so a field never carries into its neighbour.  Group totals are
//     written as per-group partial histograms (coalesced, no atomics) and a
//     second kernel adds their column sums to the bins, which keeps global
//     atomics off the hot path entirely.
//   - keys are streamed with 128-bit loads, several per work-item in flight;
//   - strided fetch_min / fetch_max / fetch_add (slot = i mod num_slots):
//     the grid is sized so that num_slots divides the grid stride whenever
//     that fits; then each work-item always meets the same slots, keeps them
//     in registers, and the group combines them once at the end.  Otherwise
//     the elements go through LDS atomics, with 16-byte loads and slots
//     tracked incrementally;
//   - strategy "global": one device-scope atomic per element, the
//     reference's pattern, kept for comparison.
#include <numeric>

#include "launch_util.hpp"

namespace acpp_impl {
namespace atomics_detail {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

enum strategy : int { strat_auto = 0, strat_privatized = 1, strat_global = 2 };

constexpr uint32_t lds_bytes_budget = 128 * 1024;

// 16-byte streaming geometry of an array: `head` scalar elements before the
// first 16-byte boundary, `nvec` aligned 16-byte chunks, then the tail.
struct geom {
  uint64_t head;
  uint64_t nvec;
  uint64_t tail_start;
  uint64_t nscalar;  // head + tail
};

template <class T>
geom make_geom(const T *p, uint64_t n) {
  constexpr uint64_t VEC = 16 / sizeof(T);
  geom g;
  g.head = ((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15) / sizeof(T);
  if (g.head > n) g.head = n;
  g.nvec = (n - g.head) / VEC;
  g.tail_start = g.head + g.nvec * VEC;
  g.nscalar = g.head + (n - g.tail_start);
  return g;
}

__device__ __forceinline__ uint64_t scalar_index(const geom &g, uint64_t s) {
  return (s < g.head) ? s : g.tail_start + (s - g.head);
}

// --------------------------------------------------------------------------
// Histogram
// --------------------------------------------------------------------------
// copies_log2 (32-bit counters only): the group keeps 2^copies_log2 copies
// of the histogram, word (bin << copies_log2) + copy, and work-item t counts
// into copy t mod 2^copies_log2.  Lanes that meet the same hot bin then hit
// adjacent words in different banks instead of one address.
template <bool PACKED, bool GLOBAL, int UNROLL>
__global__ __launch_bounds__(1024) void histogram_kernel(const uint32_t *__restrict__ keys,
                                                         geom gm, uint32_t num_bins,
                                                         uint32_t *partials, uint32_t *bins,
                                                         uint32_t copies_log2) {
  extern __shared__ uint32_t lds[];
  const uint32_t my_copy = threadIdx.x & ((1u << copies_log2) - 1u);
  // Packed counters: bin b < half in the low 16 bits of word b, bin
  // half + b in the high 16 bits, so the low (under skewed keys, hot) bins
  // sit in consecutive words and banks; adjacent bins sharing a word put
  // the two hottest Zipf bins on one address.
  const uint32_t half = (num_bins + 1) / 2;
  if constexpr (!GLOBAL) {
    const uint32_t words = PACKED ? half : (num_bins << copies_log2);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) lds[i] = 0;
    __syncthreads();
  }

  auto count = [&](uint32_t k) {
    if (k >= num_bins) return;
    if constexpr (GLOBAL) {
      __hip_atomic_fetch_add(&bins[k], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (PACKED) {
      const bool hi = k >= half;
      const uint32_t word = hi ? k - half : k;
      const uint32_t shift = hi ? 16u : 0u;
      const uint32_t old = atomicAdd(&lds[word], 1u << shift);
      if (((old >> shift) & 0xffffu) == 0x7fffu) {
        // This work-item moved the field to 0x8000: drain it.
        atomicAdd(&lds[word], 0u - (0x8000u << shift));
        __hip_atomic_fetch_add(&bins[k], 0x8000u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      atomicAdd(&lds[(k << copies_log2) + my_copy], 1u);
    }
  };

  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const u32x4 *kv = reinterpret_cast<const u32x4 *>(keys + gm.head);
  uint64_t c = gtid;
  for (; c + (UNROLL - 1) * stride < gm.nvec; c += UNROLL * stride) {
    u32x4 q[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) q[u] = __builtin_nontemporal_load(kv + c + u * stride);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      count(q[u].x);
      count(q[u].y);
      count(q[u].z);
      count(q[u].w);
    }
  }
  for (; c < gm.nvec; c += stride) {
    const u32x4 q = kv[c];
    count(q.x);
    count(q.y);
    count(q.z);
    count(q.w);
  }
  for (uint64_t s = gtid; s < gm.nscalar; s += stride) count(keys[scalar_index(gm, s)]);

  if constexpr (!GLOBAL) {
    __syncthreads();
    // Partial histogram of this group (coalesced, no atomics).
    uint32_t *mine = partials + static_cast<uint64_t>(blockIdx.x) * num_bins;
    for (uint32_t b = threadIdx.x; b < num_bins; b += blockDim.x) {
      uint32_t v;
      if constexpr (PACKED) {
        v = (b >= half) ? (lds[b - half] >> 16) : (lds[b] & 0xffffu);
      } else {
        v = 0;
        for (uint32_t c = 0; c < (1u << copies_log2); ++c) v += lds[(b << copies_log2) + c];
      }
      mine[b] = v;
    }
  }
}

// bins[b] += sum over groups of partials[g][b].  blockIdx.y takes every
// gridDim.y-th group, so the partials are read by many work-groups at once;
// the per-chunk sums are added with device-scope integer atomics (exact, so
// the order does not matter).
__global__ void histogram_combine_kernel(const uint32_t *__restrict__ partials, uint32_t groups,
                                         uint32_t num_bins, uint32_t *bins) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= num_bins) return;
  uint32_t s = 0;
  for (uint32_t g = blockIdx.y; g < groups; g += gridDim.y) s += partials[static_cast<uint64_t>(g) * num_bins + b];
  if (s) __hip_atomic_fetch_add(&bins[b], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------------------
// Strided fetch ops: target[i % num_slots] = op(target[i % num_slots], v[i])
// --------------------------------------------------------------------------
template <int Scope, class T, class Cmp>
__device__ __forceinline__ void cas_select(T *p, T v, Cmp keep_old) {
  T old = __hip_atomic_load(p, __ATOMIC_RELAXED, Scope);
  while (!keep_old(old, v)) {
    if (__hip_atomic_compare_exchange_strong(p, &old, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, Scope))
      break;
  }
}

template <class T>
struct rmw_min {
  __device__ static T identity() { return ops::minimum<T, T>::identity(); }
  __device__ static T apply(T a, T b) { return ops::minimum<T, T>::apply(a, b); }
  template <int Scope>
  __device__ static void atomic(T *p, T v) {
    if constexpr (std::is_floating_point_v<T>)
      cas_select<Scope>(p, v, [](T old, T x) { return old < x || old == x; });  // sycl::minimum rule
    else
      __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, Scope);
  }
};

template <class T>
struct rmw_max {
  __device__ static T identity() { return ops::maximum<T, T>::identity(); }
  __device__ static T apply(T a, T b) { return ops::maximum<T, T>::apply(a, b); }
  template <int Scope>
  __device__ static void atomic(T *p, T v) {
    if constexpr (std::is_floating_point_v<T>)
      cas_select<Scope>(p, v, [](T old, T x) { return old > x || old == x; });
    else
      __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, Scope);
  }
};

template <class T>
struct rmw_add {
  __device__ static T identity() { return T(0); }
  __device__ static T apply(T a, T b) { return ops::wrap_add(a, b); }
  template <int Scope>
  __device__ static void atomic(T *p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, Scope);
  }
};

constexpr int scope_wg = __HIP_MEMORY_SCOPE_WORKGROUP;
constexpr int scope_dev = __HIP_MEMORY_SCOPE_AGENT;

// Register-privatised form: requires num_slots | (VEC * total work-items),
// which the launcher arranges through the grid size.  Then element
// head + VEC*c + k of chunk c = gtid + m*stride lands in slot
// (head + VEC*gtid + k) mod num_slots for every m, so each work-item keeps
// VEC fixed slots in registers over the whole grid-stride loop.
template <class T, class R>
__global__ __launch_bounds__(1024) void strided_register_kernel(const T *__restrict__ v, geom gm,
                                                                uint32_t num_slots, T *target) {
  constexpr int VEC = 16 / sizeof(T);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const u32x4 *vv = reinterpret_cast<const u32x4 *>(v + gm.head);
  T acc[VEC];
#pragma unroll
  for (int k = 0; k < VEC; ++k) acc[k] = R::identity();
  uint64_t c = gtid;
  for (; c + 3 * stride < gm.nvec; c += 4 * stride) {
    u32x4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = __builtin_nontemporal_load(vv + c + u * stride);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      T e[VEC];
      __builtin_memcpy(e, &q[u], 16);
#pragma unroll
      for (int k = 0; k < VEC; ++k) acc[k] = R::apply(acc[k], e[k]);
    }
  }
  for (; c < gm.nvec; c += stride) {
    const u32x4 q = vv[c];
    T e[VEC];
    __builtin_memcpy(e, &q, 16);
#pragma unroll
    for (int k = 0; k < VEC; ++k) acc[k] = R::apply(acc[k], e[k]);
  }
  extern __shared__ unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  for (uint32_t i = threadIdx.x; i < num_slots; i += blockDim.x) lds[i] = R::identity();
  __syncthreads();
#pragma unroll
  for (int k = 0; k < VEC; ++k)
    R::template atomic<scope_wg>(&lds[(gm.head + VEC * gtid + k) % num_slots], acc[k]);
  for (uint64_t s = gtid; s < gm.nscalar; s += stride) {
    const uint64_t i = scalar_index(gm, s);
    R::template atomic<scope_wg>(&lds[i % num_slots], v[i]);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < num_slots; i += blockDim.x)
    R::template atomic<scope_dev>(&target[i], lds[i]);
}

// LDS-privatised form for any slot count that fits the LDS budget.  Keys
// are streamed with 16-byte loads (4 in flight) and the slot of each
// element is tracked incrementally: chunk c starts at slot
// (head + VEC c) mod num_slots, and moving c on by the grid stride adds a
// fixed step.  (One scalar load and one 64-bit modulo per element ran at a
// sixth of the register form's rate.)
template <class T, class R>
__global__ __launch_bounds__(1024) void strided_lds_kernel(const T *__restrict__ v, geom gm,
                                                           uint32_t num_slots, T *target) {
  constexpr int VEC = 16 / sizeof(T);
  constexpr int UNROLL = 4;
  extern __shared__ unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  for (uint32_t i = threadIdx.x; i < num_slots; i += blockDim.x) lds[i] = R::identity();
  __syncthreads();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const u32x4 *vv = reinterpret_cast<const u32x4 *>(v + gm.head);
  // slot of the first element of chunk c, and its advance per stride
  const uint32_t step = static_cast<uint32_t>((VEC * stride) % num_slots);
  uint32_t s0 = static_cast<uint32_t>((gm.head + VEC * gtid) % num_slots);
  auto advance = [&](uint32_t s, uint32_t by) {
    s += by;  // s, by < num_slots <= 2^31
    return s >= num_slots ? s - num_slots : s;
  };
  auto fold = [&](u32x4 q, uint32_t s) {
    T e[VEC];
    __builtin_memcpy(e, &q, 16);
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
      R::template atomic<scope_wg>(&lds[s], e[k]);
      s = advance(s, 1);
    }
  };
  uint64_t c = gtid;
  for (; c + (UNROLL - 1) * stride < gm.nvec; c += UNROLL * stride) {
    u32x4 q[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) q[u] = __builtin_nontemporal_load(vv + c + u * stride);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      fold(q[u], s0);
      s0 = advance(s0, step);
    }
  }
  for (; c < gm.nvec; c += stride) {
    fold(vv[c], s0);
    s0 = advance(s0, step);
  }
  for (uint64_t s = gtid; s < gm.nscalar; s += stride) {
    const uint64_t i = scalar_index(gm, s);
    R::template atomic<scope_wg>(&lds[i % num_slots], v[i]);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < num_slots; i += blockDim.x)
    R::template atomic<scope_dev>(&target[i], lds[i]);
}

// The reference's pattern: one device-scope atomic per element.
template <class T, class R>
__global__ __launch_bounds__(1024) void strided_global_kernel(const T *__restrict__ v, uint64_t n,
                                                              uint32_t num_slots, T *target) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    R::template atomic<scope_dev>(&target[i % num_slots], v[i]);
}

uint32_t pick_grid(uint64_t work_items, uint32_t wg, uint64_t max_groups) {
  uint64_t g = ceil_div(work_items, wg);
  if (g < 1) g = 1;
  if (g > max_groups) g = max_groups;
  return static_cast<uint32_t>(g);
}

acpp_status histogram(acpp_allocation_group *g, hipStream_t stream, const uint32_t *keys,
                      uint64_t n, uint32_t *bins, uint32_t num_bins, int strat,
                      const acpp_launch_config *cfg) {
  const char *name = "acpp_histogram_u32";
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_histogram_u32: scratch group is NULL");
  if (!bins || num_bins == 0)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_histogram_u32: bins is NULL or num_bins is 0");
  if (n > 0 && !keys) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_histogram_u32: keys is NULL");
  if (strat < strat_auto || strat > strat_global)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_histogram_u32: unknown strategy");
  if (n == 0) return clear_error();
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  const bool fits32 = static_cast<uint64_t>(num_bins) * 4 <= lds_bytes_budget;
  const bool fits16 = static_cast<uint64_t>((num_bins + 1) / 2) * 4 <= lds_bytes_budget;
  if (strat == strat_auto) strat = fits16 ? strat_privatized : strat_global;
  if (strat == strat_privatized && !fits16)
    return set_error(ACPP_ERROR_UNSUPPORTED,
                     "acpp_histogram_u32: more than 65536 bins do not fit the LDS; use the global strategy");
  const uint32_t wg = (cfg && cfg->work_group_size) ? cfg->work_group_size : 1024;
  if (!valid_work_group_size(wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_histogram_u32: bad work_group_size");
  const geom gm = make_geom(keys, n);

  if (strat == strat_global) {
    auto kernel = histogram_kernel<false, true, 4>;
    int resident = 1;
    st = occupancy(kernel, static_cast<int>(wg), &resident);
    if (st != ACPP_SUCCESS) return st;
    uint64_t groups_cap = static_cast<uint64_t>(resident) * cus;
    if (cfg && cfg->groups_per_cu)
      groups_cap = std::min<uint64_t>(groups_cap, cfg->groups_per_cu * static_cast<uint64_t>(cus));
    const uint32_t grid = pick_grid(gm.nvec + gm.nscalar, wg * 4, groups_cap);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(wg), 0, stream, keys, gm, num_bins, nullptr, bins, 0u);
    st = check_launch(name);
    return st == ACPP_SUCCESS ? clear_error() : st;
  }

  const bool packed = !fits32;
  // Up to 8 copies of 32-bit counters within the LDS budget.
  uint32_t copies_log2 = 0;
  while (!packed && copies_log2 < 3 &&
         (static_cast<uint64_t>(num_bins) << (copies_log2 + 1)) * 4 <= lds_bytes_budget)
    ++copies_log2;
  const size_t lds = packed ? static_cast<size_t>((num_bins + 1) / 2) * 4
                            : (static_cast<size_t>(num_bins) << copies_log2) * 4;
  auto kernel = packed ? histogram_kernel<true, false, 4> : histogram_kernel<false, false, 4>;
  int resident = 1;
  ACPP_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, static_cast<int>(wg), lds));
  if (resident < 1) resident = 1;
  uint64_t groups_cap = static_cast<uint64_t>(resident) * cus;
  // Each group writes a partial histogram; keep the partials to at most 1/8
  // of the key bytes (but at least one group per CU).
  const uint64_t partial_cap = std::max<uint64_t>(static_cast<uint64_t>(cus), (n / 8) / num_bins);
  groups_cap = std::min(groups_cap, partial_cap);
  // One group per CU by default: the kernel is bound by the CU's LDS atomic
  // rate (about two keys per clock), and a second resident group only adds
  // contention (2^28 Zipf keys, 4096 bins on MI355X: 252 us with one group
  // per CU, 302 us with two).
  const uint32_t gpc = (cfg && cfg->groups_per_cu) ? cfg->groups_per_cu : 1u;
  groups_cap = std::min<uint64_t>(groups_cap, gpc * static_cast<uint64_t>(cus));
  const uint32_t grid = pick_grid(gm.nvec + gm.nscalar, wg * 4, groups_cap);
  void *partials = nullptr;
  st = group_obtain(g, static_cast<size_t>(grid) * num_bins * sizeof(uint32_t), &partials);
  if (st != ACPP_SUCCESS) return st;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(wg), lds, stream, keys, gm, num_bins,
                     static_cast<uint32_t *>(partials), bins, copies_log2);
  st = check_launch(name);
  if (st != ACPP_SUCCESS) return st;
  const uint32_t chunks = std::min<uint32_t>(grid, 32u);
  hipLaunchKernelGGL(histogram_combine_kernel, dim3(static_cast<uint32_t>(ceil_div(num_bins, 256)), chunks),
                     dim3(256), 0, stream, static_cast<const uint32_t *>(partials), grid, num_bins, bins);
  st = check_launch(name);
  return st == ACPP_SUCCESS ? clear_error() : st;
}

template <class T, template <class> class RT>
acpp_status strided(acpp_allocation_group *g, hipStream_t stream, const T *values, uint64_t n,
                    T *target, uint32_t num_slots, int strat, const acpp_launch_config *cfg,
                    const char *name) {
  using R = RT<T>;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!target || num_slots == 0)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": target is NULL or num_slots is 0");
  if (n > 0 && !values) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": values is NULL");
  if (strat < strat_auto || strat > strat_global)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": unknown strategy");
  if (n == 0) return clear_error();
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  constexpr uint32_t VEC = 16 / sizeof(T);
  const uint32_t wg = (cfg && cfg->work_group_size) ? cfg->work_group_size : 256;
  if (!valid_work_group_size(wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": bad work_group_size");
  const bool fits = static_cast<uint64_t>(num_slots) * sizeof(T) <= lds_bytes_budget;
  if (strat == strat_auto) strat = fits ? strat_privatized : strat_global;
  if (strat == strat_privatized && !fits)
    return set_error(ACPP_ERROR_UNSUPPORTED, std::string(name) + ": slots do not fit the LDS; use the global strategy");
  const uint64_t cap_groups =
      static_cast<uint64_t>(cus) * ((cfg && cfg->groups_per_cu) ? cfg->groups_per_cu : 2);
  if (strat == strat_global) {
    const uint32_t grid = pick_grid(n, wg, cap_groups * 8);
    hipLaunchKernelGGL((strided_global_kernel<T, R>), dim3(grid), dim3(wg), 0, stream, values, n, num_slots, target);
  } else {
    const size_t lds = static_cast<size_t>(num_slots) * sizeof(T);
    const geom gm = make_geom(values, n);
    const uint32_t grid = pick_grid(gm.nvec + gm.nscalar, wg * 4, cap_groups);
    // The register form needs VEC * wg * grid to be a multiple of
    // num_slots: grid a multiple of `period`.  (1000 slots: period 125 with
    // 256-wide groups; with per-element LDS atomics instead, fp32 adds ran
    // at 0.8 TB/s.)
    const uint64_t period = num_slots / std::gcd<uint64_t>(static_cast<uint64_t>(VEC) * wg, num_slots);
    if (period <= grid) {
      const uint32_t grid_reg = static_cast<uint32_t>((grid / period) * period);
      hipLaunchKernelGGL((strided_register_kernel<T, R>), dim3(grid_reg), dim3(wg), lds, stream, values, gm,
                         num_slots, target);
    } else {
      hipLaunchKernelGGL((strided_lds_kernel<T, R>), dim3(grid), dim3(wg), lds, stream, values, gm,
                         num_slots, target);
    }
  }
  st = check_launch(name);
  return st == ACPP_SUCCESS ? clear_error() : st;
}

}  // namespace atomics_detail
}  // namespace acpp_impl

using namespace acpp_impl;
using namespace acpp_impl::atomics_detail;

extern "C" {

acpp_status acpp_histogram_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *keys,
                               uint64_t n, uint32_t *bins, uint32_t num_bins, int strategy,
                               const acpp_launch_config *config) {
  return histogram(scratch, static_cast<hipStream_t>(stream), keys, n, bins, num_bins, strategy, config);
}

#define ACPP_DEFINE_STRIDED(opname, RT, tag, T)                                                           \
  acpp_status acpp_atomic_fetch_##opname##_strided_##tag(acpp_allocation_group *scratch, acpp_stream_t stream, \
                                                         const T *values, uint64_t n, T *target,          \
                                                         uint32_t num_slots, int strategy,                \
                                                         const acpp_launch_config *config) {              \
    return strided<T, RT>(scratch, static_cast<hipStream_t>(stream), values, n, target, num_slots,        \
                          strategy, config, "acpp_atomic_fetch_" #opname "_strided_" #tag);              \
  }
#define ACPP_DEFINE_STRIDED_ALL(tag, T)         \
  ACPP_DEFINE_STRIDED(add, rmw_add, tag, T)     \
  ACPP_DEFINE_STRIDED(min, rmw_min, tag, T)     \
  ACPP_DEFINE_STRIDED(max, rmw_max, tag, T)

ACPP_DEFINE_STRIDED_ALL(i32, int32_t)
ACPP_DEFINE_STRIDED_ALL(u32, uint32_t)
ACPP_DEFINE_STRIDED_ALL(i64, int64_t)
ACPP_DEFINE_STRIDED_ALL(u64, uint64_t)
ACPP_DEFINE_STRIDED_ALL(f32, float)
ACPP_DEFINE_STRIDED_ALL(f64, double)

}  // extern "C"
