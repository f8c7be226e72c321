// Device-wide reduction kernel shared by reduce, transform_reduce (dot),
// sycl::reduction and the compensated sums.
//
// Reference behaviour being replaced (include/hipSYCL/algorithms/):
//   numeric.hpp:67-135 wg_model_reduction: local size 128, groups =
//     4 x compute units, a grid-stride loop of scalar loads per work-item
//     (util/memory_streaming.hpp:86-91), an LDS tree per group
//     (reduction/wg_model/group_reduction_algorithms.hpp:139-152), then
//     ceil(log_128(groups)) further launches over the group partials
//     (reduction/reduction_engine.hpp:123-142, 211-232) and a final
//     *out = op(init, r) (reduction/reduction_descriptor.hpp:36-38).
//
// MI355X design (one launch):
//   - each work-item keeps UNROLL independent accumulators and issues UNROLL
//     128-bit loads per iteration of a grid-stride loop over 16-byte chunks;
//     misaligned head/tail elements go through a scalar grid-stride loop;
//   - the work-group reduces with DPP inside each wave and one LDS exchange
//     across waves;
//   - every group stores its partial, then takes a ticket; the group that
//     draws the last ticket folds all partials in a fixed order, writes
//     *out and resets the ticket to zero for the next launch;
//   - the partials are self-validating: each is stored with the launch's
//     epoch tag (the scan's status-word layout), and the last group waits
//     until every partial carries the current tag.  So the result does not
//     depend on the partial's store becoming visible before the ticket:
//     no fence and no ordering between the two accesses is needed (a
//     release fence per group would write back the XCD's L2 on gfx950).
//   Floating-point results are deterministic for a given launch geometry.
#pragma once

#include "binary.hpp"
#include "internal.hpp"
#include "unary.hpp"
#include "acpp/device/wave.hpp"

namespace acpp_impl {
namespace reduce_detail {

constexpr int max_waves = 16;  // work-groups of up to 1024 work-items

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Monoids: how a work-item accumulates elements and how partials combine.
// ---------------------------------------------------------------------------
template <class W, class Op>
struct plain_monoid {
  using value = W;
  static constexpr bool compensated = false;
  __device__ __forceinline__ static W identity() { return Op::identity(); }
  __device__ __forceinline__ static W apply(W a, W b) { return Op::apply(a, b); }
  __device__ __forceinline__ static W insert(W acc, W x) { return Op::apply(acc, x); }
  // Dot products: plus of products (numeric.hpp:215-223 defaults).
  __device__ __forceinline__ static W insert2(W acc, W x, W y) {
    return ops::wrap_add(acc, ops::wrap_mul(x, y));
  }
};

// fp32 sum with a running TwoSum error term.
struct comp_value {
  float s;
  float c;
};

struct comp_monoid {
  using value = comp_value;
  static constexpr bool compensated = true;
  __device__ __forceinline__ static comp_value identity() { return {-0.0f, 0.0f}; }

  // Knuth's TwoSum: s + err == a + b exactly.
  __device__ __forceinline__ static float two_sum(float a, float b, float &err) {
#pragma clang fp contract(off)
    float s = a + b;
    float bb = s - a;
    err = (a - (s - bb)) + (b - bb);
    return s;
  }
  __device__ __forceinline__ static comp_value apply(comp_value a, comp_value b) {
#pragma clang fp contract(off)
    float e;
    float s = two_sum(a.s, b.s, e);
    return {s, (a.c + b.c) + e};
  }
  __device__ __forceinline__ static comp_value insert(comp_value acc, float x) {
#pragma clang fp contract(off)
    float e;
    float s = two_sum(acc.s, x, e);
    return {s, acc.c + e};
  }
  // Dot2 (Ogita, Rump, Oishi 2005): TwoProduct via FMA, then TwoSum.
  __device__ __forceinline__ static comp_value insert2(comp_value acc, float x, float y) {
#pragma clang fp contract(off)
    float p = x * y;
    float ep = __builtin_fmaf(x, y, -p);
    float e;
    float s = two_sum(acc.s, p, e);
    return {s, acc.c + (e + ep)};
  }
};

template <class W>
struct double_plus {
  using value = double;
  __device__ __forceinline__ static double identity() { return -0.0; }
  __device__ __forceinline__ static double apply(double a, double b) { return a + b; }
};

// Reduce a value over the work-group; the result is valid in wave 0.
template <class M>
__device__ __forceinline__ typename M::value block_reduce(typename M::value v,
                                                          typename M::value *s_waves) {
  const int nw = blockDim.x >> 6;
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  v = wave::reduce_full<M>(v);
  if (lane == 0) s_waves[w] = v;
  __syncthreads();
  typename M::value r = M::identity();
  if (w == 0) {
    r = (lane < nw) ? s_waves[lane] : M::identity();
    r = wave::reduce_full<M>(r);
  }
  return r;
}

// How the final value is written.
enum final_mode : int {
  final_with_init = 0,     // *out = op(init, r)                (algorithms::reduce)
  final_accumulate = 1,    // *out = op(*out, r)                (sycl::reduction)
  final_identity_init = 2, // *out = r  (== op(identity, r))   (initialize_to_identity)
  final_given_identity = 3, // *out = op(init, r), also for an empty range
                            // (sycl::reduction with a user identity and
                            // initialize_to_identity: init = that identity)
};

template <class T, class M, class OutT>
struct params {
  const T *a;
  const T *b;         // second operand of a dot product
  uint64_t head;      // scalar elements before the 16-byte aligned region
  uint64_t nvec;      // 16-byte chunks in the aligned region
  uint64_t nscalar;   // head + tail elements handled by the scalar loop
  uint64_t tail_start;
  uint64_t *partials;  // one tagged slot per work-group (partial_slot_words)
  sync_header *sync;
  OutT *out;
  typename elem_traits<T>::work init;
  int mode;
  int unary;          // element transform (unary.hpp) of the unary transform_reduce
  int binary;         // pair transform (binary.hpp) of the binary transform_reduce
  uint32_t epoch;     // tag of this launch's partials (>= 1)
  int force_give_up;  // ACPP_VARIANT_FORCE_GIVE_UP (tests of the error path)
};

// ---------------------------------------------------------------------------
// Tagged group partials.  A partial of 32 bits is one 64-bit word
// tag << 32 | bits, one of 64 bits two such words (low half, high half),
// both carrying the tag, with tag = status_tag(epoch, 0, 2) as in the scans'
// status words.  Words left by earlier launches of any kernel on the same
// scratch region carry older epochs and never read as valid.  The words are
// written and read with relaxed device-scope (cross-XCD coherent) atomics.
// ---------------------------------------------------------------------------
constexpr int partial_slot_words = 2;  // 16 bytes per work-group

__device__ __forceinline__ uint64_t partial_tag(uint32_t epoch) {
  return static_cast<uint64_t>(status_tag(epoch, 0, 2)) << 32;
}

template <class V>
__device__ __forceinline__ void publish_partial(uint64_t *slots, uint64_t idx, uint32_t epoch, V v) {
  static_assert(sizeof(V) == 4 || sizeof(V) == 8, "partials are 32 or 64 bits");
  uint64_t *slot = slots + idx * partial_slot_words;
  const uint64_t tag = partial_tag(epoch);
  if constexpr (sizeof(V) == 4) {
    __hip_atomic_store(slot, tag | __builtin_bit_cast(uint32_t, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    __hip_atomic_store(slot, tag | (bits & 0xffffffffu), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slot + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// True (and v set) once the slot holds this launch's partial.
template <class V>
__device__ __forceinline__ bool read_partial(uint64_t *slots, uint64_t idx, uint32_t epoch, V &v) {
  uint64_t *slot = slots + idx * partial_slot_words;
  const uint64_t tag = partial_tag(epoch);
  const uint64_t w0 = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if constexpr (sizeof(V) == 4) {
    if ((w0 & ~0xffffffffull) != tag) return false;
    v = __builtin_bit_cast(V, static_cast<uint32_t>(w0));
  } else {
    const uint64_t w1 = __hip_atomic_load(slot + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((w0 & ~0xffffffffull) != tag || (w1 & ~0xffffffffull) != tag) return false;
    v = __builtin_bit_cast(V, (w1 << 32) | (w0 & 0xffffffffu));
  }
  return true;
}

// The partial of work-group idx, waited for (with backoff) until it carries
// this launch's tag.  A wait that runs out of time, or any wait under
// force_give_up, reports a trip and yields `missing` instead.
template <class V>
__device__ V wait_partial(uint64_t *slots, uint64_t idx, uint32_t epoch, sync_header *sync, bool force_give_up,
                          V missing) {
  V v;
  if (!force_give_up && read_partial(slots, idx, epoch, v)) return v;
  lookback_wait wait;
  wait.begin();
  for (;;) {
    if (force_give_up || wait.failed_round()) {
      report_trip(sync);
      return missing;
    }
    if (read_partial(slots, idx, epoch, v)) return v;
  }
}

template <bool NT, class T>
__device__ __forceinline__ u32x4 load16(const T *base, uint64_t chunk) {
  if constexpr (NT)
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(base) + chunk);
  else
    return reinterpret_cast<const u32x4 *>(base)[chunk];
}

// The same load at 4-byte alignment (global_load_dwordx4 needs dword
// alignment only): the second operand of a dot product of 4- or 8-byte
// elements, whose misalignment may differ from the first operand's.
typedef unsigned int u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
template <bool NT, class T>
__device__ __forceinline__ u32x4 load16_a4(const T *base, uint64_t chunk) {
  if constexpr (NT)
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4_a4 *>(base) + chunk);
  else
    return reinterpret_cast<const u32x4_a4 *>(base)[chunk];
}

template <class T, class W, int VEC>
__device__ __forceinline__ void unpack(u32x4 q, W (&out)[VEC]) {
  using S = typename elem_traits<T>::storage;
  S tmp[VEC];
  __builtin_memcpy(tmp, &q, 16);
#pragma unroll
  for (int k = 0; k < VEC; ++k) out[k] = elem_traits<T>::to_work(__builtin_bit_cast(T, tmp[k]));
}

// XF: apply the element transform p.unary to every element before it is
// combined (unary transform_reduce); never together with DOT.
// BX (with DOT): combine p.binary(x[i], y[i]) (binary.hpp) with Op instead
// of the dot product's plus of products (binary transform_reduce).
template <class T, class M, class Op, class OutT, int UNROLL, bool DOT, bool NT = false, bool XF = false,
          bool BX = false>
__global__ __launch_bounds__(1024) void reduce_kernel(params<T, M, OutT> p) {
  static_assert(!(DOT && XF), "a dot product takes no element transform");
  static_assert(!BX || DOT, "a pair transform needs two operands");
  using W = typename elem_traits<T>::work;
  using V = typename M::value;
  constexpr int VEC = 16 / sizeof(T);
  // Sums of 1-byte integers: the result keeps only the low byte, so the sum
  // of the bytes taken as unsigned values equals the sum of the elements
  // modulo 256.  v_sad_u8 against zero adds four bytes to the accumulator
  // in one instruction (a transform first acts on the packed bytes).
  constexpr bool BYTE_SUM = sizeof(T) == 1 && sizeof(OutT) == 1 && std::is_integral_v<T> &&
                            Op::id == ops::op_plus && !DOT && std::is_same_v<V, W> && sizeof(W) == 4;

  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;

  V acc[UNROLL];
#pragma unroll
  for (int u = 0; u < UNROLL; ++u) acc[u] = M::identity();

  const T *va = p.a + p.head;
  const T *vb = DOT ? p.b + p.head : nullptr;
  uint64_t c = gtid;
  // Main loop: UNROLL independent 128-bit loads (per operand) in flight.
  for (; c + (UNROLL - 1) * stride < p.nvec; c += UNROLL * stride) {
    u32x4 qa[UNROLL];
    u32x4 qb[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      qa[u] = load16<NT>(va, c + u * stride);
      if constexpr (DOT) qb[u] = load16_a4<NT>(vb, c + u * stride);
      // keeps the loads in order ahead of the transform's branch
      if constexpr (BYTE_SUM && XF) __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (BYTE_SUM) {
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        uint32_t d[4] = {qa[u].x, qa[u].y, qa[u].z, qa[u].w};
        if constexpr (XF) unary_detail::apply_bytes<T>(p.unary, d);
        uint32_t a = static_cast<uint32_t>(acc[u]);
#pragma unroll
        for (int k = 0; k < 4; ++k) a = __builtin_amdgcn_sad_u8(d[k], 0u, a);
        acc[u] = static_cast<V>(a);
      }
      continue;
    } else if constexpr (XF && VEC > 4) {
      // Narrow elements (8 or 16 per load): one uniform switch per row.  A
      // whole batch of 8-bit elements widened at once (64 values) spilled.
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        W xa[VEC];
        unpack<T, W, VEC>(qa[u], xa);
        unary_detail::apply_row<T>(p.unary, xa);
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[u] = M::insert(acc[u], xa[k]);
      }
      continue;
    } else if constexpr (XF) {
      // One uniform switch for the whole batch.
      W xt[UNROLL][VEC];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) unpack<T, W, VEC>(qa[u], xt[u]);
      unary_detail::apply_tile<T>(p.unary, xt);
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[u] = M::insert(acc[u], xt[u][k]);
      }
      continue;
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      W xa[VEC];
      unpack<T, W, VEC>(qa[u], xa);
      if constexpr (DOT && BX) {
        W xb[VEC], t[VEC];
        unpack<T, W, VEC>(qb[u], xb);
        binary_detail::apply_row<T, W, VEC>(p.binary, xa, xb, t);
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[u] = M::insert(acc[u], t[k]);
      } else if constexpr (DOT) {
        W xb[VEC];
        unpack<T, W, VEC>(qb[u], xb);
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[u] = M::insert2(acc[u], xa[k], xb[k]);
      } else {
#pragma unroll
        for (int k = 0; k < VEC; ++k) acc[u] = M::insert(acc[u], xa[k]);
      }
    }
  }
  for (; c < p.nvec; c += stride) {
    W xa[VEC];
    unpack<T, W, VEC>(load16<NT>(va, c), xa);
    if constexpr (XF) unary_detail::apply_row<T>(p.unary, xa);
    if constexpr (DOT && BX) {
      W xb[VEC], t[VEC];
      unpack<T, W, VEC>(load16_a4<NT>(vb, c), xb);
      binary_detail::apply_row<T, W, VEC>(p.binary, xa, xb, t);
#pragma unroll
      for (int k = 0; k < VEC; ++k) acc[0] = M::insert(acc[0], t[k]);
    } else if constexpr (DOT) {
      W xb[VEC];
      unpack<T, W, VEC>(load16_a4<NT>(vb, c), xb);
#pragma unroll
      for (int k = 0; k < VEC; ++k) acc[0] = M::insert2(acc[0], xa[k], xb[k]);
    } else {
#pragma unroll
      for (int k = 0; k < VEC; ++k) acc[0] = M::insert(acc[0], xa[k]);
    }
  }
  // Scalar elements: the misaligned head and the tail.
  for (uint64_t s = gtid; s < p.nscalar; s += stride) {
    const uint64_t idx = (s < p.head) ? s : p.tail_start + (s - p.head);
    W xa = elem_traits<T>::to_work(p.a[idx]);
    if constexpr (XF) xa = unary_detail::apply_one<T>(p.unary, xa);
    if constexpr (DOT && BX) {
      acc[0] = M::insert(acc[0], binary_detail::apply_one<T, W>(p.binary, xa, elem_traits<T>::to_work(p.b[idx])));
    } else if constexpr (DOT) {
      acc[0] = M::insert2(acc[0], xa, elem_traits<T>::to_work(p.b[idx]));
    } else {
      acc[0] = M::insert(acc[0], xa);
    }
  }
  V v = acc[0];
#pragma unroll
  for (int u = 1; u < UNROLL; ++u) v = M::apply(v, acc[u]);

  __shared__ V s_waves[max_waves];
  __shared__ int s_last;
  V bv = block_reduce<M>(v, s_waves);
  if (threadIdx.x == 0) {
    // Tagged partial, then the ticket.  The last group waits for each
    // partial's tag, so the order in which the two become visible does not
    // matter; waiting for the store first (vmcnt) only makes such a wait
    // rare.
    publish_partial(p.partials, blockIdx.x, p.epoch, bv);
    wait_own_stores();
    uint32_t prev = __hip_atomic_fetch_add(&p.sync->done_counter, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;

  // Last work-group: fold the partials in a fixed order.
  const int nparts = gridDim.x;
  const bool force = p.force_give_up != 0;
  if constexpr (M::compensated) {
    // Cross-group combine of (sum, error) pairs in FP64.
    __shared__ double s_dwaves[max_waves];
    double d = -0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
      V q = wait_partial<V>(p.partials, i, p.epoch, p.sync, force, M::identity());
      d += static_cast<double>(q.s) + static_cast<double>(q.c);
    }
    d = block_reduce<double_plus<double>>(d, s_dwaves);
    if (threadIdx.x == 0) {
      double r = d;
      if (p.mode == final_with_init) r = static_cast<double>(p.init) + d;
      else if (p.mode == final_accumulate) r = static_cast<double>(*p.out) + d;
      *p.out = static_cast<OutT>(r);
      __hip_atomic_store(&p.sync->done_counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else {
    V r = M::identity();
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
      r = M::apply(r, wait_partial<V>(p.partials, i, p.epoch, p.sync, force, M::identity()));
    __syncthreads();  // s_waves is reused
    r = block_reduce<M>(r, s_waves);
    if (threadIdx.x == 0) {
      if (p.mode == final_with_init || p.mode == final_given_identity) {
        r = Op::apply(p.init, r);
      } else if (p.mode == final_accumulate) {
        r = Op::apply(static_cast<W>(*p.out), r);
      }
      *p.out = static_cast<OutT>(r);
      __hip_atomic_store(&p.sync->done_counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace reduce_detail
}  // namespace acpp_impl
