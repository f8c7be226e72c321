// Several sycl::reduction objects in one kernel (handler::parallel_for with
// more than one reduction, include/hipSYCL/sycl/handler.hpp:799-975; test
// tests/sycl/reduction.cpp:116-208 runs a plus and a multiplies reduction in
// one parallel_for).
//
// One launch reduces up to four inputs of the same length, each with its
// own operator and output: every work-group streams the inputs one after
// the other (the operator switch sits outside each streaming loop, so the
// loops are the same tight 16-byte-load loops as the single reduction),
// reduces each with DPP and one LDS exchange, and publishes one partial per
// input; the last work-group to finish (fence-free ticket, as in
// reduce_common.hpp) folds every input's partials in a fixed order and
// writes the results.  Inputs may alias (several operators over one array).
#include "launch_util.hpp"
#include "reduce_common.hpp"

namespace acpp_impl {
namespace multi_detail {

using reduce_detail::block_reduce;
using reduce_detail::max_waves;
using reduce_detail::partial_slot_words;
using reduce_detail::plain_monoid;
using reduce_detail::publish_partial;
using reduce_detail::wait_partial;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int max_inputs = 4;
constexpr int unroll = 4;

template <class T>
struct params {
  using W = typename elem_traits<T>::work;
  const T *in[max_inputs];
  T *out[max_inputs];
  int op[max_inputs];
  int count;
  uint64_t n;
  // Per input: 0 out = op(out, fold); 1 out = fold (initialize_to_identity);
  // 2 out = op(identity[k], fold) (initialize_to_identity with a given
  // identity, which an empty range yields).
  int init_mode[max_inputs];
  W identity[max_inputs];
  uint64_t *partials; // tagged slots [count][gridDim.x] (reduce_common.hpp)
  sync_header *sync;
  uint32_t epoch;     // tag of this launch's partials (>= 1)
  int force_give_up;  // ACPP_VARIANT_FORCE_GIVE_UP (tests of the error path)
};

// Calls f(Op{}) with the operator functor for `code`.
template <class W, class F>
__device__ __forceinline__ void dispatch(int code, F &&f) {
  switch (code) {
    case ops::op_plus: f(ops::plus<W, W>{}); return;
    case ops::op_multiply: f(ops::multiply<W, W>{}); return;
    case ops::op_min: f(ops::minimum<W, W>{}); return;
    case ops::op_max: f(ops::maximum<W, W>{}); return;
    default: break;
  }
  if constexpr (std::is_integral_v<W>) {
    switch (code) {
      case ops::op_bit_and: f(ops::bit_and<W, W>{}); return;
      case ops::op_bit_or: f(ops::bit_or<W, W>{}); return;
      case ops::op_bit_xor: f(ops::bit_xor<W, W>{}); return;
      case ops::op_logical_and: f(ops::logical_and<W, W>{}); return;
      default: f(ops::logical_or<W, W>{}); return;
    }
  }
}

// Fold of one input: `unroll` independent accumulators over 16-byte loads
// (`unroll` in flight), non-temporal when NT, then the scalar tail.  (With one
// accumulator and a bounds check per load, one fused int32 reduction of 2^28
// elements took 175 us against 154 us for the plain reduce.)
template <class T, class Op, bool NT>
__device__ typename elem_traits<T>::work stream_fold(const T *in, uint64_t n) {
  using W = typename elem_traits<T>::work;
  constexpr int VEC = 16 / sizeof(T);
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  W acc[unroll];
#pragma unroll
  for (int u = 0; u < unroll; ++u) acc[u] = Op::identity();
  uint64_t scalar_from = 0;
  if ((reinterpret_cast<uintptr_t>(in) & 15) == 0) {
    const uint64_t nvec = n / VEC;
    const u32x4 *src = reinterpret_cast<const u32x4 *>(in);
    auto fold = [&](u32x4 q, W &a) {
      T t[VEC];
      __builtin_memcpy(t, &q, 16);
#pragma unroll
      for (int k = 0; k < VEC; ++k) a = Op::apply(a, elem_traits<T>::to_work(t[k]));
    };
    uint64_t i = gid;
    for (; i + (unroll - 1) * stride < nvec; i += unroll * stride) {
      u32x4 q[unroll];
#pragma unroll
      for (int u = 0; u < unroll; ++u) {
        if constexpr (NT)
          q[u] = __builtin_nontemporal_load(src + i + u * stride);
        else
          q[u] = src[i + u * stride];
      }
#pragma unroll
      for (int u = 0; u < unroll; ++u) fold(q[u], acc[u]);
    }
    for (; i < nvec; i += stride) fold(src[i], acc[0]);
    scalar_from = nvec * VEC;
  }
  for (uint64_t i = scalar_from + gid; i < n; i += stride) acc[0] = Op::apply(acc[0], elem_traits<T>::to_work(in[i]));
  W v = acc[0];
#pragma unroll
  for (int u = 1; u < unroll; ++u) v = Op::apply(v, acc[u]);
  return v;
}

template <class T, bool NT>
__global__ __launch_bounds__(1024) void multi_kernel(params<T> p) {
  using W = typename elem_traits<T>::work;
  __shared__ W s_waves[max_waves];
  __shared__ bool s_last;
  for (int k = 0; k < p.count; ++k) {
    __syncthreads();  // s_waves of the previous input has been read
    dispatch<W>(p.op[k], [&](auto opv) {
      using Op = decltype(opv);
      W v = stream_fold<T, Op, NT>(p.in[k], p.n);
      v = block_reduce<plain_monoid<W, Op>>(v, s_waves);
      if (threadIdx.x == 0)
        publish_partial(p.partials, static_cast<uint64_t>(k) * gridDim.x + blockIdx.x, p.epoch, v);
    });
  }
  if (threadIdx.x == 0) {
    // The last group waits for each partial's tag; waiting for the stores
    // first only makes such a wait rare.
    wait_own_stores();
    const uint32_t prev =
        __hip_atomic_fetch_add(&p.sync->done_counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  for (int k = 0; k < p.count; ++k) {
    __syncthreads();  // s_waves of the previous input has been read
    dispatch<W>(p.op[k], [&](auto opv) {
      using Op = decltype(opv);
      W r = Op::identity();
      for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x)
        r = Op::apply(r, wait_partial<W>(p.partials, static_cast<uint64_t>(k) * gridDim.x + i, p.epoch, p.sync,
                                         p.force_give_up != 0, Op::identity()));
      r = block_reduce<plain_monoid<W, Op>>(r, s_waves);
      if (threadIdx.x == 0) {
        if (p.init_mode[k] == 0) r = Op::apply(elem_traits<T>::to_work(*p.out[k]), r);
        else if (p.init_mode[k] == 2) r = Op::apply(p.identity[k], r);
        *p.out[k] = static_cast<T>(r);
      }
    });
  }
  if (threadIdx.x == 0)
    __hip_atomic_store(&p.sync->done_counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// n == 0: outputs with initialize_to_identity become their identity (the
// given one, or the operator's); the others keep their value.
template <class T>
__global__ void identity_kernel(params<T> p) {
  using W = typename elem_traits<T>::work;
  const int k = threadIdx.x;
  if (k < p.count && p.init_mode[k] != 0)
    dispatch<W>(p.op[k], [&](auto opv) {
      *p.out[k] = static_cast<T>(p.init_mode[k] == 2 ? p.identity[k] : decltype(opv)::identity());
    });
}

// identities: NULL, or one value per input (used where the input's
// initialize_to_identity flag is set).  init_flags: one flag per input.
template <class T>
acpp_status entry(acpp_allocation_group *g, hipStream_t stream, uint32_t count, const T *const *inputs,
                  uint64_t n, T *const *outs, const acpp_algorithm_op *op_codes, const T *identities,
                  const int *init_flags, const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (count < 1 || count > static_cast<uint32_t>(max_inputs))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": count must be 1..4");
  if (!inputs || !outs || !op_codes || !init_flags)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": NULL array argument");
  params<T> p{};
  p.count = static_cast<int>(count);
  p.n = n;
  bool any_init = false;
  for (uint32_t k = 0; k < count; ++k) {
    p.init_mode[k] = init_flags[k] ? (identities ? 2 : 1) : 0;
    p.identity[k] = identities ? elem_traits<T>::to_work(identities[k]) : W{};
    any_init = any_init || init_flags[k];
  }
  for (uint32_t k = 0; k < count; ++k) {
    const int op = op_codes[k];
    if (op < ops::op_plus || op > ops::op_logical_or)
      return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": unknown operator");
    if (std::is_floating_point_v<W> && op > ops::op_max)
      return set_error(ACPP_ERROR_UNSUPPORTED, std::string(name) + ": floating-point reductions take plus, "
                                                                   "multiply, min or max");
    if (!outs[k] || (n > 0 && !inputs[k]))
      return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": NULL input or output");
    p.in[k] = inputs[k];
    p.out[k] = outs[k];
    p.op[k] = op;
  }
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  if (n == 0) {
    // reduction_descriptor.hpp:36-53: with initialize_to_identity the output
    // is the identity; without it the output keeps its value.
    if (any_init) {
      hipLaunchKernelGGL(identity_kernel<T>, dim3(1), dim3(64), 0, stream, p);
      st = check_launch(name);
      if (st != ACPP_SUCCESS) return st;
    }
    return clear_error();
  }
  const uint32_t wg = (cfg && cfg->work_group_size) ? cfg->work_group_size : 512;
  if (!valid_work_group_size(wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": bad work_group_size");
  int resident = 1;
  st = occupancy(multi_kernel<T, false>, static_cast<int>(wg), &resident);
  if (st != ACPP_SUCCESS) return st;
  const uint32_t gpc = (cfg && cfg->groups_per_cu) ? cfg->groups_per_cu : 1;
  const uint64_t cap = static_cast<uint64_t>(std::min<uint32_t>(gpc, static_cast<uint32_t>(resident))) * cus;
  constexpr int VEC = 16 / sizeof(T);
  const uint64_t needed = ceil_div(ceil_div(n, static_cast<uint64_t>(VEC)), static_cast<uint64_t>(wg) * unroll);
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(cap, needed)));
  const size_t used = sizeof(sync_header) + static_cast<size_t>(grid) * count * partial_slot_words * sizeof(uint64_t);
  void *region = nullptr;
  size_t cap_bytes = 0;
  region_state *rs = nullptr;
  st = group_obtain_sync(g, used, stream, &region, &cap_bytes, &rs);
  if (st != ACPP_SUCCESS) return st;
  region_launch rl;
  st = begin_region_launch(g, region, used, cap_bytes, rs, stream, &rl);
  if (st != ACPP_SUCCESS) return st;
  p.sync = static_cast<sync_header *>(region);
  p.partials = reinterpret_cast<uint64_t *>(static_cast<char *>(region) + sizeof(sync_header));
  p.epoch = rl.epoch;
  p.force_give_up = (cfg && (cfg->flags & ACPP_VARIANT_FORCE_GIVE_UP)) ? 1 : 0;
  // Non-temporal loads once the inputs are larger than the 256 MiB
  // last-level cache, as for the plain reductions.
  const bool nt = static_cast<uint64_t>(count) * n * sizeof(T) >= (256ull << 20);
  if (nt)
    hipLaunchKernelGGL((multi_kernel<T, true>), dim3(grid), dim3(wg), 0, stream, p);
  else
    hipLaunchKernelGGL((multi_kernel<T, false>), dim3(grid), dim3(wg), 0, stream, p);
  st = check_launch(name);
  end_region_launch(rs, rl, 0, st);
  return st == ACPP_SUCCESS ? clear_error() : st;
}

}  // namespace multi_detail
}  // namespace acpp_impl

using namespace acpp_impl;

extern "C" {

#define ACPP_DEFINE_MULTI(tag, T)                                                                       \
  acpp_status acpp_reduction_multi_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,         \
                                         uint32_t count, const T *const *inputs, uint64_t n,           \
                                         T *const *outs, const acpp_algorithm_op *ops,                 \
                                         int initialize_to_identity, const acpp_launch_config *config) { \
    const int flags[multi_detail::max_inputs] = {initialize_to_identity, initialize_to_identity,       \
                                                 initialize_to_identity, initialize_to_identity};      \
    return multi_detail::entry<T>(scratch, static_cast<hipStream_t>(stream), count, inputs, n, outs, ops, \
                                  nullptr, flags, config, "acpp_reduction_multi_" #tag);               \
  }                                                                                                     \
  acpp_status acpp_reduction_multi_identity_##tag(                                                     \
      acpp_allocation_group *scratch, acpp_stream_t stream, uint32_t count, const T *const *inputs,    \
      uint64_t n, T *const *outs, const acpp_algorithm_op *ops, const T *identities,                   \
      const int *initialize_to_identity, const acpp_launch_config *config) {                           \
    return multi_detail::entry<T>(scratch, static_cast<hipStream_t>(stream), count, inputs, n, outs, ops, \
                                  identities, initialize_to_identity, config,                          \
                                  "acpp_reduction_multi_identity_" #tag);                             \
  }
ACPP_DEFINE_MULTI(i8, int8_t)
ACPP_DEFINE_MULTI(u8, uint8_t)
ACPP_DEFINE_MULTI(i16, int16_t)
ACPP_DEFINE_MULTI(u16, uint16_t)
ACPP_DEFINE_MULTI(i32, int32_t)
ACPP_DEFINE_MULTI(u32, uint32_t)
ACPP_DEFINE_MULTI(i64, int64_t)
ACPP_DEFINE_MULTI(u64, uint64_t)
ACPP_DEFINE_MULTI(f32, float)
ACPP_DEFINE_MULTI(f64, double)

}  // extern "C"
