// Host launcher of the device-wide reduction (see reduce_common.hpp).
#pragma once

#include "launch_util.hpp"
#include "reduce_common.hpp"

namespace acpp_impl {
namespace reduce_detail {

// Element types whose 16-byte chunks are loaded with one instruction.
template <class T>
constexpr int vec_width() {
  return 16 / static_cast<int>(sizeof(typename elem_traits<T>::storage));
}

// XF: unary transform_reduce with element transform `unary` (unary.hpp).
// BX (with DOT): binary transform_reduce with pair transform `binary`
// (binary.hpp), combined with Op.
template <class T, class M, class Op, class OutT, bool DOT, bool XF = false, bool BX = false>
acpp_status launch(acpp_allocation_group *g, hipStream_t stream, const T *a,
                   const T *b, uint64_t n, OutT *out,
                   typename elem_traits<T>::work init, int mode,
                   const acpp_launch_config *cfg, const char *name, int unary = 0, int binary = 0) {
  using V = typename M::value;
  constexpr int VEC = vec_width<T>();
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;

  // Library defaults, measured on MI355X at 2^28 elements.  Every
  // reduction streams with one resident work-group per CU: 512 work-items x
  // 4 loads in flight for a single operand, 1024 x 8 for a dot product.  The
  // generic 256-wide, occupancy-filling grid was slower for every operator
  // (int32 max: 213 us against 170 us; int32 squares through the transform
  // form: 211 us against 170 us).  1-byte elements take 1024 work-items: 16
  // elements per load make the per-element arithmetic the limit, and 512
  // work-items left 2^30 int8 at 301 us against 190 us.  Inputs larger than
  // the 256 MiB last-level cache stream with non-temporal loads (except
  // 1-byte elements, below); that choice is the library's only when the
  // whole config is left at zero.  The compensated dot products keep 2 loads
  // in flight: their per-load Dot2 state makes 8 loads need 128 VGPRs and
  // spill, and 2 loads (45 VGPRs) measured fastest (2 x 2^28 fp32: 318 us
  // against 351 us with 8; bf16: 165 us against 177 us).
  constexpr bool tuned = (Op::id == ops::op_plus) && !XF && !BX;
  constexpr bool comp_dot = DOT && std::is_same_v<M, comp_monoid>;
  constexpr bool narrow8 = sizeof(typename elem_traits<T>::storage) == 1;
  const bool cfg_given = cfg && (cfg->work_group_size || cfg->items_per_thread || cfg->groups_per_cu ||
                                 cfg->flags);
  const uint32_t wg = (cfg && cfg->work_group_size) ? cfg->work_group_size : ((DOT || narrow8) ? 1024 : 512);
  if (!valid_work_group_size(wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                     std::string(name) + ": work_group_size must be a multiple of 64 in [64, 1024]");
  const uint32_t unroll = (cfg && cfg->items_per_thread) ? cfg->items_per_thread : (tuned && DOT ? (comp_dot ? 2 : 8) : 4);
  const uint32_t groups_per_cu = (cfg && cfg->groups_per_cu) ? cfg->groups_per_cu : 1;
  const uint64_t stream_bytes = n * sizeof(typename elem_traits<T>::storage) * (DOT ? 2 : 1);
  // 1-byte sums are bound by per-element arithmetic, not by the cache:
  // non-temporal loads tied at 2^30 int8 and lost at 2^28 (52.7 us against
  // 42.8 us), so they keep ordinary loads.
  const bool nt = cfg_given ? (cfg->flags & ACPP_VARIANT_NONTEMPORAL) != 0
                            : (!narrow8 && stream_bytes >= (256ull << 20));
  if (unroll != 2 && unroll != 4 && unroll != 8 && (unroll != 16 || DOT))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                     std::string(name) + ": items_per_thread must be 2, 4, 8 or 16 (16 only "
                                         "for single-operand reductions)");
  if (XF && unroll != 4)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                     std::string(name) + ": items_per_thread must be 4 with an element transform");

  // Scalar head, then the 16-byte aligned region of a, then a scalar tail.
  const uintptr_t abytes = reinterpret_cast<uintptr_t>(a);
  uint64_t head = ((16 - (abytes & 15)) & 15) / sizeof(typename elem_traits<T>::storage);
  bool vec_ok = (abytes % sizeof(typename elem_traits<T>::storage)) == 0;
  // The head aligns a; b is read with dword-aligned 16-byte loads, so for
  // 4- and 8-byte elements it only has to be element-aligned (2-byte
  // elements need b misaligned exactly like a).
  if (DOT) {
    constexpr size_t es = sizeof(typename elem_traits<T>::storage);
    const uintptr_t bbytes = reinterpret_cast<uintptr_t>(b);
    vec_ok = vec_ok && (es >= 4 ? (bbytes % es) == 0 : (bbytes & 15) == (abytes & 15));
  }
  if (!vec_ok || head > n) head = n;
  const uint64_t nvec = (n - head) / VEC;
  const uint64_t tail_start = head + nvec * VEC;
  const uint64_t nscalar = head + (n - tail_start);

  params<T, M, OutT> p;
  p.a = a;
  p.b = b;
  p.head = head;
  p.nvec = nvec;
  p.nscalar = nscalar;
  p.tail_start = tail_start;
  p.out = out;
  p.init = init;
  p.mode = mode;
  p.unary = unary;
  p.binary = binary;
  p.force_give_up = (cfg && (cfg->flags & ACPP_VARIANT_FORCE_GIVE_UP)) ? 1 : 0;

  auto run = [&](auto kernel) -> acpp_status {
    int resident = 1;
    acpp_status s = occupancy(kernel, static_cast<int>(wg), &resident);
    if (s != ACPP_SUCCESS) return s;
    uint64_t max_groups = static_cast<uint64_t>(resident) * cus;
    if (groups_per_cu) {
      uint64_t req = static_cast<uint64_t>(groups_per_cu) * cus;
      if (req < max_groups) max_groups = req;
    }
    uint64_t needed = ceil_div(nvec + nscalar, static_cast<uint64_t>(wg) * unroll);
    uint64_t grid = needed < 1 ? 1 : needed;
    if (grid > max_groups) grid = max_groups;

    // One sync region: the header (done ticket) and a tagged slot per
    // work-group for its partial.
    static_assert(sizeof(V) <= partial_slot_words * 4, "partial does not fit its slot");
    const size_t used = sizeof(sync_header) + grid * partial_slot_words * sizeof(uint64_t);
    void *region = nullptr;
    size_t cap = 0;
    region_state *rs = nullptr;
    s = group_obtain_sync(g, used, stream, &region, &cap, &rs);
    if (s != ACPP_SUCCESS) return s;
    region_launch rl;
    s = begin_region_launch(g, region, used, cap, rs, stream, &rl);
    if (s != ACPP_SUCCESS) return s;
    p.sync = static_cast<sync_header *>(region);
    p.partials = reinterpret_cast<uint64_t *>(static_cast<char *>(region) + sizeof(sync_header));
    p.epoch = rl.epoch;
    hipLaunchKernelGGL(kernel, dim3(static_cast<uint32_t>(grid)), dim3(wg), 0, stream, p);
    s = check_launch(name);
    end_region_launch(rs, rl, 0, s);
    return s;
  };

  // Only sums (the tuned hot path) get every unroll depth and the
  // non-temporal variant; other operators use 4 to bound the number of
  // kernel instantiations.
  if constexpr (tuned) {
    if (nt) {
      if (unroll == 2) return run(reduce_kernel<T, M, Op, OutT, 2, DOT, true>);
      if (unroll == 8) return run(reduce_kernel<T, M, Op, OutT, 8, DOT, true>);
      if (unroll == 16) return run(reduce_kernel<T, M, Op, OutT, 16, DOT, true>);
      return run(reduce_kernel<T, M, Op, OutT, 4, DOT, true>);
    }
    if (unroll == 2) return run(reduce_kernel<T, M, Op, OutT, 2, DOT>);
    if (unroll == 8) return run(reduce_kernel<T, M, Op, OutT, 8, DOT>);
    if (unroll == 16) return run(reduce_kernel<T, M, Op, OutT, 16, DOT>);
  }
  // Other operators and the transform forms: 4 loads in flight,
  // non-temporal above 256 MiB as for sums (2^28 int32 max: 170 us with
  // ordinary loads).
  if (nt) return run(reduce_kernel<T, M, Op, OutT, 4, DOT, true, XF, BX>);
  return run(reduce_kernel<T, M, Op, OutT, 4, DOT, false, XF, BX>);
}

// algorithms::reduce / sycl::reduction entry for element type T.
template <class T, class OutT, class InitT>
acpp_status reduce_entry(acpp_allocation_group *g, hipStream_t stream, const T *first,
                         uint64_t n, OutT *out, InitT init, int op, int mode,
                         const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  using E = std::conditional_t<std::is_floating_point_v<W>, W, T>;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": out is NULL");
  if (n > 0 && !first) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": first is NULL");
  if (n == 0 && mode == final_with_init) {
    // numeric.hpp:165-169: empty input leaves *out untouched.
    return with_op<W, E>(op, [&](auto) { return clear_error(); });
  }
  return with_op<W, E>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    acpp_status s = launch<T, plain_monoid<W, Op>, Op, OutT, false>(
        g, static_cast<hipStream_t>(stream), first, nullptr, n, out, static_cast<W>(init), mode, cfg,
        name);
    return s == ACPP_SUCCESS ? clear_error() : s;
  });
}

// Compensated sum (is_dot == false, b ignored) or dot product.
template <class T>
acpp_status compensated_entry(acpp_allocation_group *g, hipStream_t stream, const T *a,
                              const T *b, bool is_dot, uint64_t n, float *out, float init,
                              const acpp_launch_config *cfg, const char *name) {
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": out is NULL");
  if (n > 0 && (!a || (is_dot && !b)))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": input is NULL");
  if (n == 0) return clear_error();
  acpp_status s;
  using P = ops::plus<float, float>;
  if (is_dot)
    s = launch<T, comp_monoid, P, float, true>(g, stream, a, b, n, out, init, final_with_init, cfg, name);
  else
    s = launch<T, comp_monoid, P, float, false>(g, stream, a, nullptr, n, out, init, final_with_init, cfg, name);
  return s == ACPP_SUCCESS ? clear_error() : s;
}

template <class T, class OutT>
acpp_status dot_entry(acpp_allocation_group *g, hipStream_t stream, const T *a, uint64_t n,
                      const T *b, OutT *out, typename elem_traits<T>::work init,
                      const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": out is NULL");
  if (n > 0 && (!a || !b)) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": input is NULL");
  if (n == 0) return clear_error();
  using P = ops::plus<W, W>;
  acpp_status s = launch<T, plain_monoid<W, P>, P, OutT, true>(g, stream, a, b, n, out, init,
                                                               final_with_init, cfg, name);
  return s == ACPP_SUCCESS ? clear_error() : s;
}

// Unary transform_reduce (numeric.hpp:193-212): *out = op(init, op over
// unary(x[i])); an empty input leaves *out untouched (numeric.hpp:165-169).
template <class T>
acpp_status transform_reduce_entry(acpp_allocation_group *g, hipStream_t stream, const T *first,
                                   uint64_t n, T *out, T init, int op, int unary,
                                   const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": out is NULL");
  if (n > 0 && !first) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": first is NULL");
  if (!unary_detail::valid_code<T>(unary))
    return set_error(std::is_integral_v<T> || unary < 0 || unary >= unary_detail::u_count
                         ? ACPP_ERROR_INVALID_ARGUMENT
                         : ACPP_ERROR_UNSUPPORTED,
                     std::string(name) + ": unknown transform for this element type");
  return with_op<W, T>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    if (n == 0) return clear_error();
    acpp_status s;
    if (unary == unary_detail::u_identity)
      s = launch<T, plain_monoid<W, Op>, Op, T, false>(g, stream, first, nullptr, n, out,
                                                       elem_traits<T>::to_work(init), final_with_init,
                                                       cfg, name);
    else
      s = launch<T, plain_monoid<W, Op>, Op, T, false, true>(g, stream, first, nullptr, n, out,
                                                             elem_traits<T>::to_work(init),
                                                             final_with_init, cfg, name, unary);
    return s == ACPP_SUCCESS ? clear_error() : s;
  });
}

// Binary transform_reduce (numeric.hpp:170-191) with reduction operator
// `op` (plus, multiplies, minimum or maximum) and pair transform `binary`
// (binary.hpp): *out = op(init, op over binary(a[i], b[i])); an empty input
// leaves *out untouched (numeric.hpp:165-169).  plus with multiplies is
// the dot product's own kernel.
template <class T>
acpp_status binary_transform_reduce_entry(acpp_allocation_group *g, hipStream_t stream, const T *a, uint64_t n,
                                          const T *b, T *out, T init, int op, int binary,
                                          const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": out is NULL");
  if (n > 0 && (!a || !b)) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": input is NULL");
  if (!binary_detail::valid_code<T>(binary))
    return set_error(std::is_integral_v<T> || binary < 0 || binary >= binary_detail::b_count
                         ? ACPP_ERROR_INVALID_ARGUMENT
                         : ACPP_ERROR_UNSUPPORTED,
                     std::string(name) + ": unknown pair transform for this element type");
  if (op < ops::op_plus || op > ops::op_logical_or)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": unknown operator");
  if (op > ops::op_max)
    return set_error(ACPP_ERROR_UNSUPPORTED,
                     std::string(name) + ": the binary form reduces with plus, multiplies, minimum or maximum");
  if (n == 0) return clear_error();
  if (op == ops::op_plus && binary == binary_detail::b_multiplies)
    return dot_entry<T, T>(g, stream, a, n, b, out, elem_traits<T>::to_work(init), cfg, name);
  return with_op<W, T>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    if constexpr (Op::id > ops::op_max) {
      return set_error(ACPP_ERROR_UNSUPPORTED, std::string(name) + ": unsupported operator");
    } else {
      acpp_status s = launch<T, plain_monoid<W, Op>, Op, T, true, false, true>(
          g, stream, a, b, n, out, elem_traits<T>::to_work(init), final_with_init, cfg, name, 0, binary);
      return s == ACPP_SUCCESS ? clear_error() : s;
    }
  });
}

}  // namespace reduce_detail
}  // namespace acpp_impl
