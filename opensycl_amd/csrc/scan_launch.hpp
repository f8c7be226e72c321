// Host launcher of the decoupled look-back scan (see scan_common.hpp).
#pragma once

#include "launch_util.hpp"
#include "scan_common.hpp"

namespace acpp_impl {
namespace scan_detail {

// Default rows (16-byte vectors) per work-item of the generic kernel: 8,
// or 4 for 1-byte elements (16 per vector, widened to 32 bits: 8 rows would
// not fit the registers).  Fewer rows mean more tiles per byte, and the
// look-back chain, not the memory, then bounds the scan: 2^28 int8 took
// 153 us with 2 rows.  8-row tiles of 1-byte elements kept as raw bytes in
// registers (and unpacked twice) measured slower than these 4-row tiles:
// 145-151 us against 127-134 us at 2^28 int8 and uint8.  The 8-bit scans
// (operators other than the logical ones) take 8-row tiles on packed
// 16-bit pairs instead (scan8_kernel).
template <class T>
constexpr int default_rows() {
  return sizeof(T) >= 2 ? 8 : 4;
}

// XF: the transform scans (element transform `unary`, see unary.hpp).
// fixed: a caller-provided sync region (the no-allocation entry point) in
// place of one from the group `g` (then unused).
template <class T, class Op, bool XF = false>
acpp_status launch(acpp_allocation_group *g, hipStream_t stream, const T *in, uint64_t n,
                   T *out, bool exclusive, bool has_init, typename elem_traits<T>::work init,
                   const acpp_launch_config *cfg, const char *name, int unary = 0,
                   const fixed_region *fixed = nullptr) {
  using W = typename elem_traits<T>::work;
  constexpr int VEC = 16 / sizeof(T);
  int cus = 0;
  acpp_status st = ACPP_SUCCESS;
  if (fixed) {
    cus = fixed->compute_units;
  } else {
    st = enter_device(g, &cus);
    if (st != ACPP_SUCCESS) return st;
  }

  // 8-bit scans and transform scans with operators other than the logical
  // ones run the packed-pair kernel (8 rows) unless the caller asks for 4
  // rows or the work-group-wide look-back, which only the generic kernel has.
  constexpr bool packed8 = packed8_covers<T, Op>();
  const bool block_lb = cfg && (cfg->flags & variant_block_lookback);
  const int dflt_rows = (packed8 && !block_lb) ? packed8_rows : default_rows<T>();

  // Default work-group size: 1024-wide tiles (fewest tiles, so the fewest
  // look-back links) when the input still gives every CU a few tiles,
  // otherwise 256-wide tiles for parallelism on small inputs.
  uint32_t wg = 256;
  if (cfg && cfg->work_group_size) {
    wg = cfg->work_group_size;
  } else {
    const int rows0 = (cfg && cfg->items_per_thread) ? static_cast<int>(cfg->items_per_thread) : dflt_rows;
    if (ceil_div(n, static_cast<uint64_t>(1024) * rows0 * VEC) >= 4ull * static_cast<uint64_t>(cus)) wg = 1024;
  }
  if (!valid_work_group_size(wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                     std::string(name) + ": work_group_size must be a multiple of 64 in [64, 1024]");
  int rows = (cfg && cfg->items_per_thread) ? static_cast<int>(cfg->items_per_thread) : dflt_rows;
  constexpr bool tunable = (Op::id == ops::op_plus) && sizeof(T) >= 4 && !XF;
  if (rows != dflt_rows) {
    const bool ok = tunable ? (rows == 2 || rows == 4 || rows == 8)
                            : (packed8 && rows == default_rows<T>());
    if (!ok)
      return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                       std::string(name) + ": items_per_thread (rows per work-item) must be " +
                           std::to_string(dflt_rows) +
                           (tunable ? ", 2 or 4"
                                    : packed8 ? (block_lb ? " with the work-group look-back" : " or 4")
                                              : " for this element type and operator"));
  }

  const uint64_t tile_elems = static_cast<uint64_t>(wg) * rows * VEC;
  const uint64_t num_tiles = ceil_div(n, tile_elems);
  if (num_tiles >= (1ull << 31))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": problem too large");

  using SIO = status_io<W>;
  static_assert(SIO::words <= status_stride_words, "status slot too small");
  const uint64_t status_bytes = num_tiles * status_stride_words * sizeof(uint64_t);
  const size_t used = sizeof(sync_header) + status_bytes;
  void *region = nullptr;
  size_t cap = 0;
  region_state *rs = nullptr;
  region_launch rl;
  if (fixed) {
    if (used > fixed->capacity)
      return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                       std::string(name) + ": " + std::to_string(num_tiles) + " tiles need " +
                           std::to_string(used) + " bytes of scratch; the static scratch holds " +
                           std::to_string(fixed->capacity) + " (at most " +
                           std::to_string((fixed->capacity - sizeof(sync_header)) /
                                          (status_stride_words * sizeof(uint64_t))) +
                           " tiles of " + std::to_string(tile_elems) + " elements)");
    region = fixed->ptr;
    cap = fixed->capacity;
    rs = fixed->rs;
    st = begin_region_launch_flag(fixed->trip_flag, region, used, cap, rs, stream, &rl);
  } else {
    st = group_obtain_sync(g, used, stream, &region, &cap, &rs);
    if (st != ACPP_SUCCESS) return st;
    st = begin_region_launch(g, region, used, cap, rs, stream, &rl);
  }
  if (st != ACPP_SUCCESS) return st;

  params<T> p;
  p.in = in;
  p.out = out;
  p.n = n;
  p.init = init;
  p.has_init = has_init ? 1 : 0;
  p.exclusive = exclusive ? 1 : 0;
  // 16-byte vectors need 16-byte alignment for 1- and 2-byte elements, only
  // element (4-byte) alignment for 4- and 8-byte elements.
  p.vec_io = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) &
              (sizeof(T) >= 4 ? 3u : 15u)) == 0;
  p.block_lookback = (cfg && (cfg->flags & variant_block_lookback)) ? 1 : 0;
  p.force_give_up = (cfg && (cfg->flags & ACPP_VARIANT_FORCE_GIVE_UP)) ? 1 : 0;
  p.unary = unary;
  p.num_tiles = static_cast<uint32_t>(num_tiles);
  p.epoch = rl.epoch;
  p.ticket_base = rl.ticket_base;
  p.sync = static_cast<sync_header *>(region);
  p.status = reinterpret_cast<uint64_t *>(static_cast<char *>(region) + sizeof(sync_header));

  auto run = [&](auto kernel) -> acpp_status {
    hipLaunchKernelGGL(kernel, dim3(static_cast<uint32_t>(num_tiles)), dim3(wg), 0, stream, p);
    acpp_status s = check_launch(name);
    end_region_launch(rs, rl, static_cast<uint32_t>(num_tiles), s);
    return s;
  };
  if constexpr (packed8) {
    if (rows == packed8_rows) return run(scan8_kernel<T, Op, XF>);
  }
  if constexpr (tunable) {
    if (rows == 2) return run(scan_kernel<T, Op, 2>);
    if (rows == 4) return run(scan_kernel<T, Op, 4>);
    if (cfg && (cfg->flags & ACPP_VARIANT_NONTEMPORAL)) return run(scan_kernel<T, Op, default_rows<T>(), false, true>);
  }
  return run(scan_kernel<T, Op, default_rows<T>(), XF>);
}

template <class T>
acpp_status scan_entry(acpp_allocation_group *g, hipStream_t stream, const T *first, uint64_t n,
                       T *d_first, int op, bool exclusive, const T *init_ptr, T init_value,
                       const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (n > 0 && (!first || !d_first))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": first or d_first is NULL");
  bool has_init = exclusive || init_ptr != nullptr;
  T init = exclusive ? init_value : (init_ptr ? *init_ptr : T{});
  return with_op<W, T>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    // decoupled_lookback_scan.hpp:565-566: an empty problem enqueues nothing.
    if (n == 0) return clear_error();
    acpp_status s = launch<T, Op>(g, stream, first, n, d_first, exclusive, has_init,
                                  elem_traits<T>::to_work(init), cfg, name);
    return s == ACPP_SUCCESS ? clear_error() : s;
  });
}

// transform_inclusive_scan / transform_exclusive_scan (numeric.hpp:299-327):
// the scan of unary(x[i]) for a transform code of unary.hpp.
template <class T>
acpp_status transform_scan_entry(acpp_allocation_group *g, hipStream_t stream, const T *first,
                                 uint64_t n, T *d_first, int op, int unary, bool exclusive,
                                 const T *init_ptr, T init_value, const acpp_launch_config *cfg,
                                 const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (n > 0 && (!first || !d_first))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": first or d_first is NULL");
  if (!unary_detail::valid_code<T>(unary))
    return set_error(std::is_integral_v<T> || unary < 0 || unary >= unary_detail::u_count
                         ? ACPP_ERROR_INVALID_ARGUMENT
                         : ACPP_ERROR_UNSUPPORTED,
                     std::string(name) + ": unknown transform for this element type");
  bool has_init = exclusive || init_ptr != nullptr;
  T init = exclusive ? init_value : (init_ptr ? *init_ptr : T{});
  return with_op<W, T>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    // scan.hpp:85-132 on decoupled_lookback_scan.hpp:565-566: an empty
    // problem enqueues nothing.
    if (n == 0) return clear_error();
    acpp_status s;
    if (unary == unary_detail::u_identity)
      s = launch<T, Op>(g, stream, first, n, d_first, exclusive, has_init, elem_traits<T>::to_work(init),
                        cfg, name);
    else
      s = launch<T, Op, true>(g, stream, first, n, d_first, exclusive, has_init,
                              elem_traits<T>::to_work(init), cfg, name, unary);
    return s == ACPP_SUCCESS ? clear_error() : s;
  });
}

}  // namespace scan_detail
}  // namespace acpp_impl
