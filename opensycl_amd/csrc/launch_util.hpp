// Host-side helpers shared by the launchers.
#pragma once

#include <mutex>
#include <unordered_map>

#include "internal.hpp"

namespace acpp_impl {

// Resident work-groups per CU for `kernel` at `block` work-items, cached.
template <class K>
acpp_status occupancy(K kernel, int block, int *out) {
  static std::mutex m;
  static std::unordered_map<uint64_t, int> cache;
  const uint64_t key = (reinterpret_cast<uint64_t>(reinterpret_cast<const void *>(kernel)) << 11) ^
                       static_cast<uint64_t>(block);
  {
    std::lock_guard<std::mutex> lock{m};
    auto it = cache.find(key);
    if (it != cache.end()) {
      *out = it->second;
      return ACPP_SUCCESS;
    }
  }
  int blocks = 0;
  ACPP_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, block, 0));
  if (blocks < 1) blocks = 1;
  std::lock_guard<std::mutex> lock{m};
  cache[key] = blocks;
  *out = blocks;
  return ACPP_SUCCESS;
}

inline bool valid_work_group_size(uint32_t wg) {
  return wg >= 64 && wg <= 1024 && (wg % 64) == 0;
}

// Calls f(OpType{}) for the runtime operator `op` on work type W whose
// element (storage) type is E.  Bitwise and logical operators exist only for
// integers.
template <class W, class E, class F>
acpp_status with_op(int op, F &&f) {
  switch (op) {
    case ops::op_plus: return f(ops::plus<W, E>{});
    case ops::op_multiply: return f(ops::multiply<W, E>{});
    case ops::op_min: return f(ops::minimum<W, E>{});
    case ops::op_max: return f(ops::maximum<W, E>{});
    default: break;
  }
  if constexpr (std::is_integral_v<W>) {
    switch (op) {
      case ops::op_bit_and: return f(ops::bit_and<W, E>{});
      case ops::op_bit_or: return f(ops::bit_or<W, E>{});
      case ops::op_bit_xor: return f(ops::bit_xor<W, E>{});
      case ops::op_logical_and: return f(ops::logical_and<W, E>{});
      case ops::op_logical_or: return f(ops::logical_or<W, E>{});
      default: break;
    }
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "unknown algorithm op " + std::to_string(op));
  } else {
    if (op >= ops::op_bit_and && op <= ops::op_logical_or)
      return set_error(ACPP_ERROR_UNSUPPORTED,
                       "bitwise and logical operators need an integer element type");
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "unknown algorithm op " + std::to_string(op));
  }
}

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace acpp_impl
