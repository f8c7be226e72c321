// Wave64 cross-lane primitives for gfx950 (CDNA4).
//
// A CDNA wavefront has 64 lanes arranged as 4 rows of 16.  Data-parallel
// primitives (DPP) move a register between lanes inside the VALU pipe without
// touching LDS:
//   row_shr:k     lane i of a row reads lane i-k of the same row
//   row_bcast:15  every lane of row r+1 reads lane 15 of row r
//   row_bcast:31  every lane of rows 2 and 3 reads lane 31
//   wave_shr:1    lane i reads lane i-1 of the whole wave
// A lane whose source does not exist keeps the "old" operand, which we set to
// the operator's identity.  Six DPP steps give an inclusive scan of the whole
// wave (Hillis-Steele inside each row, then two broadcasts across rows).
//
// The scan is also correct for a partially filled wave as long as the active
// lanes form a prefix 0..k-1 (the shape of the last wave of a work-group
// whose size is not a multiple of 64): every active lane only ever reads
// lower lanes, which are active too.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace acpp {
namespace wave {

constexpr int size = 64;

// DPP control words (GFX9 encoding).
enum : int {
  row_shr1 = 0x111,
  row_shr2 = 0x112,
  row_shr4 = 0x114,
  row_shr8 = 0x118,
  wave_shr1 = 0x138,
  row_bcast15 = 0x142,
  row_bcast31 = 0x143,
};

template <int Ctrl, int RowMask, int BankMask, class T>
__device__ __forceinline__ T dpp(T old, T src) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "DPP moves 32 or 64 bits");
  if constexpr (sizeof(T) == 4) {
    int r = __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                        __builtin_bit_cast(int, src), Ctrl,
                                        RowMask, BankMask, false);
    return __builtin_bit_cast(T, r);
  } else {
    long long r = __builtin_amdgcn_update_dpp(
        __builtin_bit_cast(long long, old), __builtin_bit_cast(long long, src),
        Ctrl, RowMask, BankMask, false);
    return __builtin_bit_cast(T, r);
  }
}

__device__ __forceinline__ int lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Value of lane `l` (uniform across the wave).
template <class T>
__device__ __forceinline__ T read_lane(T x, int l) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "readlane moves 32 or 64 bits");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
  } else {
    uint64_t u = __builtin_bit_cast(uint64_t, x);
    uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffu), l);
    uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), l);
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}

// Value of the first active lane (uniform across the wave).
template <class T>
__device__ __forceinline__ T read_first_lane(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
  } else {
    uint64_t u = __builtin_bit_cast(uint64_t, x);
    uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<int>(u & 0xffffffffu));
    uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32));
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}

__device__ __forceinline__ uint64_t ballot(bool p) {
  return __builtin_amdgcn_ballot_w64(p);
}

// Index of the highest active lane.
__device__ __forceinline__ int last_active_lane() {
  uint64_t exec = __builtin_amdgcn_read_exec();
  return 63 - __builtin_clzll(exec);
}

// Inclusive scan across the wave, lane order = element order.  Op::apply is
// called as apply(earlier, later).
template <class Op, class W>
__device__ __forceinline__ W inclusive_scan(W x) {
  const W id = Op::identity();
  x = Op::apply(dpp<row_shr1, 0xf, 0xf>(id, x), x);
  x = Op::apply(dpp<row_shr2, 0xf, 0xf>(id, x), x);
  x = Op::apply(dpp<row_shr4, 0xf, 0xf>(id, x), x);
  x = Op::apply(dpp<row_shr8, 0xf, 0xf>(id, x), x);
  x = Op::apply(dpp<row_bcast15, 0xa, 0xf>(id, x), x);
  x = Op::apply(dpp<row_bcast31, 0xc, 0xf>(id, x), x);
  return x;
}

// Shift an inclusive scan by one lane to make it exclusive (lane 0 gets the
// identity).
template <class Op, class W>
__device__ __forceinline__ W shift_up(W incl) {
  return dpp<wave_shr1, 0xf, 0xf>(Op::identity(), incl);
}

// Reduction of a full wave; the result is uniform.
template <class Op, class W>
__device__ __forceinline__ W reduce_full(W x) {
  return read_lane(inclusive_scan<Op>(x), 63);
}

// Reduction over the active lanes (a prefix of the wave); the result is
// uniform.
template <class Op, class W>
__device__ __forceinline__ W reduce_active(W x) {
  return read_lane(inclusive_scan<Op>(x), last_active_lane());
}

}  // namespace wave
}  // namespace acpp
