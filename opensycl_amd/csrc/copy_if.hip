// Stream compaction: algorithms::copy_if
// (include/hipSYCL/algorithms/algorithm.hpp:178-243).
//
// Reference: an exclusive scan (generate_scan_process) of size_t flags
// pred(x[i]) ? 1 : 0, whose result processor writes x[i] to
// d_first[scan[i]] and, at the last element, the number of elements copied
// (algorithm.hpp:194-242).  An empty range writes 0 to *num_elements_copied
// and returns (algorithm.hpp:184-188).
//
// MI355X design: one launch on the decoupled look-back machinery of the scan
// (scan_common.hpp).  A tile is B work-items x R rows x one 16-byte vector;
// each work-item evaluates the predicate on its vectors into a bitmask,
// counts with popcount, the work-group ranks the selected elements with DPP
// wave scans and one LDS exchange, one wave looks back over the tile
// counts (64-bit, one 128-byte status line per tile) for the tile's output
// offset.  Meanwhile every wave compacts the selected elements of each of
// its rows into that row's own LDS segment (128 KiB of segments at the
// default 1024 x 8 tile); the segments are then stored with consecutive
// lanes on consecutive output elements (coalesced whatever the selection
// pattern).  The last tile writes the total count.
//
// The predicate is data, not a functor (a functor cannot cross the C-ABI):
// comparison with a constant, a bit-mask test, or a stencil array for any
// predicate the caller evaluates with its own kernel.
#include "launch_util.hpp"
#include "scan_common.hpp"

namespace acpp_impl {
namespace copy_if_detail {

using scan_detail::st_aggregate;
using scan_detail::st_prefix;
using scan_detail::status_stride_words;
using scan_detail::u32x4;

template <class T>
struct params {
  const T *in;
  const uint8_t *stencil;  // ACPP_PRED_STENCIL
  T *out;
  uint64_t *num_copied;     // may be NULL
  uint64_t n;
  T operand;
  T mask;
  int vec_io;
  int stencil_vec;  // stencil aligned to VEC bytes: one load per 16-byte data vector
  uint32_t num_tiles;
  uint32_t epoch;
  uint32_t ticket_base;
  int force_give_up;  // ACPP_VARIANT_FORCE_GIVE_UP (tests of the error path)
  sync_header *sync;
  uint64_t *status;
};

template <int PRED, class T>
__device__ __forceinline__ bool eval(const params<T> &p, T x, uint64_t idx) {
  if constexpr (PRED == ACPP_PRED_NONZERO) return x != T(0);
  else if constexpr (PRED == ACPP_PRED_EQ) return x == p.operand;
  else if constexpr (PRED == ACPP_PRED_NE) return x != p.operand;
  else if constexpr (PRED == ACPP_PRED_LT) return x < p.operand;
  else if constexpr (PRED == ACPP_PRED_LE) return x <= p.operand;
  else if constexpr (PRED == ACPP_PRED_GT) return x > p.operand;
  else if constexpr (PRED == ACPP_PRED_GE) return x >= p.operand;
  else if constexpr (PRED == ACPP_PRED_MASK_EQ) {
    if constexpr (std::is_integral_v<T>) return (x & p.mask) == p.operand;
    else return false;
  } else {
    return p.stencil[idx] != 0;  // ACPP_PRED_STENCIL
  }
}

// Selection bits of the tile: bit k of fm[r] = element k of row r is copied.
template <int PRED, class T, int R, int VEC>
__device__ __forceinline__ void select_bits(const params<T> &p, const T (&v)[R][VEC], uint64_t base, bool full,
                                            uint32_t (&fm)[R]) {
  const int B = blockDim.x;
  if constexpr (PRED == ACPP_PRED_STENCIL && (VEC == 2 || VEC == 4)) {
    if (full && p.stencil_vec) {
      // The VEC stencil bytes of a data vector in one load.
      using SW = std::conditional_t<VEC == 4, uint32_t, uint16_t>;
      const SW *sw = reinterpret_cast<const SW *>(p.stencil + base);
      SW bytes[R];
#pragma unroll
      for (int r = 0; r < R; ++r) bytes[r] = sw[r * B + threadIdx.x];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        fm[r] = 0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) fm[r] |= (((bytes[r] >> (8 * k)) & 0xffu) != 0u ? 1u : 0u) << k;
      }
      return;
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    fm[r] = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
      const uint64_t idx = base + static_cast<uint64_t>(r * B + threadIdx.x) * VEC + k;
      if ((full || idx < p.n) && eval<PRED>(p, v[r][k], idx)) fm[r] |= 1u << k;
    }
  }
}

// Element k of a 16-byte vector of T.
template <class T>
__device__ __forceinline__ T elem(const u32x4 &q, int k) {
  T tmp[16 / sizeof(T)];
  __builtin_memcpy(tmp, &q, 16);
  return tmp[k];
}

// select_bits for 1- and 2-byte elements kept packed in their vectors (q);
// in_range: the tile's elements before the end of the input.
template <int PRED, class T, int R>
__device__ __forceinline__ void select_bits_packed(const params<T> &p, const u32x4 (&q)[R], uint64_t base,
                                                   uint32_t in_range, bool full, uint32_t (&fm)[R]) {
  constexpr int VEC = 16 / sizeof(T);
  const int B = blockDim.x;
  if constexpr (PRED == ACPP_PRED_STENCIL) {
    if (full && p.stencil_vec) {
      // The VEC stencil bytes of a data vector in one load.
      using SW = std::conditional_t<VEC == 16, u32x4, uint64_t>;
      const SW *sw = reinterpret_cast<const SW *>(p.stencil + base);
      SW bytes[R];
#pragma unroll
      for (int r = 0; r < R; ++r) bytes[r] = sw[r * B + threadIdx.x];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        uint32_t words[4] = {};
        __builtin_memcpy(words, &bytes[r], sizeof(SW));
        fm[r] = 0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) fm[r] |= (((words[k >> 2] >> (8 * (k & 3))) & 0xffu) != 0u ? 1u : 0u) << k;
      }
      return;
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    fm[r] = 0;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
      const uint32_t i = static_cast<uint32_t>(r * B + threadIdx.x) * VEC + k;
      if ((full || i < in_range) && eval<PRED>(p, elem<T>(q[r], k), base + i)) fm[r] |= 1u << k;
    }
  }
}

// Shared words of a copy_if work-group.
template <int R>
struct tile_shared {
  uint32_t tile;
  uint32_t wave[R * scan_detail::max_waves];  // per (row, wave): counts, then offsets
  uint32_t agg;
  uint64_t excl;
};

// One tile.  PRED and FULL are compile-time: a full tile has no bounds
// checks and one predicate's code, since with one work-group per CU (the
// LDS segments take 128 KiB) this phase is not overlapped with memory and
// its instruction count is on the critical path.
template <int PRED, bool FULL, class T, int R>
__device__ __forceinline__ void copy_if_tile(const params<T> &p, uint32_t tile, tile_shared<R> &sh, T *s_stage) {
  using CountOp = ops::plus<uint32_t, uint32_t>;
  using SIO = scan_detail::status_io<uint64_t>;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int SEG = 64 * VEC;  // LDS room of one wave row's selection
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int nw = B >> 6;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const uint64_t tile_elems = static_cast<uint64_t>(B) * R * VEC;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;

  // 1- and 2-byte elements stay packed in their 16-byte vectors (four
  // registers a row) and are extracted where they are used: 8 rows of 8 or
  // 16 unpacked elements spill.
  constexpr bool NARROW = sizeof(T) < 4;
  u32x4 q[R];
  T v[NARROW ? 1 : R][VEC];
  if constexpr (FULL) {
    const scan_detail::u32x4_a4 *src = reinterpret_cast<const scan_detail::u32x4_a4 *>(p.in + base);
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = src[r * B + tid];
    if constexpr (!NARROW) {
#pragma unroll
      for (int r = 0; r < R; ++r) __builtin_memcpy(v[r], &q[r], 16);
    }
  } else if constexpr (NARROW) {
    const uint32_t in_range = static_cast<uint32_t>(p.n > base ? min(p.n - base, tile_elems) : 0);
    const T *in_tile = p.in + base;
    using U = std::make_unsigned_t<T>;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t first = static_cast<uint32_t>(r * B + tid) * VEC;
      if (p.vec_io && first + VEC <= in_range) {
        // a whole vector inside the input: one load
        q[r] = *reinterpret_cast<const scan_detail::u32x4_a4 *>(in_tile + first);
      } else {
        // the vector across the end (or an input not 4-byte aligned):
        // element by element, each ORed into its word
        uint32_t d[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
          constexpr int sz = static_cast<int>(sizeof(T));
          if (first + k < in_range)
            d[(k * sz) >> 2] |= static_cast<uint32_t>(static_cast<U>(in_tile[first + k])) << (8 * ((k * sz) & 3));
        }
        q[r] = u32x4{d[0], d[1], d[2], d[3]};
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < VEC; ++k) {
        const uint64_t idx = base + static_cast<uint64_t>(r * B + tid) * VEC + k;
        v[r][k] = (idx < p.n) ? p.in[idx] : T(0);
      }
  }
  uint32_t fm[R];
  if constexpr (NARROW) {
    const uint32_t in_range = FULL ? static_cast<uint32_t>(tile_elems)
                                   : static_cast<uint32_t>(p.n > base ? min(p.n - base, tile_elems) : 0);
    select_bits_packed<PRED>(p, q, base, in_range, FULL, fm);
    // The compaction below extracts the elements again from the packed
    // vectors instead of keeping the predicate's extracted copies alive.
#pragma unroll
    for (int r = 0; r < R; ++r) {
      uint32_t a = q[r].x, b = q[r].y, c = q[r].z, d = q[r].w;
      asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
      q[r] = u32x4{a, b, c, d};
    }
  } else {
    select_bits<PRED>(p, v, base, FULL, fm);
  }
  // Ranks inside each wave row; the (row, wave) counts go to LDS.
  uint32_t wexcl[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t c = static_cast<uint32_t>(__builtin_popcount(fm[r]));
    const uint32_t incl = wave::inclusive_scan<CountOp>(c);
    wexcl[r] = incl - c;
    if (lane == 63) sh.wave[r * nw + w] = incl;
  }
  __syncthreads();
  // Wave 0 turns the (row, wave) counts into offsets and publishes the
  // tile's count before anything else, since every later tile's look-back
  // may be waiting for it.
  uint32_t agg0 = 0;
  if (w == 0) {
    agg0 = scan_detail::tile_prefixes<CountOp, uint32_t>(R * nw, sh.wave);
    if (lane == 0) {
      if (tile == 0) {
        if (p.num_tiles > 1) SIO::publish(p.status, 0, p.epoch, st_prefix, static_cast<uint64_t>(agg0));
        sh.excl = 0;
      } else {
        SIO::publish(p.status, tile, p.epoch, st_aggregate, static_cast<uint64_t>(agg0));
      }
      sh.agg = agg0;
    }
  }
  // Compaction: each wave row's selected elements go, in order, to that
  // row's own LDS segment (room for all 64 x VEC of them), at the rank
  // inside the wave row.  It needs no tile-wide offsets, so it runs while
  // wave 0 looks back; the stores below then copy segment after segment.
#pragma unroll
  for (int r = 0; r < R; ++r) {
    uint32_t pos = static_cast<uint32_t>((r * nw + w) * SEG) + wexcl[r];
    const uint32_t bits = fm[r];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
      if (bits & (1u << k)) {
        if constexpr (NARROW) s_stage[pos++] = elem<T>(q[r], k);
        else s_stage[pos++] = v[r][k];
      }
    }
  }
  // Wave 0 looks back for the tile's output offset.
  if (w == 0 && tile != 0) {
    wait_own_stores();  // the tile's count is out before the first poll
    const uint64_t e = scan_detail::lookback<ops::plus<uint64_t, uint64_t>, uint64_t, scan_detail::lookback_slots>(
        p.status, tile, p.epoch, p.sync, p.force_give_up != 0);
    if (lane == 0) {
      if (tile + 1 < p.num_tiles) SIO::publish(p.status, tile, p.epoch, st_prefix, e + agg0);
      sh.excl = e;
    }
  }
  __syncthreads();
  // Each wave stores its own segments: a segment's elements are contiguous
  // in the output, at the offset of its (row, wave) pair, and consecutive
  // lanes take consecutive elements (coalesced stores, conflict-free LDS
  // reads).
  const uint32_t count = __builtin_amdgcn_readfirstlane(sh.agg);
  const uint64_t ex = sh.excl;
  const uint64_t tile_prefix =
      (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(ex >> 32))) << 32) |
      __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(ex));
  if (tile + 1 == p.num_tiles && tid == 0 && p.num_copied) *p.num_copied = tile_prefix + count;
  T *dst = p.out + tile_prefix;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int seg = r * nw + w;
    const uint32_t off = sh.wave[seg];
    const uint32_t end = (seg + 1 < R * nw) ? sh.wave[seg + 1] : count;
    const T *src = s_stage + seg * SEG;
    if constexpr (sizeof(T) == 1) {
      // 4-byte stores of EPW elements: the elements before the segment's
      // first 4-byte aligned output address and those after its last whole
      // word are stored one by one (the words around them hold elements of
      // neighbouring segments); each word's source bytes are two aligned
      // LDS words shifted together.  (2-byte elements measured slower this
      // way: one store per element stays.)
      constexpr uint32_t EPW = 4 / sizeof(T);
      const uint32_t len = end - off;
      T *d = dst + off;
      const uint32_t mis = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(d) / sizeof(T)) % EPW;
      const uint32_t head = min(len, (EPW - mis) % EPW);
      if (lane < head) d[lane] = src[lane];
      const uint32_t words = (len - head) / EPW;
      uint32_t *dw = reinterpret_cast<uint32_t *>(d + head);
      const uint32_t *lds32 = reinterpret_cast<const uint32_t *>(s_stage);
      const uint32_t b0 = static_cast<uint32_t>((seg * SEG + head) * sizeof(T));  // LDS byte of word 0
      for (uint32_t j = lane; j < words; j += 64) {
        const uint32_t b = b0 + j * 4;
        const uint32_t sh = b & 3u;
        const uint32_t lo = lds32[b >> 2];
        const uint32_t hi = lds32[(b >> 2) + (sh ? 1u : 0u)];  // in the segment when sh != 0
        dw[j] = __builtin_amdgcn_alignbyte(hi, lo, sh);
      }
      const uint32_t t0 = head + words * EPW;
      if (lane < len - t0) d[t0 + lane] = src[t0 + lane];
    } else {
      for (uint32_t j = lane; j < end - off; j += 64) dst[off + j] = src[j];
    }
  }
}

template <class T, int R, int PRED>
__global__ __launch_bounds__(1024) void copy_if_kernel(params<T> p) {
  constexpr int VEC = 16 / sizeof(T);
  __shared__ tile_shared<R> sh;
  extern __shared__ unsigned char s_stage_raw[];  // R x nw segments of 64 x VEC elements (B * R * VEC)
  T *s_stage = reinterpret_cast<T *>(s_stage_raw);

  if (threadIdx.x == 0)
    sh.tile = __hip_atomic_fetch_add(&p.sync->tile_counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
              p.ticket_base;
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(sh.tile);
  if (tile >= p.num_tiles) return;
  const uint64_t tile_elems = static_cast<uint64_t>(blockDim.x) * R * VEC;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;
  if (p.vec_io && base + tile_elems <= p.n)
    copy_if_tile<PRED, true>(p, tile, sh, s_stage);
  else
    copy_if_tile<PRED, false>(p, tile, sh, s_stage);
}

__global__ void store_zero_kernel(uint64_t *p) { *p = 0; }

template <class T>
constexpr int rows() {
  return 8;
}

template <class T>
acpp_status entry(acpp_allocation_group *g, hipStream_t stream, const T *in, uint64_t n, const uint8_t *stencil,
                  T *out, int pred, T operand, T mask, uint64_t *num_copied, const acpp_launch_config *cfg,
                  const char *name) {
  constexpr int VEC = 16 / sizeof(T);
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (pred < ACPP_PRED_NONZERO || pred > ACPP_PRED_STENCIL)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": unknown predicate");
  if (pred == ACPP_PRED_MASK_EQ && !std::is_integral_v<T>)
    return set_error(ACPP_ERROR_UNSUPPORTED, std::string(name) + ": mask predicates take integer elements");
  if (n > 0 && (!in || !out || (pred == ACPP_PRED_STENCIL && !stencil)))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": NULL pointer argument");
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  if (n == 0) {
    // algorithm.hpp:184-188: the count of an empty range is 0.
    // Written by a kernel: asynchronous, and allowed under stream capture.
    if (num_copied) {
      hipLaunchKernelGGL(store_zero_kernel, dim3(1), dim3(1), 0, stream, num_copied);
      st = check_launch(name);
      if (st != ACPP_SUCCESS) return st;
    }
    return clear_error();
  }
  const uint32_t wg = (cfg && cfg->work_group_size) ? cfg->work_group_size : 1024;
  if (!valid_work_group_size(wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": bad work_group_size");
  if (cfg && cfg->items_per_thread && static_cast<int>(cfg->items_per_thread) != rows<T>())
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": items_per_thread must be 8 (or 0)");
  const uint64_t tile_elems = static_cast<uint64_t>(wg) * rows<T>() * VEC;
  const uint64_t num_tiles = ceil_div(n, tile_elems);
  if (num_tiles >= (1ull << 31)) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": problem too large");

  void *region = nullptr;
  size_t cap = 0;
  region_state *rs = nullptr;
  const size_t used = sizeof(sync_header) + num_tiles * status_stride_words * sizeof(uint64_t);
  st = group_obtain_sync(g, used, stream, &region, &cap, &rs);
  if (st != ACPP_SUCCESS) return st;
  region_launch rl;
  st = begin_region_launch(g, region, used, cap, rs, stream, &rl);
  if (st != ACPP_SUCCESS) return st;
  params<T> p;
  p.in = in;
  p.stencil = stencil;
  p.out = out;
  p.num_copied = num_copied;
  p.n = n;
  p.stencil_vec = stencil && (reinterpret_cast<uintptr_t>(stencil) % VEC) == 0;
  p.operand = operand;
  p.mask = mask;
  // 16-byte input loads need only element alignment for 4- and 8-byte
  // elements (see scan_detail::u32x4_a4).
  p.vec_io = (reinterpret_cast<uintptr_t>(in) & 3) == 0;
  p.num_tiles = static_cast<uint32_t>(num_tiles);
  p.epoch = rl.epoch;
  p.ticket_base = rl.ticket_base;
  p.force_give_up = (cfg && (cfg->flags & ACPP_VARIANT_FORCE_GIVE_UP)) ? 1 : 0;
  p.sync = static_cast<sync_header *>(region);
  p.status = reinterpret_cast<uint64_t *>(static_cast<char *>(region) + sizeof(sync_header));
  const size_t stage_bytes = static_cast<size_t>(tile_elems) * sizeof(T);  // 128 KiB at 1024 x 8
  const dim3 grid(static_cast<uint32_t>(num_tiles));
  constexpr int R = rows<T>();
  switch (pred) {
    case ACPP_PRED_NONZERO:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_NONZERO>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_EQ:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_EQ>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_NE:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_NE>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_LT:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_LT>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_LE:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_LE>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_GT:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_GT>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_GE:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_GE>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    case ACPP_PRED_MASK_EQ:
      // rejected above for floating-point elements
      if constexpr (std::is_integral_v<T>)
        hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_MASK_EQ>), grid, dim3(wg), stage_bytes, stream, p);
      break;
    default:
      hipLaunchKernelGGL((copy_if_kernel<T, R, ACPP_PRED_STENCIL>), grid, dim3(wg), stage_bytes, stream, p);
      break;
  }
  st = check_launch(name);
  end_region_launch(rs, rl, static_cast<uint32_t>(num_tiles), st);
  return st == ACPP_SUCCESS ? clear_error() : st;
}

}  // namespace copy_if_detail
}  // namespace acpp_impl

using namespace acpp_impl;

extern "C" {

#define ACPP_DEFINE_COPY_IF(tag, T)                                                                       \
  acpp_status acpp_copy_if_##tag(acpp_allocation_group *scratch, acpp_stream_t stream, const T *first,   \
                                 uint64_t n, T *d_first, int pred, T operand, T mask,                    \
                                 uint64_t *num_copied, const acpp_launch_config *config) {               \
    if (pred == ACPP_PRED_STENCIL)                                                                        \
      return set_error(ACPP_ERROR_INVALID_ARGUMENT, "acpp_copy_if_" #tag ": use acpp_copy_if_stencil_" #tag); \
    return copy_if_detail::entry<T>(scratch, static_cast<hipStream_t>(stream), first, n, nullptr, d_first,  \
                                    pred, operand, mask, num_copied, config, "acpp_copy_if_" #tag);        \
  }                                                                                                       \
  acpp_status acpp_copy_if_stencil_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,           \
                                         const T *first, uint64_t n, const uint8_t *stencil, T *d_first, \
                                         uint64_t *num_copied, const acpp_launch_config *config) {       \
    return copy_if_detail::entry<T>(scratch, static_cast<hipStream_t>(stream), first, n, stencil, d_first,  \
                                    ACPP_PRED_STENCIL, T(0), T(0), num_copied, config,                    \
                                    "acpp_copy_if_stencil_" #tag);                                         \
  }

ACPP_DEFINE_COPY_IF(i8, int8_t)
ACPP_DEFINE_COPY_IF(u8, uint8_t)
ACPP_DEFINE_COPY_IF(i16, int16_t)
ACPP_DEFINE_COPY_IF(u16, uint16_t)
ACPP_DEFINE_COPY_IF(i32, int32_t)
ACPP_DEFINE_COPY_IF(u32, uint32_t)
ACPP_DEFINE_COPY_IF(i64, int64_t)
ACPP_DEFINE_COPY_IF(u64, uint64_t)
ACPP_DEFINE_COPY_IF(f32, float)
ACPP_DEFINE_COPY_IF(f64, double)

}  // extern "C"
