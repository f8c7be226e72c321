// Single-pass decoupled look-back scan for gfx950.
//
// Reference behaviour being replaced
// (include/hipSYCL/algorithms/scan/decoupled_lookback_scan.hpp):
//   - a status-initialisation kernel (584-589), then the scan kernel (633-642);
//   - work-groups of 128 work-items x 16 invocations; each invocation is a
//     Kogge-Stone scan in LDS with 2 barriers per step plus a broadcast
//     (49-72, 122-135, 181-217): ~277 barriers per 2048-element tile;
//   - dynamic tile ids from an atomic counter (379-391) so that every tile's
//     predecessors have started;
//   - per-tile status {invalid, aggregate, prefix} published with acq_rel
//     atomics next to separate value arrays (30-47, 407-447), and a serial
//     single-lane look-back (219-257) combining op(earlier, later).
//
// MI355X design (one launch, no initialisation kernel):
//   - a tile is B work-items x R rows x one 16-byte vector per work-item per
//     row; the R loads per work-item are issued back to back (coalesced
//     128-bit loads, R x 16 bytes in flight per lane);
//   - each work-item scans its vector in registers, each wave scans the
//     per-lane totals with DPP, and one LDS exchange of the wave totals
//     (one barrier) gives every work-item its offset inside the tile;
//   - the status of a tile is self-validating: value bits and a tag
//     (status_tag in internal.hpp) share one 64-bit word (two words for 64-bit
//     values, each carrying the tag), written and read with relaxed
//     device-scope atomics.  Words from earlier launches carry an older
//     epoch and read as "invalid", so nothing has to be cleared;
//   - each tile's status has a 128-byte line of its own.  Every tile in
//     flight polls the statuses of its predecessors, so packed statuses
//     put all those polls on a handful of lines; on MI355X spreading them
//     out took the 2^28 int32 scan from ~490 us to ~410 us;
//   - the look-back is done by one wave: 64 predecessors are polled at
//     once, the nearest "prefix" among them is found with a ballot and the
//     window is folded with DPP, with no work-group barrier per poll.  The
//     wave first waits for its own aggregate store to complete
//     (wait_own_stores in internal.hpp).  A work-group-wide window
//     (blockDim.x tiles per poll) is kept as a variant;
//   - nothing is reset at the end of a launch: the host tracks, per scratch
//     region, the epoch of the last launch and how many tile tickets have
//     been drawn, and passes both as kernel arguments (a per-tile "done"
//     counter would add a second same-address atomic per tile);
//   - a poll round that finds an unfinished predecessor backs off before
//     the next, and a wall-clock watchdog keeps a look-back from spinning
//     forever (lookback_wait in internal.hpp).
#pragma once

#include "internal.hpp"
#include "unary.hpp"

namespace acpp_impl {
namespace scan_detail {

constexpr int max_waves = 16;
// Status slots polled per lane and look-back round (window 64 x this).
// Wider windows put more polls on the memory system than they save in
// rounds: with all loads in flight together, the 2^28 int32 scan took
// 377 us with 64 and 384 us with 128 on MI355X.
constexpr int lookback_slots = 1;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// The same 16-byte vector with only 4-byte alignment: global_load_dwordx4 /
// global_store_dwordx4 need dword alignment only, so tiles of 4- and
// 8-byte elements take the vector path at any element-aligned address
// (a view such as x[1:]); the element-wise path took 1.75x as long.
typedef unsigned int u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

enum : uint32_t { st_invalid = 0, st_aggregate = 1, st_prefix = 2 };

__device__ __forceinline__ uint64_t atomic_load_u64(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_u64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile status encoding.
template <class W, int Bytes = sizeof(W)>
struct status_io;

// 64-bit words per tile status slot: one 128-byte line.
constexpr int status_stride_words = 16;

template <class W>
struct status_io<W, 4> {
  static constexpr int words = 1;
  __device__ __forceinline__ static void publish(uint64_t *st, uint32_t tile, uint32_t epoch,
                                                 uint32_t flag, W v) {
    const uint64_t tag = status_tag(epoch, 0, flag);
    atomic_store_u64(st + static_cast<uint64_t>(tile) * status_stride_words,
                     (tag << 32) | __builtin_bit_cast(uint32_t, v));
  }
  __device__ __forceinline__ static uint32_t read(const uint64_t *st, uint32_t tile,
                                                  uint32_t epoch, W &v) {
    const uint64_t w = atomic_load_u64(st + static_cast<uint64_t>(tile) * status_stride_words);
    const uint32_t tag = static_cast<uint32_t>(w >> 32);
    if (tag_epoch(tag) != epoch) return st_invalid;
    v = __builtin_bit_cast(W, static_cast<uint32_t>(w));
    return tag_status(tag);
  }
};

template <class W>
struct status_io<W, 8> {
  static constexpr int words = 2;
  __device__ __forceinline__ static void publish(uint64_t *st, uint32_t tile, uint32_t epoch,
                                                 uint32_t flag, W v) {
    const uint64_t tag = status_tag(epoch, 0, flag);
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    uint64_t *slot = st + static_cast<uint64_t>(tile) * status_stride_words;
    atomic_store_u64(slot, (tag << 32) | (bits & 0xffffffffu));
    atomic_store_u64(slot + 1, (tag << 32) | (bits >> 32));
  }
  __device__ __forceinline__ static uint32_t read(const uint64_t *st, uint32_t tile,
                                                  uint32_t epoch, W &v) {
    const uint64_t *slot = st + static_cast<uint64_t>(tile) * status_stride_words;
    const uint64_t w0 = atomic_load_u64(slot);
    const uint64_t w1 = atomic_load_u64(slot + 1);
    const uint32_t t0 = static_cast<uint32_t>(w0 >> 32);
    const uint32_t t1 = static_cast<uint32_t>(w1 >> 32);
    // Both halves must come from the same publication.
    if (t0 != t1 || tag_epoch(t0) != epoch) return st_invalid;
    v = __builtin_bit_cast(W, (w1 << 32) | (w0 & 0xffffffffu));
    return tag_status(t0);
  }
};

// Exclusive prefix of `tile`: op over all predecessors, found by a wave-wide
// look-back with a window of 64 * K predecessors.  Lane l polls tiles
// tile - 1 - l - 64 k (k < K) with all K loads in flight at once; the
// nearest "prefix" in the window is found with ballots, and the window up
// to it is folded with DPP, with no work-group barrier per poll.  Must be
// called by a whole wave; the result is uniform.
template <class Op, class W, int K>
__device__ W lookback(const uint64_t *status, uint32_t tile, uint32_t epoch,
                      sync_header *sync, bool force_give_up) {
  using SIO = status_io<W>;
  const int lane = wave::lane_id();
  W excl = Op::identity();
  int64_t base = static_cast<int64_t>(tile) - 1;
  lookback_wait wait;
  wait.begin();
  bool give_up = force_give_up;  // uniform
  if (give_up && lane == 0) report_trip(sync);
  for (;;) {
    W val[K];
    int p;
    for (;;) {
      uint32_t flag[K];
      // Branch-free polls, so that all K loads are in flight together:
      // positions before tile 0 read tile 0's slot and act as an empty
      // prefix.
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int64_t j = base - lane - 64 * k;
        W x = Op::identity();
        const uint32_t f = SIO::read(status, j >= 0 ? static_cast<uint32_t>(j) : 0u, epoch, x);
        flag[k] = (j >= 0 && !give_up) ? f : static_cast<uint32_t>(st_prefix);
        val[k] = (j >= 0) ? x : Op::identity();
      }
      // p: window position of the nearest prefix (64 * K: none).  Every
      // position up to it must hold a valid status.
      p = 64 * K;
      bool bad = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint64_t pm = wave::ballot(flag[k] == st_prefix);
        const uint64_t im = wave::ballot(flag[k] == st_invalid);
        if (p == 64 * K) {
          const int pk = pm ? __builtin_ctzll(pm) : 64;
          const uint64_t need = (pk >= 63) ? ~0ull : ((2ull << pk) - 1);
          if (im & need) bad = true;
          if (pm) p = 64 * k + pk;
        }
      }
      if (!bad) break;
      if (wait.failed_round()) {
        give_up = true;
        if (lane == 0) report_trip(sync);
      }
    }
    W x = Op::identity();
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (64 * k + lane <= p) x = Op::apply(x, val[k]);
    const W window = wave::reduce_full<Op>(x);
    excl = Op::apply(window, excl);
    if (p < 64 * K) return excl;
    base -= 64 * K;
  }
}

// Exclusive prefixes, in element order, of a tile's per-(row, wave) totals:
// s_tot[r * nw + w] holds the totals on entry and the exclusive prefixes on
// return (P = R * nw <= 256 entries).  Called by one whole wave, which scans
// them with DPP (1, 2 or 4 consecutive entries per lane); returns the tile
// aggregate, uniform across the wave.  This replaces a walk over all R * nw
// totals by every work-item (128 dependent LDS reads and combines per
// work-item for 1024 x 8 tiles), which also delayed the aggregate's
// publication.
template <class Op, class W>
__device__ W tile_prefixes(int P, W *s_tot) {
  const int lane = wave::lane_id();
  const int per = P > 128 ? 4 : P > 64 ? 2 : 1;
  const int a = lane * per;
  W v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = (j < per && a + j < P) ? s_tot[a + j] : Op::identity();
  W sum = v[0];
#pragma unroll
  for (int j = 1; j < 4; ++j)
    if (j < per) sum = Op::apply(sum, v[j]);
  const W incl = wave::inclusive_scan<Op>(sum);
  W excl = wave::shift_up<Op>(incl);
  const W agg = wave::read_lane(incl, 63);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j < per && a + j < P) s_tot[a + j] = excl;
    excl = Op::apply(excl, v[j]);
  }
  return agg;
}

// Exclusive prefix of `tile` with a look-back window as wide as the whole
// work-group: work-item t polls tile (tile - 1 - t), so one round trip to
// memory covers blockDim.x predecessors.  The number of tiles a look-back
// must cover grows with the number of tiles in flight, and every poll costs a
// device-scope round trip, so the window width bounds the scan's throughput
// at about (window / poll latency) tiles per unit time.  Must be called by
// the whole work-group; the result is uniform.
template <class Op, class W>
__device__ W block_lookback(const uint64_t *status, uint32_t tile, uint32_t epoch,
                            uint64_t *s_masks, W *s_red, sync_header *sync, bool force_give_up) {
  using SIO = status_io<W>;
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int nw = B >> 6;
  const int w = tid >> 6;
  const int lane = tid & 63;
  W excl = Op::identity();
  int64_t base = static_cast<int64_t>(tile) - 1;
  __shared__ int s_give_up;
  lookback_wait wait;
  wait.begin();
  bool give_up = force_give_up;  // uniform across the work-group
  if (give_up && tid == 0) report_trip(sync);
  for (;;) {
    const int64_t j = base - tid;  // work-item 0 = nearest predecessor
    W val = Op::identity();
    int p;
    for (;;) {
      uint32_t flag = st_prefix;  // positions before tile 0 act as an empty prefix
      if (j >= 0) flag = SIO::read(status, static_cast<uint32_t>(j), epoch, val);
      if (j < 0) val = Op::identity();
      if (give_up) flag = st_prefix;
      const uint64_t pm = wave::ballot(flag == st_prefix);
      const uint64_t im = wave::ballot(flag == st_invalid);
      if (lane == 0) {
        s_masks[w] = pm;
        s_masks[max_waves + w] = im;
      }
      __syncthreads();
      p = B;
      for (int ww = 0; ww < nw; ++ww) {
        const uint64_t m = s_masks[ww];
        if (m) {
          p = ww * 64 + __builtin_ctzll(m);
          break;
        }
      }
      bool bad = false;
      for (int ww = 0; ww < nw; ++ww) {
        const int lo = ww * 64;
        uint64_t need;
        if (p >= lo + 63) need = ~0ull;
        else if (p < lo) need = 0;
        else need = (2ull << (p - lo)) - 1;
        bad |= (s_masks[max_waves + ww] & need) != 0;
      }
      __syncthreads();  // the masks are rewritten by the next poll
      if (!bad) break;
      // Every work-item backs off; work-item 0's clock decides for all.
      const bool late = wait.failed_round();
      if (tid == 0) {
        s_give_up = late ? 1 : 0;
        if (late) report_trip(sync);
      }
      __syncthreads();
      give_up = s_give_up != 0;
    }
    // Fold the window [nearest .. p] and broadcast it.
    W x = (tid <= p) ? val : Op::identity();
    x = wave::reduce_full<Op>(x);
    if (lane == 0) s_red[w] = x;
    __syncthreads();
    W window = Op::identity();
    for (int ww = 0; ww < nw; ++ww) window = Op::apply(window, s_red[ww]);
    __syncthreads();  // s_red is rewritten by the next window
    excl = Op::apply(window, excl);
    if (p < B) return excl;
    base -= B;
  }
}

// Launch variants (acpp_launch_config::flags).
enum : uint32_t {
  variant_block_lookback = 1u,  // look back with the whole work-group
};

template <class T>
struct params {
  using W = typename elem_traits<T>::work;
  const T *in;
  T *out;
  uint64_t n;
  W init;
  int has_init;
  int exclusive;
  int vec_io;  // 16-byte vector loads and stores (see u32x4_a4)
  int block_lookback;
  int force_give_up;     // ACPP_VARIANT_FORCE_GIVE_UP (tests of the error path)
  int unary;             // element transform (unary.hpp) of the transform scans
  uint32_t num_tiles;
  uint32_t epoch;        // tag of this launch's status words (>= 1)
  uint32_t ticket_base;  // tile ticket counter value when this launch starts
  sync_header *sync;
  uint64_t *status;
};

// Shared words of a scan work-group.
template <class W, int R>
struct scan_shared {
  uint32_t tile;
  W wave[R * max_waves];  // per (row, wave): totals, then exclusive offsets
  W excl;
  W agg;
  uint64_t masks[2 * max_waves];
  W red[max_waves];
};

// One tile.  FULL tiles (all R x B vectors in range, 16-byte I/O allowed)
// and the partial last tile are separate compile-time paths: with no join
// after the loads, the compiler waits for each row's load only when that
// row is processed (vmcnt(7), (6), ...), so the local and wave scans of
// the first rows overlap the arrival of the later ones.  With one join it
// waited for all R loads first (2^28 int32: 368-377 us split, 373-379 us
// joined, same runs; 2^20: 11.4 against 12.0 us).
// XF: apply the element transform p.unary to every input element (the
// transform scans); false for the plain scans.
// NT: full tiles are read and written with non-temporal loads and stores.
template <bool FULL, class T, class Op, int R, bool XF, bool NT>
__device__ __forceinline__ void scan_tile(const params<T> &p, uint32_t tile,
                                          scan_shared<typename elem_traits<T>::work, R> &sh) {
  using W = typename elem_traits<T>::work;
  using SIO = status_io<W>;
  constexpr int VEC = 16 / sizeof(T);
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int nw = B >> 6;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const uint32_t epoch = p.epoch;
  const uint64_t tile_elems = static_cast<uint64_t>(B) * R * VEC;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;
  // Elements of the tile in range (partial tiles; < tile_elems there), so
  // that bounds checks compare 32-bit offsets inside the tile.
  const uint32_t in_range = FULL ? static_cast<uint32_t>(tile_elems)
                                 : static_cast<uint32_t>(p.n > base ? min(p.n - base, tile_elems) : 0);

  // The R loads are issued back to back, then converted to the work type.
  u32x4 q[R];
  if constexpr (FULL) {
    const u32x4_a4 *src = reinterpret_cast<const u32x4_a4 *>(p.in + base);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if constexpr (NT)
        q[r] = __builtin_nontemporal_load(src + r * B + tid);
      else
        q[r] = src[r * B + tid];
      // With a transform, the loads are otherwise issued in any order ahead
      // of the transform's branch, and the first row waits for nearly all.
      if constexpr (XF) __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      T tmp[VEC];
#pragma unroll
      for (int k = 0; k < VEC; ++k) {
        const uint32_t i = static_cast<uint32_t>(r * B + tid) * VEC + k;
        tmp[k] = (i < in_range) ? p.in[base + i] : T{};
      }
      __builtin_memcpy(&q[r], tmp, 16);
    }
  }
  W v[R][VEC];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    T tmp[VEC];
    __builtin_memcpy(tmp, &q[r], 16);
#pragma unroll
    for (int k = 0; k < VEC; ++k) v[r][k] = elem_traits<T>::to_work(tmp[k]);
  }

  // Per row: the work-item-local inclusive scan of the row vector, then the
  // wave scan of the per-lane row totals.
  W wexcl[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    // One uniform switch per row, so that each row is transformed as its
    // load arrives (a whole tile of widened 8- or 16-bit elements under one
    // switch also spilled to scratch).
    if constexpr (XF) unary_detail::apply_row<T>(p.unary, v[r]);
    if constexpr (!FULL) {
      // Elements past the end act as the identity.
#pragma unroll
      for (int k = 0; k < VEC; ++k) {
        const uint32_t i = static_cast<uint32_t>(r * B + tid) * VEC + k;
        if (i >= in_range) v[r][k] = Op::identity();
      }
    }
#pragma unroll
    for (int k = 1; k < VEC; ++k) v[r][k] = Op::apply(v[r][k - 1], v[r][k]);
    const W incl = wave::inclusive_scan<Op>(v[r][VEC - 1]);
    wexcl[r] = wave::shift_up<Op>(incl);
    if (lane == 63) sh.wave[r * nw + w] = incl;
  }
  __syncthreads();
  // Wave 0 turns the (row, wave) totals into offsets and publishes the
  // tile's aggregate at once; with the one-wave look-back it then looks
  // back, while the other waves wait at the barrier.
  if (w == 0) {
    const W agg = tile_prefixes<Op, W>(R * nw, sh.wave);
    if (tile == 0) {
      const W first = p.has_init ? p.init : Op::identity();
      if (lane == 0) {
        if (p.num_tiles > 1) SIO::publish(p.status, 0, epoch, st_prefix, Op::apply(first, agg));
        sh.excl = first;
      }
    } else {
      if (lane == 0) SIO::publish(p.status, tile, epoch, st_aggregate, agg);
      if (!p.block_lookback) {
        wait_own_stores();  // the aggregate is out before the first poll
        const W excl = lookback<Op, W, lookback_slots>(p.status, tile, epoch, p.sync, p.force_give_up != 0);
        if (lane == 0) {
          if (tile + 1 < p.num_tiles)
            SIO::publish(p.status, tile, epoch, st_prefix, Op::apply(excl, agg));
          sh.excl = excl;
        }
      }
    }
    if (lane == 0) sh.agg = agg;
  }
  __syncthreads();
  // Offsets of (row, wave) inside the tile, in element order.
  W before[R];
#pragma unroll
  for (int r = 0; r < R; ++r) before[r] = sh.wave[r * nw + w];

  W tile_prefix;
  if (tile != 0 && p.block_lookback) {
    const W agg = sh.agg;
    tile_prefix = block_lookback<Op, W>(p.status, tile, epoch, sh.masks, sh.red, p.sync,
                                       p.force_give_up != 0);
    if (tid == 0 && tile + 1 < p.num_tiles)
      SIO::publish(p.status, tile, epoch, st_prefix, Op::apply(tile_prefix, agg));
  } else {
    tile_prefix = sh.excl;
  }

#pragma unroll
  for (int r = 0; r < R; ++r) {
    const W P = Op::apply(tile_prefix, Op::apply(before[r], wexcl[r]));
    T o[VEC];
    if (p.exclusive) {
      o[0] = static_cast<T>(P);
#pragma unroll
      for (int k = 1; k < VEC; ++k) o[k] = static_cast<T>(Op::apply(P, v[r][k - 1]));
      // d[0] is init itself (logical operators would turn op(init, identity)
      // into 0/1).
      if (r == 0 && tile == 0 && tid == 0) o[0] = static_cast<T>(p.init);
    } else {
#pragma unroll
      for (int k = 0; k < VEC; ++k) o[k] = static_cast<T>(Op::apply(P, v[r][k]));
      // Without init the first element is passed through untouched
      // (std::inclusive_scan semantics).  op(identity, x) == x holds for
      // every operator except the logical ones, which turn x into 0/1.
      if (r == 0 && tile == 0 && tid == 0 && !p.has_init) o[0] = static_cast<T>(v[0][0]);
    }
    if constexpr (FULL) {
      u32x4 qo;
      __builtin_memcpy(&qo, o, 16);
      // Ordinary stores by default: non-temporal output stores alone
      // measured no faster (380 us against 377 us at 2^28 int32).  The NT
      // variant streams both directions, as the fastest tile copy does.
      if constexpr (NT)
        __builtin_nontemporal_store(qo, reinterpret_cast<u32x4_a4 *>(p.out + base) + r * B + tid);
      else
        reinterpret_cast<u32x4_a4 *>(p.out + base)[r * B + tid] = qo;
    } else {
#pragma unroll
      for (int k = 0; k < VEC; ++k) {
        const uint32_t i = static_cast<uint32_t>(r * B + tid) * VEC + k;
        if (i < in_range) p.out[base + i] = o[k];
      }
    }
  }
}

template <class T, class Op, int R, bool XF = false, bool NT = false>
__global__ __launch_bounds__(1024) void scan_kernel(params<T> p) {
  using W = typename elem_traits<T>::work;
  constexpr int VEC = 16 / sizeof(T);
  __shared__ scan_shared<W, R> sh;

  // Dynamic tile ids: tiles are handed out in the order work-groups start,
  // so every predecessor of a tile is already running.  The counter is never
  // reset; the host passes the value it had before this launch.
  if (threadIdx.x == 0)
    sh.tile = __hip_atomic_fetch_add(&p.sync->tile_counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
              p.ticket_base;
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(sh.tile);
  // Tickets beyond this launch's tiles only appear if the host-side ticket
  // count went out of step with the counter; never touch memory for them.
  if (tile >= p.num_tiles) return;
  const uint64_t tile_elems = static_cast<uint64_t>(blockDim.x) * R * VEC;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;
  if (p.vec_io && base + tile_elems <= p.n)
    scan_tile<true, T, Op, R, XF, NT>(p, tile, sh);
  else
    scan_tile<false, T, Op, R, XF, NT>(p, tile, sh);
}

// --------------------------------------------------------------------------
// 8-bit scans on packed 16-bit pairs (plus, multiply, minimum, maximum and
// the bitwise operators).
//
// The generic kernel widens every 8-bit element to a 32-bit work value, so
// a 16-element vector takes 16 registers and only 4-row tiles fit.  Here
// register j of a row holds element j in its low and element j + 8 in its
// high 16 bits (one v_perm_b32 each), zero-extended, and the row's
// inclusive scan runs on both halves at once with packed 16-bit operations
// (v_pk_add_u16, v_pk_mul_lo_u16, v_pk_min_u16, ...), the high half then
// combined with the low half's total.  A row takes 8 registers, so 8-row
// (128 KiB) tiles fit, with about half the instructions per element.
// Every result's low byte is the 8-bit result: sums and products are exact
// modulo 2^16 (then modulo 2^32 across lanes and tiles), hence modulo 256;
// the bitwise operators never leave the low byte; minimum and maximum of
// signed bytes run on the bytes with their top bit flipped (x ^ 0x80),
// which orders them as unsigned values, and the output is flipped back.
// 2^28 int8 plus: 109-111 us against 125-133 us for the generic 4-row
// kernel (MI355X, same runs).
// Transform scans (XF) apply the element transform to the raw bytes of a
// row, four bytes per 32-bit operation (unary_detail::apply_bytes), before
// the flip; a partial tile transforms its loaded bytes and only then fills
// the bytes past the end with the identity.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u32(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }

constexpr int packed8_rows = 8;

// Whether the packed kernel covers operator Op on 8-bit element type T.
template <class T, class Op>
constexpr bool packed8_covers() {
  return sizeof(T) == 1 && Op::id >= ops::op_plus && Op::id <= ops::op_bit_xor;
}

// The operator on packed pairs and on the 32-bit row totals; `ident` is
// the identity byte (also what elements past the end read as), in the
// flipped domain for signed minimum / maximum.
template <class T, class Op>
struct packed8_op {
  static constexpr int id = Op::id;
  static constexpr bool flip = (id == ops::op_min || id == ops::op_max) && std::is_signed_v<T>;
  using Op32 = std::conditional_t<
      id == ops::op_plus, ops::plus<uint32_t, uint32_t>,
      std::conditional_t<
          id == ops::op_multiply, ops::multiply<uint32_t, uint32_t>,
          std::conditional_t<
              id == ops::op_min, ops::minimum<uint32_t, uint32_t>,
              std::conditional_t<id == ops::op_max, ops::maximum<uint32_t, uint32_t>,
                                 std::conditional_t<id == ops::op_bit_and, ops::bit_and<uint32_t, uint32_t>,
                                                    std::conditional_t<id == ops::op_bit_or,
                                                                       ops::bit_or<uint32_t, uint32_t>,
                                                                       ops::bit_xor<uint32_t, uint32_t>>>>>>>;
  static constexpr uint32_t ident = (id == ops::op_multiply) ? 1u
                                    : (id == ops::op_min || id == ops::op_bit_and) ? 0xffu
                                                                                     : 0u;
  __device__ __forceinline__ static u16x2 apply(u16x2 a, u16x2 b) {
    if constexpr (id == ops::op_plus) return a + b;
    else if constexpr (id == ops::op_multiply) return a * b;
    else if constexpr (id == ops::op_min) return __builtin_elementwise_min(a, b);
    else if constexpr (id == ops::op_max) return __builtin_elementwise_max(a, b);
    else if constexpr (id == ops::op_bit_and) return a & b;
    else if constexpr (id == ops::op_bit_or) return a | b;
    else return a ^ b;
  }
};

template <bool FULL, class T, class Op, int R, bool XF>
__device__ __forceinline__ void scan8_tile(const params<T> &p, uint32_t tile, uint32_t *s_wave, uint32_t &s_excl) {
  static_assert(packed8_covers<T, Op>(), "packed 8-bit scan");
  using PO = packed8_op<T, Op>;
  using Op32 = typename PO::Op32;
  using SIO = status_io<uint32_t>;
  constexpr uint32_t flip_mask = PO::flip ? 0x80808080u : 0u;
  // elements past the end: the identity, stored before the flip
  constexpr uint8_t fill = static_cast<uint8_t>(PO::ident ^ (PO::flip ? 0x80u : 0u));
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int nw = B >> 6;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const uint32_t epoch = p.epoch;
  const uint64_t tile_elems = static_cast<uint64_t>(B) * R * 16;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;
  const uint32_t in_range = FULL ? static_cast<uint32_t>(tile_elems)
                                 : static_cast<uint32_t>(p.n > base ? min(p.n - base, tile_elems) : 0);
  const uint8_t *in8 = reinterpret_cast<const uint8_t *>(p.in) + base;
  uint8_t *out8 = reinterpret_cast<uint8_t *>(p.out) + base;

  u32x4 q[R];
  if constexpr (FULL) {
    const u32x4 *src = reinterpret_cast<const u32x4 *>(in8);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      q[r] = src[r * B + tid];
      // With a transform, the loads are otherwise issued in any order ahead
      // of the transform's branch, and the first row waits for nearly all.
      if constexpr (XF) __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      constexpr uint32_t fill4 = 0x01010101u * fill;
      if constexpr (XF) {
        // transform the loaded bytes, then fill the bytes past the end
        uint32_t d[4] = {0u, 0u, 0u, 0u};
        const uint32_t first = static_cast<uint32_t>(r * B + tid) * 16;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (first + k < in_range) d[k >> 2] |= static_cast<uint32_t>(in8[first + k]) << (8 * (k & 3));
        }
        unary_detail::apply_bytes<T>(p.unary, d);
        const uint32_t cnt = in_range > first ? min(in_range - first, 16u) : 0u;
#pragma unroll
        for (int k4 = 0; k4 < 4; ++k4) {
          const uint32_t vb = cnt > 4u * k4 ? min(cnt - 4u * k4, 4u) : 0u;  // loaded bytes in word k4
          const uint32_t m = vb >= 4u ? ~0u : ((1u << (8u * vb)) - 1u);
          d[k4] = (d[k4] & m) | (fill4 & ~m);
        }
        q[r] = u32x4{d[0], d[1], d[2], d[3]};
      } else {
        uint32_t d[4] = {fill4, fill4, fill4, fill4};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const uint32_t i = static_cast<uint32_t>(r * B + tid) * 16 + k;
          if (i < in_range)
            d[k >> 2] = (d[k >> 2] & ~(0xffu << (8 * (k & 3)))) | (static_cast<uint32_t>(in8[i]) << (8 * (k & 3)));
        }
        q[r] = u32x4{d[0], d[1], d[2], d[3]};
      }
    }
  }
  u16x2 P[R][8];
  uint32_t wexcl[R];
  const u16x2 id2 = {static_cast<unsigned short>(PO::ident), static_cast<unsigned short>(PO::ident)};
#pragma unroll
  for (int r = 0; r < R; ++r) {
    uint32_t d[4] = {q[r].x, q[r].y, q[r].z, q[r].w};
    if constexpr (XF && FULL) unary_detail::apply_bytes<T>(p.unary, d);
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] ^= flip_mask;
    // pair j = (byte j, byte j + 8), zero-extended to 16 bits each
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t sel = (j & 3) | (12u << 8) | ((4u + (j & 3)) << 16) | (12u << 24);
      P[r][j] = as_u16x2(__builtin_amdgcn_perm(d[(j >> 2) + 2], d[j >> 2], sel));
    }
#pragma unroll
    for (int j = 1; j < 8; ++j) P[r][j] = PO::apply(P[r][j - 1], P[r][j]);
    // the high half continues after element 7
    const u16x2 fix = {id2.x, P[r][7].x};
#pragma unroll
    for (int j = 0; j < 8; ++j) P[r][j] = PO::apply(fix, P[r][j]);
    const uint32_t incl = wave::inclusive_scan<Op32>(static_cast<uint32_t>(P[r][7].y));
    wexcl[r] = wave::shift_up<Op32>(incl);
    if (lane == 63) s_wave[r * nw + w] = incl;
  }
  __syncthreads();
  if (w == 0) {
    const uint32_t agg = tile_prefixes<Op32, uint32_t>(R * nw, s_wave);
    if (tile == 0) {
      const uint32_t first =
          p.has_init ? ((static_cast<uint32_t>(p.init) & 0xffu) ^ (flip_mask & 0xffu)) : Op32::identity();
      if (lane == 0) {
        if (p.num_tiles > 1) SIO::publish(p.status, 0, epoch, st_prefix, Op32::apply(first, agg));
        s_excl = first;
      }
    } else {
      if (lane == 0) SIO::publish(p.status, tile, epoch, st_aggregate, agg);
      wait_own_stores();  // the aggregate is out before the first poll
      const uint32_t excl = lookback<Op32, uint32_t, lookback_slots>(p.status, tile, epoch, p.sync,
                                                                     p.force_give_up != 0);
      if (lane == 0) {
        if (tile + 1 < p.num_tiles) SIO::publish(p.status, tile, epoch, st_prefix, Op32::apply(excl, agg));
        s_excl = excl;
      }
    }
  }
  __syncthreads();
  const uint32_t tile_prefix = s_excl;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t P32 = Op32::apply(tile_prefix, Op32::apply(s_wave[r * nw + w], wexcl[r]));
    const unsigned short p16 = static_cast<unsigned short>(P32);
    const u16x2 pv = {p16, p16};
    u16x2 O[8];
    if (p.exclusive) {
      // element j takes the inclusive value of element j - 1; element 8
      // (high half of pair 0) that of element 7
      const u16x2 e0 = {id2.x, P[r][7].x};
      O[0] = PO::apply(pv, e0);
#pragma unroll
      for (int j = 1; j < 8; ++j) O[j] = PO::apply(pv, P[r][j - 1]);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) O[j] = PO::apply(pv, P[r][j]);
    }
    // low bytes back in element order: [e0 e1 e8 e9] and [e2 e3 e10 e11]
    // give dwords 0 and 2; likewise pairs 4..7 give dwords 1 and 3
    const uint32_t a0 = __builtin_amdgcn_perm(as_u32(O[1]), as_u32(O[0]), 0x06020400u);
    const uint32_t a1 = __builtin_amdgcn_perm(as_u32(O[3]), as_u32(O[2]), 0x06020400u);
    const uint32_t a2 = __builtin_amdgcn_perm(as_u32(O[5]), as_u32(O[4]), 0x06020400u);
    const uint32_t a3 = __builtin_amdgcn_perm(as_u32(O[7]), as_u32(O[6]), 0x06020400u);
    const uint32_t o[4] = {__builtin_amdgcn_perm(a1, a0, 0x05040100u) ^ flip_mask,
                           __builtin_amdgcn_perm(a3, a2, 0x05040100u) ^ flip_mask,
                           __builtin_amdgcn_perm(a1, a0, 0x07060302u) ^ flip_mask,
                           __builtin_amdgcn_perm(a3, a2, 0x07060302u) ^ flip_mask};
    if constexpr (FULL) {
      reinterpret_cast<u32x4 *>(out8)[r * B + tid] = u32x4{o[0], o[1], o[2], o[3]};
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint32_t i = static_cast<uint32_t>(r * B + tid) * 16 + k;
        if (i < in_range) out8[i] = static_cast<uint8_t>(o[k >> 2] >> (8 * (k & 3)));
      }
    }
  }
}

template <class T, class Op, bool XF>
__global__ __launch_bounds__(1024) void scan8_kernel(params<T> p) {
  constexpr int R = packed8_rows;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wave[R * max_waves];
  __shared__ uint32_t s_excl;
  if (threadIdx.x == 0)
    s_tile = __hip_atomic_fetch_add(&p.sync->tile_counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
             p.ticket_base;
  __syncthreads();
  const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
  if (tile >= p.num_tiles) return;
  const uint64_t tile_elems = static_cast<uint64_t>(blockDim.x) * R * 16;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;
  if (p.vec_io && base + tile_elems <= p.n)
    scan8_tile<true, T, Op, R, XF>(p, tile, s_wave, s_excl);
  else
    scan8_tile<false, T, Op, R, XF>(p, tile, s_wave, s_excl);
}

}  // namespace scan_detail
}  // namespace acpp_impl
