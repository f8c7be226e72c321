// Segmented scans and reductions: thousands of independent scans /
// reductions over [offsets[s], offsets[s+1]) in one launch.
//
// The reference has no segmented form; the nearest equivalents are one
// joint_* group algorithm call per segment
// (include/hipSYCL/sycl/libkernel/generic/hiplike/group_functions.hpp:214-287,
// 361-432) or one algorithms::inclusive_scan / reduce call per segment
// (include/hipSYCL/algorithms/numeric.hpp:226-279), each a separate launch.
//
// MI355X design:
//   - the concatenated elements are cut into equal tiles, independent of the
//     segment lengths, so log-normal segment lengths cost no load imbalance;
//   - a setup kernel builds the device work queue: for every tile the first
//     segment it touches (each non-empty segment writes the tiles whose first
//     element it contains), and writes init into empty segments' results;
//   - persistent work-groups (one wave of resident groups) draw tiles from
//     a ticket counter in order; every tile marks its segment starts in an
//     LDS bitmap and runs a segmented scan over (head flag, value) pairs:
//     (f1,v1) . (f2,v2) = (f1|f2, f2 ? v2 : v1 op v2), in registers, across
//     each wave with DPP and across waves through LDS;
//   - tiles chain through a decoupled look-back on the pair operator.  The
//     operator is not commutative, so the window is folded in tile order,
//     and a look-back stops at the nearest predecessor that contains a
//     segment start (its aggregate already discards everything before it),
//     so it is short whenever segments are shorter than a tile;
//   - scans write every element.  Reductions skip the look-back: each tile
//     writes the totals of the segments it holds whole and two partial
//     folds, and a fix-up kernel joins the partials of segments that span
//     tiles (segmented_reduce_kernel below);
//   - the setup and main kernels can be captured into a hipGraph once and
//     replayed (acpp_segmented_plan_*).
#include <memory>

#include "launch_util.hpp"

namespace acpp_impl {
namespace seg_detail {

constexpr int max_waves = 16;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// 16-byte vector accesses need dword alignment only; every segmented element
// type is 4 or 8 bytes, so any element-aligned view takes the vector path.
typedef unsigned int u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

enum : uint32_t { st_invalid = 0, st_aggregate = 1, st_prefix = 2 };

// Status tags: status_tag() in internal.hpp (epoch << 3 | head flag << 2 |
// status), shared with the scans.

__device__ __forceinline__ uint64_t ld64(const uint64_t *p) {
  return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class W, int Bytes = sizeof(W)>
struct seg_status;

// Each tile's status has a 128-byte line of its own, so that the polls of
// all tiles in flight do not pile onto a few lines.
constexpr uint64_t seg_stride_words = 16;

template <class W>
struct seg_status<W, 4> {
  static constexpr int words = 1;
  __device__ static void publish(uint64_t *st, uint32_t tile, uint32_t epoch, uint32_t status,
                                 uint32_t f, W v) {
    const uint64_t tag = status_tag(epoch, f, status);
    st64(st + tile * seg_stride_words, (tag << 32) | __builtin_bit_cast(uint32_t, v));
  }
  __device__ static uint32_t read(const uint64_t *st, uint32_t tile, uint32_t epoch, uint32_t &f,
                                  W &v) {
    const uint64_t w = ld64(st + tile * seg_stride_words);
    const uint32_t tag = static_cast<uint32_t>(w >> 32);
    if (tag_epoch(tag) != epoch) return st_invalid;
    f = tag_head(tag);
    v = __builtin_bit_cast(W, static_cast<uint32_t>(w));
    return tag_status(tag);
  }
};

template <class W>
struct seg_status<W, 8> {
  static constexpr int words = 2;
  __device__ static void publish(uint64_t *st, uint32_t tile, uint32_t epoch, uint32_t status,
                                 uint32_t f, W v) {
    const uint64_t tag = status_tag(epoch, f, status);
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    uint64_t *slot = st + tile * seg_stride_words;
    st64(slot, (tag << 32) | (bits & 0xffffffffu));
    st64(slot + 1, (tag << 32) | (bits >> 32));
  }
  __device__ static uint32_t read(const uint64_t *st, uint32_t tile, uint32_t epoch, uint32_t &f,
                                  W &v) {
    const uint64_t *slot = st + tile * seg_stride_words;
    const uint64_t w0 = ld64(slot);
    const uint64_t w1 = ld64(slot + 1);
    const uint32_t t0 = static_cast<uint32_t>(w0 >> 32);
    const uint32_t t1 = static_cast<uint32_t>(w1 >> 32);
    if (t0 != t1 || tag_epoch(t0) != epoch) return st_invalid;
    f = tag_head(t0);
    v = __builtin_bit_cast(W, (w1 << 32) | (w0 & 0xffffffffu));
    return tag_status(t0);
  }
};

// (f, v) pair operator, earlier . later
template <class Op, class W>
__device__ __forceinline__ void combine(uint32_t fa, W va, uint32_t &fb, W &vb) {
  vb = fb ? vb : Op::apply(va, vb);
  fb |= fa;
}

// Inclusive scan of pairs across the wave (lane order = element order).
template <class Op, class W>
__device__ __forceinline__ void wave_pair_scan(uint32_t &f, W &v) {
  const W id = Op::identity();
#define ACPP_SEG_STEP(CTRL, RM)                                  \
  {                                                              \
    const uint32_t fo = wave::dpp<CTRL, RM, 0xf>(0u, f);         \
    const W vo = wave::dpp<CTRL, RM, 0xf>(id, v);                \
    combine<Op>(fo, vo, f, v);                                   \
  }
  ACPP_SEG_STEP(wave::row_shr1, 0xf)
  ACPP_SEG_STEP(wave::row_shr2, 0xf)
  ACPP_SEG_STEP(wave::row_shr4, 0xf)
  ACPP_SEG_STEP(wave::row_shr8, 0xf)
  ACPP_SEG_STEP(wave::row_bcast15, 0xa)
  ACPP_SEG_STEP(wave::row_bcast31, 0xc)
#undef ACPP_SEG_STEP
}

// Exclusive prefixes, in element order, of a tile's per-(row, wave) pair
// totals.  s_tf / s_tv[r * nw + w] hold the totals on entry and the
// exclusive prefixes on return (P = R * nw <= 128 entries).  Called by one
// whole wave, which scans them with DPP (two entries per lane above 64);
// returns the tile aggregate, uniform across the wave.  Replaces a serial
// walk over all R * nw totals by every work-item.
template <class Op, class W>
__device__ void tile_pair_prefixes(int P, uint32_t *s_tf, W *s_tv, uint32_t &agg_f, W &agg_v) {
  const int lane = threadIdx.x & 63;
  const W id = Op::identity();
  const int per = P > 64 ? 2 : 1;
  const int a = lane * per;
  uint32_t f0 = 0, f1 = 0;
  W v0 = id, v1 = id;
  if (a < P) {
    f0 = s_tf[a];
    v0 = s_tv[a];
  }
  if (per == 2 && a + 1 < P) {
    f1 = s_tf[a + 1];
    v1 = s_tv[a + 1];
  }
  uint32_t sf = f1;  // lane total: entry a . entry a + 1
  W sv = v1;
  combine<Op>(f0, v0, sf, sv);
  wave_pair_scan<Op>(sf, sv);
  const uint32_t xf = wave::dpp<wave::wave_shr1, 0xf, 0xf>(0u, sf);
  const W xv = wave::dpp<wave::wave_shr1, 0xf, 0xf>(id, sv);
  agg_f = wave::read_lane(sf, 63);
  agg_v = wave::read_lane(sv, 63);
  if (a < P) {
    s_tf[a] = xf;
    s_tv[a] = xv;
  }
  if (per == 2 && a + 1 < P) {
    uint32_t bf = f0;
    W bv = v0;
    combine<Op>(xf, xv, bf, bv);
    s_tf[a + 1] = bf;
    s_tv[a + 1] = bv;
  }
}

// Ordered look-back by one wave: lane l polls tile (wbase - 63 + l), lane 63
// the nearest predecessor.  It stops at the nearest tile that is a prefix
// or an aggregate containing a segment start, and folds the window from
// there in tile order.  Must be called by a whole wave; result uniform.
template <class Op, class W>
__device__ void wave_pair_lookback(const uint64_t *status, uint32_t tile, uint32_t epoch,
                                   sync_header *sync, bool force_give_up, uint32_t &out_f, W &out_v) {
  using SIO = seg_status<W>;
  const int lane = threadIdx.x & 63;
  const W id = Op::identity();
  int64_t wbase = static_cast<int64_t>(tile) - 1;
  uint32_t ef = 0;  // accumulated prefix of tiles nearer than the window
  W ev = id;
  lookback_wait wait;
  wait.begin();
  bool give_up = force_give_up;  // uniform
  if (give_up && lane == 0) report_trip(sync);
  for (;;) {
    const int64_t j = wbase - (63 - lane);
    uint32_t lf = 0;
    W lv = id;
    int q;
    for (;;) {
      uint32_t st = st_prefix;  // positions before tile 0 act as an empty prefix
      lf = 0;
      lv = id;
      if (j >= 0) st = SIO::read(status, static_cast<uint32_t>(j), epoch, lf, lv);
      if (give_up) st = st_prefix;
      const bool term = (st == st_prefix) || (st == st_aggregate && lf);
      const uint64_t tm = wave::ballot(term);
      q = tm ? 63 - __builtin_clzll(tm) : -1;
      const uint64_t need = (q <= 0) ? ~0ull : (~0ull << q);  // lanes q..63
      const uint64_t bad = wave::ballot(st == st_invalid) & need;
      if (bad == 0) break;
      if (wait.failed_round()) {
        give_up = true;
        if (lane == 0) report_trip(sync);
      }
    }
    uint32_t xf = (lane >= q) ? lf : 0u;
    W xv = (lane >= q) ? lv : id;
    wave_pair_scan<Op>(xf, xv);
    uint32_t wf = wave::read_lane(xf, 63);
    W wv = wave::read_lane(xv, 63);
    // window (earlier) . accumulated (later)
    combine<Op>(wf, wv, ef, ev);
    if (q >= 0) break;
    wbase -= 64;
  }
  out_f = ef;
  out_v = ev;
}

template <class T>
struct params {
  using W = typename elem_traits<T>::work;
  const T *in;
  T *out;
  const uint64_t *offsets;  // num_segments + 1 entries
  const uint32_t *tile_seg;  // first segment of every tile (num_tiles + 1 entries)
  const uint32_t *heads;     // bit e set: a non-empty segment starts at element e
  uint64_t head_words_total;  // 32-bit words of the bitmap
  int block_lookback;        // 1: work-group-wide look-back window
  int force_give_up;         // ACPP_VARIANT_FORCE_GIVE_UP (tests of the error path)
  uint64_t num_segments;
  uint64_t total;
  W init;
  int reduce;  // 0: inclusive scan per segment, 1: reduce per segment
  int vec_io;
  uint32_t num_tiles;
  uint32_t epoch;
  uint32_t ticket_base;
  sync_header *sync;
  uint64_t *status;
};

__device__ __forceinline__ uint64_t clamp_off(const uint64_t *offsets, uint64_t s, uint64_t total) {
  const uint64_t o = offsets[s];
  return o < total ? o : total;
}

// Setup: tile_seg[t] = the non-empty segment containing element t*tile_elems
// (num_segments for t == num_tiles), the head bitmap for the scans (bit b
// for the first element b of every non-empty segment; the bitmap is cleared
// beforehand; none for the reductions, heads_out NULL), and init into empty
// segments' results.
template <class T>
__global__ void setup_kernel(params<T> p, uint32_t *tile_seg_out, uint32_t *heads_out, uint64_t tile_elems) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < p.num_segments;
       s += stride) {
    const uint64_t b = clamp_off(p.offsets, s, p.total);
    const uint64_t e = clamp_off(p.offsets, s + 1, p.total);
    if (e <= b) {
      if (p.reduce) p.out[s] = static_cast<T>(p.init);
      continue;
    }
    if (heads_out) atomicOr(&heads_out[b >> 5], 1u << (b & 31));
    // tiles whose first element lies in [b, e)
    const uint64_t t0 = (b + tile_elems - 1) / tile_elems;
    const uint64_t t1 = (e + tile_elems - 1) / tile_elems;
    for (uint64_t t = t0; t < t1; ++t) tile_seg_out[t] = static_cast<uint32_t>(s);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tile_seg_out[p.num_tiles] = static_cast<uint32_t>(p.num_segments);
}

// Largest s in [lo, hi] with offsets[s] <= e (offsets[lo] <= e holds).
__device__ __forceinline__ uint64_t find_segment(const uint64_t *offsets, uint64_t lo, uint64_t hi,
                                                 uint64_t e, uint64_t total) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) / 2;
    if (clamp_off(offsets, mid, total) <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// threadIdx.x through an empty asm statement: a value the compiler cannot
// move, so what is derived from it inside a loop stays inside the loop.
__device__ __forceinline__ int loop_local_thread_index() {
  int t = static_cast<int>(threadIdx.x);
  asm volatile("" : "+v"(t));
  return t;
}

// Head bits of a work-item's R row vectors from the tile's LDS bitmap, all in
// one word: bit r * VEC + k = element k of row r starts a segment.  VEC
// consecutive bits never straddle a bitmap word (VEC divides 32).
template <int R, int VEC>
__device__ __forceinline__ uint32_t row_head_bits(const uint32_t *s_heads, int tid, int B) {
  static_assert(R * VEC <= 32, "head bits of a work-item must fit one word");
  uint32_t fm = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t pos = static_cast<uint32_t>(r * B + tid) * VEC;
    fm |= ((s_heads[pos >> 5] >> (pos & 31)) & ((1u << VEC) - 1u)) << (r * VEC);
  }
  return fm;
}

// Wave scans of the (head flag, total) pairs of each row, after the
// work-item-local scans.  Lane 63 leaves the wave's row total in
// s_tf / s_tv[r * nw + w]; every work-item folds its wave-exclusive prefix
// into the row's elements up to the row's first segment start and returns
// bit r set when a segment starts in an earlier lane of row r (the elements
// then hold their full in-segment fold already).  Folding at once keeps one
// bit per row live instead of a (flag, value) pair.
template <class Op, int R, int VEC, class W>
__device__ __forceinline__ uint32_t fold_wave_prefixes(uint32_t fm, W (&v)[R][VEC], uint32_t *s_tf, W *s_tv,
                                                       int nw, int w, int lane) {
  const W id = Op::identity();
  uint32_t xcut = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t rowf = (fm >> (r * VEC)) & ((1u << VEC) - 1u);
    uint32_t tf = rowf != 0u;
    W tv = v[r][VEC - 1];
    wave_pair_scan<Op>(tf, tv);
    if (lane == 63) {
      s_tf[r * nw + w] = tf;
      s_tv[r * nw + w] = tv;
    }
    const uint32_t ef = wave::dpp<wave::wave_shr1, 0xf, 0xf>(0u, tf);
    const W ev = wave::dpp<wave::wave_shr1, 0xf, 0xf>(id, tv);
    xcut |= ef << r;
#pragma unroll
    for (int k = 0; k < VEC; ++k)
      if ((rowf & ((2u << k) - 1u)) == 0u) v[r][k] = Op::apply(ev, v[r][k]);
  }
  return xcut;
}

template <class T, class Op, int R>
__global__ __launch_bounds__(1024) void segmented_kernel(params<T> p) {
  using W = typename elem_traits<T>::work;
  using SIO = seg_status<W>;
  constexpr int VEC = 16 / sizeof(T);
  const int B = blockDim.x;
  const int nw = B >> 6;
  const uint64_t tile_elems = static_cast<uint64_t>(B) * R * VEC;
  const W id = Op::identity();

  extern __shared__ uint32_t s_heads[];  // tile_elems + 1 bits, in 2 x blockDim.x words
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_tf[R * max_waves];  // per (row, wave): totals, then exclusive prefixes
  __shared__ W s_tv[R * max_waves];
  __shared__ uint32_t s_aggf;
  __shared__ W s_aggv;
  __shared__ uint32_t s_lbf[max_waves];
  __shared__ W s_lbv[max_waves];
  __shared__ int s_q[max_waves];
  __shared__ int s_give_up;
  __shared__ uint64_t s_invalid[max_waves];
  const uint32_t head_words = static_cast<uint32_t>((tile_elems + 1 + 31) / 32);

  // A ticket is drawn only when the work-group is ready to work on the
  // tile: every later tile's look-back waits for this tile's aggregate, so a
  // ticket held while the group is busy with another tile stalls the whole
  // chain (drawing one tile ahead made the 65,536-segment int64 scan ~3x
  // slower on MI355X).
  for (;;) {
    // The work-item's indices are re-derived in every iteration, so that the
    // per-row addresses and bit positions computed from them are too: hoisted
    // out of the loop they held ~60 VGPRs for the whole kernel and made
    // 8-row tiles spill.
    const int tid = loop_local_thread_index();
    const int w = tid >> 6;
    const int lane = tid & 63;
    __syncthreads();  // LDS of the previous tile is free again
    if (tid == 0)
      s_tile = __hip_atomic_fetch_add(&p.sync->tile_counter, 1u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) -
               p.ticket_base;
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= p.num_tiles) return;  // queue drained

    const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;

    // Full tiles and the partial last tile as separate compile-time paths:
    // with no join after the loads, each row waits only for its own load.
    const bool full = p.vec_io && (base + tile_elems <= p.total);
    auto tile_body = [&](auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
      // elements of the tile in range, as a 32-bit offset inside the tile
      const uint32_t in_range =
          FULL ? static_cast<uint32_t>(tile_elems) : static_cast<uint32_t>(min(p.total - base, tile_elems));
      // Head bits of elements [base, base + tile_elems] (base is a multiple
      // of 32), loaded just before the tile's data in the same block, so that
      // waiting for them does not wait for the data (loads retire in order).
      // Both loads are unconditional (addresses clamped into the bitmap) and
      // the words are masked only where they go to LDS, with opaque masks: a
      // load under a branch, or a select the compiler turned back into one,
      // made it wait for every outstanding load, the tile's data included,
      // before the head bits could go to LDS.
      uint32_t hword0, hword1, hmask0, hmask1;
      {
        const uint64_t last = p.head_words_total - 1;
        const uint64_t i0 = (base >> 5) + tid;
        const uint64_t i1 = i0 + B;
        hword0 = p.heads[i0 < last ? i0 : last];
        hword1 = p.heads[i1 < last ? i1 : last];
        hmask0 = (static_cast<uint32_t>(tid) < head_words && i0 <= last) ? ~0u : 0u;
        hmask1 = (static_cast<uint32_t>(tid + B) < head_words && i1 <= last) ? ~0u : 0u;
        asm volatile("" : "+v"(hmask0), "+v"(hmask1));
      }
      __builtin_amdgcn_sched_barrier(0);
      W v[R][VEC];
      u32x4 q[R];
      if constexpr (FULL) {
        const u32x4_a4 *src = reinterpret_cast<const u32x4_a4 *>(p.in + base);
#pragma unroll
        for (int r = 0; r < R; ++r) q[r] = src[r * B + tid];
        // Converted at once: the copies only rename registers (4- and 8-byte
        // elements), and the loaded vectors then never live beside v.
#pragma unroll
        for (int r = 0; r < R; ++r) {
          T tmp[VEC];
          __builtin_memcpy(tmp, &q[r], 16);
#pragma unroll
          for (int k = 0; k < VEC; ++k) v[r][k] = elem_traits<T>::to_work(tmp[k]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int k = 0; k < VEC; ++k) {
            const uint32_t li = static_cast<uint32_t>(r * B + tid) * VEC + k;
            v[r][k] = (li < in_range) ? elem_traits<T>::to_work(p.in[base + li]) : id;
          }
      }
      // Unconditional stores (2 B words of LDS; the words past head_words
      // are zero): a store under a branch made the compiler wait for the
      // tile's data as well.
      s_heads[tid] = hword0 & hmask0;
      s_heads[tid + B] = hword1 & hmask1;
      __syncthreads();

      auto head_at = [&](uint64_t pos) -> uint32_t {
        return (s_heads[pos >> 5] >> (pos & 31)) & 1u;
      };

      // Head bits of the work-item's elements, all rows in one word: bit
      // r * VEC + k = element k of row r starts a segment.
      const uint32_t fm = row_head_bits<R, VEC>(s_heads, tid, B);

      // Work-item-local segmented scan of each row vector: an element that
      // starts a segment keeps its value, the others add the running value.
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 1; k < VEC; ++k)
          v[r][k] = ((fm >> (r * VEC + k)) & 1u) ? v[r][k] : Op::apply(v[r][k - 1], v[r][k]);

      // Wave scans of the row totals; the wave-exclusive prefix is folded into
      // the row's elements at once (up to the row's first segment start), so
      // only one bit per row stays live: bit r of xcut = a segment starts in
      // an earlier lane of this wave's row r.
      const uint32_t xcut = fold_wave_prefixes<Op, R, VEC>(fm, v, s_tf, s_tv, nw, w, lane);
      __syncthreads();
      // Wave 0 turns the (row, wave) totals into exclusive prefixes and
      // publishes the tile's aggregate at once.  A tile that contains a
      // segment start publishes it as its prefix: it is final for everything
      // after the tile.
      if (w == 0) {
        uint32_t af;
        W av;
        tile_pair_prefixes<Op, W>(R * nw, s_tf, s_tv, af, av);
        if (lane == 0) {
          s_aggf = af;
          s_aggv = av;
          if (tile == 0) {
            if (p.num_tiles > 1) SIO::publish(p.status, 0, p.epoch, st_prefix, af, av);
          } else {
            SIO::publish(p.status, tile, p.epoch, af ? st_prefix : st_aggregate, af, av);
          }
        }
      }
      __syncthreads();
      // s_tf / s_tv now hold the pairs before each (row, wave) inside the
      // tile, in element order; they are read again for the results (the
      // next tile rewrites them only after the barrier at the top of the loop).
      const uint32_t agg_f = s_aggf;
      const W agg_v = s_aggv;

      // Tile prefix: ordered look-back (one wave, or the whole work-group).
      uint32_t pf = 0;
      W pv = id;
      if (tile != 0) {
        // Only elements before the tile's first segment start need the
        // prefix; a tile that begins with a segment start needs none.
        if (!head_at(0) && !p.block_lookback) {
          // One wave looks back; its first lane publishes the tile's prefix.
          if (w == 0) {
            uint32_t ef;
            W ev;
            wait_own_stores();  // the aggregate is out before the first poll
            wave_pair_lookback<Op, W>(p.status, tile, p.epoch, p.sync, p.force_give_up != 0, ef, ev);
            if (lane == 0) {
              s_lbf[0] = ef;
              s_lbv[0] = ev;
              if (!agg_f && tile + 1 < p.num_tiles) {
                uint32_t of = agg_f;
                W ov = agg_v;
                combine<Op>(ef, ev, of, ov);
                SIO::publish(p.status, tile, p.epoch, st_prefix, of, ov);
              }
            }
          }
          __syncthreads();
          pf = s_lbf[0];
          pv = s_lbv[0];
        } else if (!head_at(0)) {
          // thread t polls tile (base_t - (B - 1 - t)); thread B-1 = nearest.
          int64_t wbase = static_cast<int64_t>(tile) - 1;
          uint32_t ef = 0;  // accumulated prefix of tiles nearer than the window
          W ev = id;
          lookback_wait wait;
          wait.begin();
          bool give_up = p.force_give_up != 0;  // uniform across the work-group
          if (give_up && tid == 0) report_trip(p.sync);
          for (;;) {
            const int64_t j = wbase - (B - 1 - tid);
            uint32_t lf = 0;
            W lv = id;
            int q;
            for (;;) {
              uint32_t st = st_prefix;
              lf = 0;
              lv = id;
              if (j >= 0) st = SIO::read(p.status, static_cast<uint32_t>(j), p.epoch, lf, lv);
              if (give_up) st = st_prefix;
              const bool term = (st == st_prefix) || (st == st_aggregate && lf);
              const uint64_t tm = wave::ballot(term);
              const uint64_t im = wave::ballot(st == st_invalid);
              if (lane == 0) {
                s_q[w] = tm ? (w * 64 + 63 - __builtin_clzll(tm)) : -1;
                s_invalid[w] = im;
              }
              __syncthreads();
              q = -1;
              for (int ww = nw - 1; ww >= 0; --ww)
                if (s_q[ww] >= 0) {
                  q = s_q[ww];
                  break;
                }
              bool bad = false;
              for (int ww = 0; ww < nw; ++ww) {
                const int lo = ww * 64;
                uint64_t need;
                if (q <= lo) need = ~0ull;
                else if (q > lo + 63) need = 0;
                else need = ~0ull << (q - lo);
                bad |= (s_invalid[ww] & need) != 0;
              }
              __syncthreads();
              if (!bad) break;
              // Every work-item backs off; work-item 0's clock decides for all.
              const bool late = wait.failed_round();
              if (tid == 0) {
                s_give_up = late ? 1 : 0;
                if (late) report_trip(p.sync);
              }
              __syncthreads();
              give_up = s_give_up != 0;
            }
            // Fold window tiles [q .. B-1] in tile order.
            uint32_t xf2 = (tid >= q) ? lf : 0;
            W xv2 = (tid >= q) ? lv : id;
            wave_pair_scan<Op>(xf2, xv2);
            if (lane == 63) {
              s_lbf[w] = xf2;
              s_lbv[w] = xv2;
            }
            __syncthreads();
            uint32_t wf = 0;
            W wv = id;
            for (int ww = 0; ww < nw; ++ww) {
              uint32_t yf = s_lbf[ww];
              W yv = s_lbv[ww];
              combine<Op>(wf, wv, yf, yv);
              wf = yf;
              wv = yv;
            }
            __syncthreads();
            // window (earlier) . accumulated (later)
            combine<Op>(wf, wv, ef, ev);
            if (q >= 0) break;
            wbase -= B;
          }
          pf = ef;
          pv = ev;
          if (tid == 0 && !agg_f && tile + 1 < p.num_tiles) {
            uint32_t of = agg_f;
            W ov = agg_v;
            combine<Op>(pf, pv, of, ov);
            SIO::publish(p.status, tile, p.epoch, st_prefix, of, ov);
          }
        }
      }

      // Results: the elements already hold the fold from the start of their
      // wave's row (or of their segment); the rest is P = tile prefix . before
      // (row r, wave w), cut off by a segment start in an earlier lane or
      // among elements 0..k of the vector.
      (void)pf;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint32_t bfr = s_tf[r * nw + w];
        const W bvr = s_tv[r * nw + w];
        const W P = bfr ? bvr : Op::apply(pv, bvr);
        const uint32_t rowf = (fm >> (r * VEC)) & ((1u << VEC) - 1u);
        const bool wcut = (xcut >> r) & 1u;
        W o[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
          const bool cut = wcut || (rowf & ((2u << k) - 1u)) != 0u;
          o[k] = cut ? v[r][k] : Op::apply(P, v[r][k]);
        }
        if constexpr (FULL) {
          T ot[VEC];
#pragma unroll
          for (int k = 0; k < VEC; ++k) ot[k] = static_cast<T>(o[k]);
          u32x4 qv;
          __builtin_memcpy(&qv, ot, 16);
          reinterpret_cast<u32x4_a4 *>(p.out + base)[r * B + tid] = qv;
        } else {
#pragma unroll
          for (int k = 0; k < VEC; ++k) {
            const uint32_t li = static_cast<uint32_t>(r * B + tid) * VEC + k;
            if (li < in_range) p.out[base + li] = static_cast<T>(o[k]);
          }
        }
      }
    };
    if (full)
      tile_body(std::true_type{});
    else
      tile_body(std::false_type{});
  }
}

// --------------------------------------------------------------------------
// Segmented reduction without a look-back.
//
// A reduction needs only each segment's total, so tiles do not wait for
// each other.  Every tile (one work-group, launched in any order) runs the
// same in-tile (head flag, value) scan as the scan kernel and then
//   - writes op(init, total) for every segment that starts and ends inside
//     the tile;
//   - stores head_part[tile]: the fold of the elements before its first
//     segment start (the whole tile if it has none);
//   - stores tail_part[tile]: the fold from its last segment start to the
//     tile's end, when that segment continues into the next tile.
// A fix-up kernel then gives every segment that spans tiles tb < ... <= tl
// the total tail_part[tb] . head_part[tb+1] . ... . head_part[tl], folded in
// tile order (deterministic).  The tiles strictly inside such a segment hold
// no segment start, so their head_part is the whole tile.
// --------------------------------------------------------------------------
template <class T, class Op, int R>
__global__ __launch_bounds__(1024) void segmented_reduce_kernel(params<T> p,
                                                                typename elem_traits<T>::work *head_part,
                                                                typename elem_traits<T>::work *tail_part) {
  using W = typename elem_traits<T>::work;
  constexpr int VEC = 16 / sizeof(T);
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int nw = B >> 6;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const uint64_t tile_elems = static_cast<uint64_t>(B) * R * VEC;
  const W id = Op::identity();

  extern __shared__ uint32_t s_heads[];  // tile_elems + 1 bits
  __shared__ uint32_t s_tf[R * max_waves];  // per (row, wave): totals, then exclusive prefixes
  __shared__ W s_tv[R * max_waves];
  // offsets[s_lo .. s_lo + off_cap) of the segments touching the tile: the
  // binary search that numbers a finished segment then runs in LDS instead
  // of as a chain of dependent global loads.
  constexpr uint32_t off_cap = 1024;
  __shared__ uint64_t s_off[off_cap];
  const uint32_t head_words = static_cast<uint32_t>((tile_elems + 1 + 31) / 32);

  const uint32_t tile = blockIdx.x;
  if (tile >= p.num_tiles) return;
  const uint64_t base = static_cast<uint64_t>(tile) * tile_elems;
  uint64_t s_lo = p.tile_seg[tile];
  uint64_t s_hi = p.tile_seg[tile + 1];

  const bool full = p.vec_io && (base + tile_elems <= p.total);
  W v[R][VEC];
  u32x4 q[R];
  if (full) {
    const u32x4_a4 *src = reinterpret_cast<const u32x4_a4 *>(p.in + base);
#pragma unroll
    for (int r = 0; r < R; ++r) q[r] = src[r * B + tid];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < VEC; ++k) {
        const uint64_t idx = base + static_cast<uint64_t>(r * B + tid) * VEC + k;
        v[r][k] = (idx < p.total) ? elem_traits<T>::to_work(p.in[idx]) : id;
      }
  }
  for (uint32_t i = tid; i < head_words; i += B) s_heads[i] = 0;
  if (s_lo >= p.num_segments) s_lo = p.num_segments - 1;
  if (s_hi >= p.num_segments) s_hi = p.num_segments - 1;
  if (s_hi < s_lo) s_hi = s_lo;
  // Segments s_lo..s_hi touch the tile; their offsets and the next one's.
  const uint64_t n_seg = s_hi - s_lo + 1;
  const bool off_in_lds = n_seg + 1 <= off_cap;
  if (off_in_lds)
    for (uint64_t i = tid; i <= n_seg; i += B) s_off[i] = clamp_off(p.offsets, s_lo + i, p.total);
  __syncthreads();
  // The tile's head bits (tile elements 0..tile_elems), marked from the
  // offsets: a reduction needs no global head bitmap, so its plan has no
  // bitmap to clear and fill.
  for (uint64_t i = tid; i < n_seg; i += B) {
    const uint64_t b = off_in_lds ? s_off[i] : clamp_off(p.offsets, s_lo + i, p.total);
    const uint64_t e = off_in_lds ? s_off[i + 1] : clamp_off(p.offsets, s_lo + i + 1, p.total);
    if (e > b && b >= base && b - base <= tile_elems) {
      const uint64_t pos = b - base;
      atomicOr(&s_heads[pos >> 5], 1u << (pos & 31));
    }
  }
  __syncthreads();
  if (full) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      T tmp[VEC];
      __builtin_memcpy(tmp, &q[r], 16);
#pragma unroll
      for (int k = 0; k < VEC; ++k) v[r][k] = elem_traits<T>::to_work(tmp[k]);
    }
  }

  auto head_at = [&](uint64_t pos) -> uint32_t {
    return (s_heads[pos >> 5] >> (pos & 31)) & 1u;
  };
  // Largest s in [s_lo, s_hi] with offsets[s] <= e.
  auto segment_of = [&](uint64_t e) -> uint64_t {
    if (!off_in_lds) return find_segment(p.offsets, s_lo, s_hi, e, p.total);
    uint64_t lo = 0, hi = s_hi - s_lo;
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) / 2;
      if (s_off[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    return s_lo + lo;
  };

  const uint32_t fm = row_head_bits<R, VEC>(s_heads, tid, B);
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 1; k < VEC; ++k)
      v[r][k] = ((fm >> (r * VEC + k)) & 1u) ? v[r][k] : Op::apply(v[r][k - 1], v[r][k]);

  const uint32_t xcut = fold_wave_prefixes<Op, R, VEC>(fm, v, s_tf, s_tv, nw, w, lane);
  __syncthreads();
  if (w == 0) {
    uint32_t af;
    W av;
    tile_pair_prefixes<Op, W>(R * nw, s_tf, s_tv, af, av);
    (void)af;
    (void)av;
  }
  __syncthreads();
  const uint64_t tile_left = p.total - base;  // > 0: the tile holds elements
#pragma unroll
  for (int r = 0; r < R; ++r) {
    // The elements hold the fold from their wave's row start (or segment
    // start); Pv is the (row r, wave w) prefix inside the tile, cut off by
    // a segment start in an earlier lane.
    const uint32_t rowf = (fm >> (r * VEC)) & ((1u << VEC) - 1u);
    const bool wcut = (xcut >> r) & 1u;
    const uint32_t Pf = s_tf[r * nw + w] | (wcut ? 1u : 0u);
    const W Pv = s_tv[r * nw + w];
    const uint64_t l0 = static_cast<uint64_t>(r * B + tid) * VEC;  // index inside the tile
    // Bit k: element k of the vector is the last of its segment in this
    // tile (a segment start follows, or the data or the tile ends there).
    uint32_t ends = (rowf >> 1) | (head_at(l0 + VEC) << (VEC - 1));
    if (l0 + VEC == tile_elems) ends |= 1u << (VEC - 1);
    if (l0 < tile_left && tile_left <= l0 + VEC) ends |= 1u << (tile_left - 1 - l0);
    if (tile_left <= l0) ends = 0;
    else if (tile_left < l0 + VEC) ends &= (1u << (tile_left - l0)) - 1u;
    if (ends == 0) continue;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
      if (!((ends >> k) & 1u)) continue;
      const uint64_t e = base + l0 + k;
      // a segment start among elements 0..k of the vector cuts the prefix off
      const bool cut = wcut || (rowf & ((2u << k) - 1u)) != 0u;
      const W val = cut ? v[r][k] : Op::apply(Pv, v[r][k]);
      const bool next_head = (e + 1 == p.total) || head_at(l0 + k + 1);
      if (cut || Pf) {  // a segment start lies in [tile start, e]
        if (next_head) {
          const uint64_t s = segment_of(e);
          p.out[s] = static_cast<T>(Op::apply(p.init, val));
        } else {
          tail_part[tile] = val;  // the tile's last element; the segment goes on
        }
      } else {
        head_part[tile] = val;
      }
    }
  }
}

// op(init, tail_part[tb] . head_part[tb+1] . ... . head_part[tl]) for every
// segment whose elements span tiles tb < tl.
template <class T, class Op>
__global__ void segmented_fixup_kernel(params<T> p, const typename elem_traits<T>::work *head_part,
                                       const typename elem_traits<T>::work *tail_part, uint64_t tile_elems) {
  using W = typename elem_traits<T>::work;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < p.num_segments;
       s += stride) {
    const uint64_t b = clamp_off(p.offsets, s, p.total);
    const uint64_t e = clamp_off(p.offsets, s + 1, p.total);
    if (e <= b) continue;
    const uint64_t tb = b / tile_elems;
    const uint64_t tl = (e - 1) / tile_elems;
    if (tb == tl || tl >= p.num_tiles) continue;
    W acc = tail_part[tb];
    uint64_t t = tb + 1;
    // Eight partials in flight per step, folded in tile order.
    for (; t + 8 <= tl + 1; t += 8) {
      W h[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) h[u] = head_part[t + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = Op::apply(acc, h[u]);
    }
    for (; t <= tl; ++t) acc = Op::apply(acc, head_part[t]);
    p.out[s] = static_cast<T>(Op::apply(p.init, acc));
  }
}

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------
// Default tile when the caller gives no configuration, by total element
// count and element size.  Read off the first 400 of 500 segmented shapes
// up to 2^28 elements timed on MI355X (the tuner's timing database); on the
// other 100 it has a mean regret of ~4 % against the best of the 15
// configurations, where any single fixed tile has 15-70 %.
//   scan:   n < 2^16: 256 x 2;  n < 2^20: 512 x 2 (8-byte: 1024 x 2);
//           n < 2^26: 512 x 8;  otherwise 1024 x 8
//   reduce: n < 2^16: 128 x 2;  n < 2^20: 256 x 2;  n < 2^24: 1024 x 2;
//           otherwise 256 x 4 (8-byte: 512 x 4)
// Large scans take 8-row tiles (two 512-wide or one 1024-wide group per CU,
// 123 VGPRs for int64); large reductions take 4-row tiles, which leave room
// for several groups per CU (the reduction's groups never wait for each
// other).
inline uint32_t default_work_group(bool reduce, uint64_t total, int elem_bytes) {
  if (total < (1ull << 16)) return reduce ? 128 : 256;
  if (total < (1ull << 20)) return reduce ? 256 : (elem_bytes == 8 ? 1024 : 512);
  if (reduce) return total < (1ull << 24) ? 1024 : (elem_bytes == 8 ? 512 : 256);
  return total < (1ull << 26) ? 512 : 1024;
}
inline int default_rows(bool reduce, uint64_t total) {
  if (total < (1ull << 20)) return 2;
  if (reduce) return total < (1ull << 24) ? 2 : 4;
  return 8;
}

template <class T, class Op>
struct launch_plan {
  params<T> p;
  uint32_t wg = 256;
  int rows = 8;
  uint32_t grid = 1;
  size_t lds = 0;
  uint64_t tile_elems = 0;
  uint32_t *tile_seg = nullptr;
  uint32_t *heads = nullptr;
  uint64_t head_words = 0;
  void *region = nullptr;
  region_state *rs = nullptr;
  size_t region_used = 0;
  size_t region_cap = 0;
};

template <class T, class Op, int R>
acpp_status launch_rows(launch_plan<T, Op> &lp, hipStream_t stream, const char *name) {
  acpp_status st = fill_zero(lp.heads, lp.head_words * sizeof(uint32_t), stream);
  if (st != ACPP_SUCCESS) return st;
  hipLaunchKernelGGL(setup_kernel<T>, dim3(static_cast<uint32_t>(std::min<uint64_t>(ceil_div(lp.p.num_segments + 1, 256), 4096))),
                     dim3(256), 0, stream, lp.p, lp.tile_seg, lp.heads, lp.tile_elems);
  st = check_launch(name);
  if (st != ACPP_SUCCESS) return st;
  hipLaunchKernelGGL((segmented_kernel<T, Op, R>), dim3(lp.grid), dim3(lp.wg), lp.lds, stream, lp.p);
  return check_launch(name);
}

// Segmented reduction: set up (tile -> first segment, init into empty
// segments), one work-group per tile (its head bits marked in LDS from the
// offsets, so no global bitmap is cleared or filled), then the fix-up of
// segments that span tiles.
template <class T, class Op>
acpp_status enqueue_reduce(acpp_allocation_group *g, launch_plan<T, Op> &lp, hipStream_t stream, const T *in,
                           T *out, const uint64_t *offsets, uint64_t num_segments, uint64_t total,
                           typename elem_traits<T>::work init, uint64_t num_tiles, const char *name) {
  using W = typename elem_traits<T>::work;
  void *parts = nullptr;
  acpp_status st = group_obtain(g, 2 * num_tiles * sizeof(W), &parts);
  if (st != ACPP_SUCCESS) return st;
  W *head_part = static_cast<W *>(parts);
  W *tail_part = head_part + num_tiles;
  params<T> &p = lp.p;
  p.in = in;
  p.out = out;
  p.offsets = offsets;
  p.tile_seg = lp.tile_seg;
  p.heads = nullptr;  // the reduce kernel marks its heads from the offsets
  p.head_words_total = 0;
  p.block_lookback = 0;
  p.force_give_up = 0;
  p.num_segments = num_segments;
  p.total = total;
  p.init = init;
  p.reduce = 1;
  p.vec_io = (reinterpret_cast<uintptr_t>(in) & 3) == 0;
  p.num_tiles = static_cast<uint32_t>(num_tiles);
  p.epoch = 0;
  p.ticket_base = 0;
  p.sync = nullptr;
  p.status = nullptr;
  lp.grid = static_cast<uint32_t>(num_tiles);

  hipLaunchKernelGGL(setup_kernel<T>, dim3(static_cast<uint32_t>(std::min<uint64_t>(ceil_div(num_segments + 1, 256), 4096))),
                     dim3(256), 0, stream, p, lp.tile_seg, static_cast<uint32_t *>(nullptr), lp.tile_elems);
  st = check_launch(name);
  if (st != ACPP_SUCCESS) return st;
  if (total > 0) {
    if (lp.rows == 2)
      hipLaunchKernelGGL((segmented_reduce_kernel<T, Op, 2>), dim3(lp.grid), dim3(lp.wg), lp.lds, stream, p,
                         head_part, tail_part);
    else if (lp.rows == 4)
      hipLaunchKernelGGL((segmented_reduce_kernel<T, Op, 4>), dim3(lp.grid), dim3(lp.wg), lp.lds, stream, p,
                         head_part, tail_part);
    else
      hipLaunchKernelGGL((segmented_reduce_kernel<T, Op, 8>), dim3(lp.grid), dim3(lp.wg), lp.lds, stream, p,
                         head_part, tail_part);
    st = check_launch(name);
    if (st != ACPP_SUCCESS) return st;
    hipLaunchKernelGGL((segmented_fixup_kernel<T, Op>),
                       dim3(static_cast<uint32_t>(std::min<uint64_t>(ceil_div(num_segments, 256), 4096))), dim3(256), 0,
                       stream, p, head_part, tail_part, lp.tile_elems);
    st = check_launch(name);
  }
  return st;
}

template <class T, class Op>
acpp_status enqueue(acpp_allocation_group *g, hipStream_t stream, const T *in, T *out,
                    const uint64_t *offsets, uint64_t num_segments, uint64_t total,
                    typename elem_traits<T>::work init, bool reduce,
                    const acpp_launch_config *cfg, const char *name) {
  static_assert(sizeof(T) >= 4, "16-byte vector accesses at 4-byte alignment (u32x4_a4)");
  constexpr int VEC = 16 / sizeof(T);
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  launch_plan<T, Op> lp;
  lp.wg = (cfg && cfg->work_group_size) ? cfg->work_group_size : default_work_group(reduce, total, static_cast<int>(sizeof(T)));
  if (!valid_work_group_size(lp.wg))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": bad work_group_size");
  lp.rows = (cfg && cfg->items_per_thread) ? static_cast<int>(cfg->items_per_thread) : default_rows(reduce, total);
  if (lp.rows != 2 && lp.rows != 4 && lp.rows != 8)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": items_per_thread must be 2, 4 or 8");
  lp.tile_elems = static_cast<uint64_t>(lp.wg) * lp.rows * VEC;
  const uint64_t num_tiles = std::max<uint64_t>(1, ceil_div(total, lp.tile_elems));
  if (num_tiles >= (1ull << 31) || num_segments >= (1ull << 32))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": problem too large");
  // the scan kernel stores 2 x wg words of head bits (>= the tile_elems + 1
  // bits it needs, since rows x 16-byte vector <= 32 elements)
  lp.lds = std::max<size_t>(static_cast<size_t>((lp.tile_elems + 1 + 31) / 32), 2 * static_cast<size_t>(lp.wg)) * 4;

  void *ts = nullptr;
  st = group_obtain(g, (num_tiles + 1) * sizeof(uint32_t), &ts);
  if (st != ACPP_SUCCESS) return st;
  lp.tile_seg = static_cast<uint32_t *>(ts);
  if (reduce) return enqueue_reduce<T, Op>(g, lp, stream, in, out, offsets, num_segments, total, init, num_tiles, name);
  // Head bitmap over elements [0, total] plus one spare word (scans).
  lp.head_words = (total + 1 + 31) / 32 + 1;
  void *hb = nullptr;
  st = group_obtain(g, lp.head_words * sizeof(uint32_t), &hb);
  if (st != ACPP_SUCCESS) return st;
  lp.heads = static_cast<uint32_t *>(hb);
  using SIO = seg_status<typename elem_traits<T>::work>;
  static_assert(SIO::words <= seg_stride_words, "status slot too small");
  lp.region_used = sizeof(sync_header) + num_tiles * seg_stride_words * sizeof(uint64_t);
  st = group_obtain_sync(g, lp.region_used, stream, &lp.region, &lp.region_cap, &lp.rs);
  if (st != ACPP_SUCCESS) return st;

  region_launch rl;
  st = begin_region_launch(g, lp.region, lp.region_used, lp.region_cap, lp.rs, stream, &rl);
  if (st != ACPP_SUCCESS) return st;

  params<T> &p = lp.p;
  p.in = in;
  p.out = out;
  p.offsets = offsets;
  p.tile_seg = lp.tile_seg;
  p.heads = lp.heads;
  p.head_words_total = lp.head_words;
  p.block_lookback = (cfg && (cfg->flags & ACPP_VARIANT_BLOCK_LOOKBACK)) ? 1 : 0;
  p.force_give_up = (cfg && (cfg->flags & ACPP_VARIANT_FORCE_GIVE_UP)) ? 1 : 0;
  p.num_segments = num_segments;
  p.total = total;
  p.init = init;
  p.reduce = reduce ? 1 : 0;
  p.vec_io = ((reinterpret_cast<uintptr_t>(in) | (reduce ? 0 : reinterpret_cast<uintptr_t>(out))) & 3) == 0;
  p.num_tiles = static_cast<uint32_t>(num_tiles);
  p.epoch = rl.epoch;
  p.ticket_base = rl.ticket_base;
  p.sync = static_cast<sync_header *>(lp.region);
  p.status = reinterpret_cast<uint64_t *>(static_cast<char *>(lp.region) + sizeof(sync_header));

  // Persistent grid: one wave of resident work-groups, each drawing tiles
  // from the queue until it is empty.  With ACPP_VARIANT_TILE_PER_GROUP the
  // grid has one work-group per tile (the hardware dispatcher then keeps
  // the CUs full while groups wait in their look-back), still fed by the
  // same queue.  Every work-group draws one ticket past the last tile
  // before it exits; the ticket base counts those too.
  const bool per_tile = cfg && (cfg->flags & ACPP_VARIANT_TILE_PER_GROUP);
  auto pick = [&](auto kernel) -> acpp_status {
    if (per_tile) {
      lp.grid = static_cast<uint32_t>(num_tiles);
      return ACPP_SUCCESS;
    }
    int resident = 1;
    ACPP_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, static_cast<int>(lp.wg), lp.lds));
    if (resident < 1) resident = 1;
    uint64_t cap = static_cast<uint64_t>(resident) * cus;
    if (cfg && cfg->groups_per_cu) cap = std::min<uint64_t>(cap, static_cast<uint64_t>(cfg->groups_per_cu) * cus);
    lp.grid = static_cast<uint32_t>(std::min<uint64_t>(cap, num_tiles));
    return ACPP_SUCCESS;
  };
  acpp_status s;
  if (lp.rows == 2) {
    if ((s = pick(segmented_kernel<T, Op, 2>)) != ACPP_SUCCESS) return s;
    s = launch_rows<T, Op, 2>(lp, stream, name);
  } else if (lp.rows == 4) {
    if ((s = pick(segmented_kernel<T, Op, 4>)) != ACPP_SUCCESS) return s;
    s = launch_rows<T, Op, 4>(lp, stream, name);
  } else {
    if ((s = pick(segmented_kernel<T, Op, 8>)) != ACPP_SUCCESS) return s;
    s = launch_rows<T, Op, 8>(lp, stream, name);
  }
  // every work-group draws one ticket past the last tile before it exits
  end_region_launch(lp.rs, rl, static_cast<uint32_t>(num_tiles) + lp.grid, s);
  return s;
}

template <class T>
acpp_status entry(acpp_allocation_group *g, hipStream_t stream, const T *in, T *out,
                  const uint64_t *offsets, uint64_t num_segments, uint64_t total, T init, int op,
                  bool reduce, const acpp_launch_config *cfg, const char *name) {
  using W = typename elem_traits<T>::work;
  if (!g) return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": scratch group is NULL");
  if (num_segments > 0 && !offsets)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": offsets is NULL");
  if ((total > 0 && !in) || ((total > 0 || (reduce && num_segments > 0)) && !out))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": in or out is NULL");
  if (op < ops::op_plus || op > ops::op_max)
    return set_error(ACPP_ERROR_UNSUPPORTED, std::string(name) + ": segmented forms support plus, multiply, min and max");
  if (num_segments == 0) return clear_error();
  return with_op<W, T>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    if constexpr (Op::id > ops::op_max) {
      return set_error(ACPP_ERROR_UNSUPPORTED, "unsupported operator");
    } else {
      acpp_status s = enqueue<T, Op>(g, stream, in, out, offsets, num_segments, total,
                                     elem_traits<T>::to_work(init), reduce, cfg, name);
      return s == ACPP_SUCCESS ? clear_error() : s;
    }
  });
}

}  // namespace seg_detail
}  // namespace acpp_impl

using namespace acpp_impl;
using namespace acpp_impl::seg_detail;

// --------------------------------------------------------------------------
// Graph plans
// --------------------------------------------------------------------------
struct acpp_segmented_plan {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
};

namespace {
template <class F>
acpp_status capture_plan(acpp_allocation_group *g, F &&enqueue_fn, acpp_segmented_plan **out) {
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "plan output pointer is NULL");
  int cus = 0;
  acpp_status st = enter_device(g, &cus);
  if (st != ACPP_SUCCESS) return st;
  hipStream_t cs = nullptr;
  ACPP_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  std::unique_ptr<acpp_segmented_plan> plan(new acpp_segmented_plan);
  plan->device = group_device(g);
  // Relaxed mode: the first capture of a plan may have to allocate its
  // scratch (hipMalloc through the allocation cache); the allocation is made
  // at once and is not part of the graph.
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return hip_error(e, "hipStreamBeginCapture");
  }
  acpp_status s = enqueue_fn(cs);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(cs, &graph);
  (void)hipStreamDestroy(cs);
  if (s != ACPP_SUCCESS) {
    if (graph) (void)hipGraphDestroy(graph);
    return s;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  plan->graph = graph;
  e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_error(e, "hipGraphInstantiate");
  }
  *out = plan.release();
  return clear_error();
}
}  // namespace

extern "C" {

#define ACPP_DEFINE_SEGMENTED(tag, T)                                                                 \
  acpp_status acpp_segmented_inclusive_scan_##tag(acpp_allocation_group *scratch, acpp_stream_t stream, \
                                                   const T *first, uint64_t n, T *d_first,             \
                                                   const uint64_t *offsets, uint64_t num_segments,     \
                                                   acpp_algorithm_op op,                               \
                                                   const acpp_launch_config *config) {                 \
    return entry<T>(scratch, static_cast<hipStream_t>(stream), first, d_first, offsets, num_segments,  \
                    n, T{}, op, false, config, "acpp_segmented_inclusive_scan_" #tag);                \
  }                                                                                                   \
  acpp_status acpp_segmented_reduce_##tag(acpp_allocation_group *scratch, acpp_stream_t stream,       \
                                          const T *first, uint64_t n, const uint64_t *offsets,        \
                                          uint64_t num_segments, T *out, T init,                      \
                                          acpp_algorithm_op op, const acpp_launch_config *config) {   \
    return entry<T>(scratch, static_cast<hipStream_t>(stream), first, out, offsets, num_segments, n,   \
                    init, op, true, config, "acpp_segmented_reduce_" #tag);                           \
  }                                                                                                   \
  acpp_status acpp_segmented_plan_inclusive_scan_##tag(                                                \
      acpp_allocation_group *scratch, const T *first, uint64_t n, T *d_first,                         \
      const uint64_t *offsets, uint64_t num_segments, acpp_algorithm_op op,                           \
      const acpp_launch_config *config, acpp_segmented_plan **plan) {                                 \
    if (!scratch) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "scratch group is NULL");             \
    return capture_plan(                                                                              \
        scratch,                                                                                      \
        [&](hipStream_t cs) {                                                                         \
          return entry<T>(scratch, cs, first, d_first, offsets, num_segments, n, T{}, op, false,      \
                          config, "acpp_segmented_plan_inclusive_scan_" #tag);                       \
        },                                                                                            \
        plan);                                                                                        \
  }                                                                                                   \
  acpp_status acpp_segmented_plan_reduce_##tag(                                                        \
      acpp_allocation_group *scratch, const T *first, uint64_t n, const uint64_t *offsets,            \
      uint64_t num_segments, T *out, T init, acpp_algorithm_op op, const acpp_launch_config *config, \
      acpp_segmented_plan **plan) {                                                                   \
    if (!scratch) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "scratch group is NULL");             \
    return capture_plan(                                                                              \
        scratch,                                                                                      \
        [&](hipStream_t cs) {                                                                         \
          return entry<T>(scratch, cs, first, out, offsets, num_segments, n, init, op, true, config,   \
                          "acpp_segmented_plan_reduce_" #tag);                                        \
        },                                                                                            \
        plan);                                                                                        \
  }

ACPP_DEFINE_SEGMENTED(i32, int32_t)
ACPP_DEFINE_SEGMENTED(i64, int64_t)
ACPP_DEFINE_SEGMENTED(f32, float)
ACPP_DEFINE_SEGMENTED(f64, double)

acpp_status acpp_segmented_plan_launch(acpp_segmented_plan *plan, acpp_stream_t stream) {
  if (!plan || !plan->exec) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "plan is NULL");
  int cur = 0;
  ACPP_CHECK_HIP(hipGetDevice(&cur));
  if (cur != plan->device) ACPP_CHECK_HIP(hipSetDevice(plan->device));
  ACPP_CHECK_HIP(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)));
  return clear_error();
}

acpp_status acpp_segmented_plan_destroy(acpp_segmented_plan *plan) {
  if (!plan) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "plan is NULL");
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  delete plan;
  return clear_error();
}

}  // extern "C"
