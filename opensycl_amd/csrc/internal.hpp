// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#include <string>
#include <type_traits>

#include "acpp_mi355x.h"
#include "acpp/device/ops.hpp"
#include "acpp/device/wave.hpp"

namespace acpp_impl {

namespace ops = ::acpp::ops;
namespace wave = ::acpp::wave;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
acpp_status set_error(acpp_status code, const std::string &msg);
acpp_status clear_error();
acpp_status hip_error(hipError_t e, const char *what);

#define ACPP_CHECK_HIP(expr)                                                   \
  do {                                                                         \
    hipError_t acpp_e_ = (expr);                                               \
    if (acpp_e_ != hipSuccess) return ::acpp_impl::hip_error(acpp_e_, #expr);  \
  } while (0)

// Check the launch that was just enqueued.
acpp_status check_launch(const char *kernel_name);

// ---------------------------------------------------------------------------
// Synchronisation regions.
//
// A sync region is scratch memory that is zero when first handed out and
// that every kernel using it leaves in a reusable state:
//   - counters (tile ticket, done counter, group ticket) are reset to zero by
//     the last work-group of the launch;
//   - scan status words carry the launch's epoch, so words written by an
//     earlier launch are recognised as stale without clearing them.
// This removes the reference's separate status-initialisation kernel
// (decoupled_lookback_scan.hpp:584-589) and any memset from the hot path.
// ---------------------------------------------------------------------------
struct sync_header {
  uint32_t epoch;           // unused by eager launches (epochs are kernel arguments)
  uint32_t tile_counter;    // dynamic tile ticket, monotonic across launches
  uint32_t done_counter;    // group ticket of the single-launch reductions
  uint32_t watchdog_trips;  // waits that gave up (never expected)
  // Device address of the pinned host word that makes a give-up visible to
  // the host without a device read (host_trip_flag in runtime.cpp); written
  // into every region whenever the region is cleared (reset_sync_region).
  uint32_t *host_trip_flag;
  uint32_t pad[58];         // header occupies 256 bytes
};

static_assert(sizeof(sync_header) == 256, "sync header is 256 bytes");
static_assert(offsetof(sync_header, host_trip_flag) == 16, "trip flag pointer is the second 16-byte chunk");

// A wait inside a kernel gave up: count it in the region and raise the
// host-visible flag, which the host reports as ACPP_ERROR_WATCHDOG at its
// next synchronisation point (acpp_check_asynchronous_errors).  The flag is
// written with a plain system-scope store (no read-modify-write on host
// memory).
// The flag is stored through a global-address-space pointer: a flat store
// (the generic pointer) counts on two wait counters out of order, and in a
// kernel that holds one anywhere the compiler waits for every outstanding
// load where one would do.
__device__ __forceinline__ void report_trip(sync_header *sync) {
  __hip_atomic_fetch_add(&sync->watchdog_trips, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t *flag = sync->host_trip_flag;
  if (flag) {
    using global_u32 = __attribute__((address_space(1))) uint32_t;
    __hip_atomic_store((global_u32 *)flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Waiting inside a decoupled look-back.
//   - Backoff: a poll round that finds an unfinished predecessor sleeps
//     before polling again, twice as long each time up to 64 s_sleep(1)
//     units (~4k clocks).  With thousands of small tiles in flight,
//     back-to-back polls flood the memory system and starve the very tiles
//     being waited for: 65,536-segment scans with 256-element tiles averaged
//     over 2e5 polls per look-back on MI355X without it.
//   - Watchdog: the wait is bounded in wall-clock time (the steady 100 MHz
//     counter), not in polls, so a slow but live chain never trips it.
//     After lookback_timeout_ticks the look-back gives up, reports a trip
//     (report_trip) and continues with what it has, so every kernel still
//     ends; its results are then wrong, and the host reports the trip as an
//     error (ACPP_ERROR_WATCHDOG) at its next synchronisation point.  It
//     only guards against a device that stops making progress.
//   - ACPP_VARIANT_FORCE_GIVE_UP (a launch-config flag for tests) makes
//     every such wait give up at once, so the error path can be exercised
//     with kernels that still end normally.
constexpr uint64_t lookback_timeout_ticks = 1ull << 28;  // ~2.7 s at 100 MHz
constexpr uint32_t lookback_max_backoff = 64;

struct lookback_wait {
  uint64_t start;
  uint32_t backoff;
  __device__ __forceinline__ void begin() {
    start = static_cast<uint64_t>(wall_clock64());
    backoff = 0;
  }
  // After an unsuccessful round: back off, then report whether the wait has
  // run out of time.  The result is uniform across the wave.
  __device__ __forceinline__ bool failed_round() {
    backoff = backoff ? (backoff < lookback_max_backoff ? 2 * backoff : backoff) : 1;
    for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(1);
    return static_cast<uint64_t>(wall_clock64()) - start > lookback_timeout_ticks;
  }
};

// Wait until this wave's own vector memory operations have completed
// (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their no-wait maxima).  On
// gfx950 the counter covers stores, so a look-back that calls this after
// publishing its tile's aggregate starts polling only once that status word
// has been written: polls issued right behind the store kept it queued, and
// every later tile waits for it.  Measured on MI355X (2^28 int32 exclusive
// scan): 383.6 -> 377.0 us.
__device__ __forceinline__ void wait_own_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Tag of a look-back status word (the upper 32 bits of every 64-bit status
// word): epoch << 3 | head flag << 2 | status (2 bits).  All look-back
// kernels (scans, copy_if, segmented scans) use this one layout because they
// share cached scratch regions and each region's epoch counter: a word left
// by an earlier launch of any of them carries an older epoch and never
// reads as valid.  (With per-kernel layouts, a segmented-scan word of epoch
// e decoded as a scan word of epoch 2e + head, and was taken as live.)
__host__ __device__ constexpr uint32_t status_tag(uint32_t epoch, uint32_t head, uint32_t status) {
  return (epoch << 3) | (head << 2) | status;
}
__host__ __device__ constexpr uint32_t tag_epoch(uint32_t tag) { return tag >> 3; }
__host__ __device__ constexpr uint32_t tag_head(uint32_t tag) { return (tag >> 2) & 1u; }
__host__ __device__ constexpr uint32_t tag_status(uint32_t tag) { return tag & 3u; }

// Epochs live in 29 bits of a tag; 0 is never a live epoch.
constexpr uint32_t max_epoch = (1u << 29) - 1;

}  // namespace acpp_impl

// Internal part of the allocation group used by the kernels.
struct acpp_allocation_group;
namespace acpp_impl {

// Host-side state of a sync region, kept with the region while it lives in
// the allocation cache.  Eager launches pass the epoch and the ticket base as
// kernel arguments: the host knows how many launches used the region and
// how many tile tickets each one drew, so no kernel has to reset anything.
struct region_state {
  enum mode_t : int { fresh = 0, host_managed = 1, graph = 2 };
  int mode = fresh;
  uint32_t epoch = 0;        // epoch of the last launch (host_managed)
  uint32_t ticket_base = 0;  // tile ticket counter value before the next launch
  // A captured launch uses the region: every replay of that graph clears
  // the region and draws tickets from 0, unseen by the host, so eager
  // launches clear it first as long as the graph may run.  Set at capture;
  // cleared when the region is handed out by the cache again (a graph must
  // not outlive the group that holds its regions).
  bool captured = false;
};

// Obtain ordinary scratch from the group.
acpp_status group_obtain(acpp_allocation_group *g, size_t bytes, void **out);
// Epoch and ticket base of one launch on a sync region.  Eager launches
// continue the region's host-side count.  Under stream capture the graph may
// be replayed any number of times, so the captured work clears the region
// first (with a kernel) and always runs with epoch 1 and ticket base 0.
struct region_launch {
  uint32_t epoch = 1;
  uint32_t ticket_base = 0;
  bool capturing = false;
};
// Prepare `region` (of `capacity` bytes, `used` of them needed) for a launch.
acpp_status begin_region_launch(acpp_allocation_group *g, void *region, size_t used, size_t capacity,
                                region_state *rs, hipStream_t stream, region_launch *out);
// The same for a region that belongs to no allocation group (the library's
// static no-allocation scratch): `trip_flag` (may be NULL) is stored in the
// header whenever the region is cleared.
acpp_status begin_region_launch_flag(uint32_t *trip_flag, void *region, size_t used, size_t capacity,
                                     region_state *rs, hipStream_t stream, region_launch *out);

// A sync region handed to a launcher instead of one from an allocation
// group.  The caller makes its device current and knows its CU count.
struct fixed_region {
  void *ptr;            // device address
  size_t capacity;      // bytes
  region_state *rs;
  uint32_t *trip_flag;  // host trip flag of the device, or NULL
  int compute_units;
};
// Record the launch: `tickets` tile tickets drawn; a failed launch makes the
// region start over next time (it may or may not have run).
void end_region_launch(region_state *rs, const region_launch &l, uint32_t tickets, acpp_status launch_status);
// Clear the first `bytes` of a sync region with a kernel and store the
// device's host trip flag address in its header.
acpp_status reset_sync_region(void *region, size_t bytes, int device, hipStream_t stream);
acpp_status reset_sync_region_flag(void *region, size_t bytes, uint32_t *trip_flag, hipStream_t stream);
// Pinned, device-visible host word of `device` raised by report_trip
// (allocated on first use), and the same without allocating (NULL if no
// call has allocated it yet).
uint32_t *host_trip_flag(int device);
uint32_t *host_trip_flag_if_allocated(int device);
// Obtain a sync region of at least `bytes` (header included).  Fresh regions
// are zeroed on `stream` before first use.  `capacity_bytes` receives the
// region's real size, `state` its host-side state.
acpp_status group_obtain_sync(acpp_allocation_group *g, size_t bytes,
                              hipStream_t stream, void **out,
                              size_t *capacity_bytes, region_state **state = nullptr);

// Zero `bytes` of device memory with a kernel (also inside stream capture,
// where memset nodes are avoided).
acpp_status fill_zero(void *p, size_t bytes, hipStream_t stream);

// True when `stream` is being captured into a graph.
bool stream_is_capturing(hipStream_t stream);
int group_device(const acpp_allocation_group *g);
// Make the group's device current; returns its compute-unit count.
acpp_status enter_device(acpp_allocation_group *g, int *compute_units);

// ---------------------------------------------------------------------------
// Element types
// ---------------------------------------------------------------------------
struct bf16_t {
  uint16_t bits;
};

template <class T>
struct elem_traits {
  using storage = T;
  // Type the kernels compute in.
  using work = std::conditional_t<
      (std::is_integral_v<T> && sizeof(T) < 4),
      std::conditional_t<std::is_signed_v<T>, int32_t, uint32_t>, T>;
  __host__ __device__ static work to_work(T x) { return static_cast<work>(x); }
  __host__ __device__ static T from_work(work x) { return static_cast<T>(x); }
};

template <>
struct elem_traits<bf16_t> {
  using storage = uint16_t;
  using work = float;
  __host__ __device__ static float to_work(bf16_t x) {
    return __builtin_bit_cast(float, static_cast<uint32_t>(x.bits) << 16);
  }
};

}  // namespace acpp_impl
