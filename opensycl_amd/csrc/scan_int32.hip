// C-ABI entry points: inclusive_scan / exclusive_scan for 32-bit integers,
// and the exclusive int32 scan that allocates nothing.
#include <mutex>

#include "scan_launch.hpp"

using namespace acpp_impl;
using namespace acpp_impl::scan_detail;

namespace {

// Scratch of acpp_exclusive_scan_i32_noalloc: one sync region (header plus
// a 128-byte status line per tile) in a file-scope device array instead of
// memory from an allocation group.  2 MiB of status lines: 16384 tiles, i.e.
// up to 2^29 int32 elements with the default 1024 x 8-row tiles.  Each GPU
// holds its own copy of the array (its device address comes from
// hipGetSymbolAddress on that device) until the process ends.
constexpr uint64_t noalloc_max_tiles = 16384;
constexpr size_t noalloc_bytes =
    sizeof(sync_header) + noalloc_max_tiles * status_stride_words * sizeof(uint64_t);
static __device__ __align__(256) unsigned char scratch[noalloc_bytes];

struct noalloc_device {
  void *ptr = nullptr;  // this device's address of `scratch` (never the host-side name)
  int compute_units = 0;
  region_state rs;
};
constexpr int noalloc_max_devices = 64;
std::mutex g_noalloc_mutex;  // host bookkeeping of the calls (epoch, ticket base)
noalloc_device g_noalloc[noalloc_max_devices];

// Once per device, with that device current: the array's device address and
// the CU count.  The region starts in "graph" mode, so the first launch
// clears it with a kernel before use.
acpp_status noalloc_setup(int dev, noalloc_device **out) {
  noalloc_device &d = g_noalloc[dev];
  if (!d.ptr) {
    void *p = nullptr;
    ACPP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(ACPP_ERROR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    int cus = 0;
    ACPP_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    d.compute_units = cus;
    d.rs.mode = region_state::graph;
    d.ptr = p;
  }
  *out = &d;
  return ACPP_SUCCESS;
}

}  // namespace

extern "C" {

acpp_status acpp_inclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, acpp_algorithm_op op, const int32_t *init, const acpp_launch_config *config) {
  return scan_entry<int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, int32_t(0), config, "acpp_inclusive_scan_i32");
}
acpp_status acpp_exclusive_scan_i32(acpp_allocation_group *scratch, acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<int32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_i32");
}

acpp_status acpp_inclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, acpp_algorithm_op op, const uint32_t *init, const acpp_launch_config *config) {
  return scan_entry<uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, false, init, uint32_t(0), config, "acpp_inclusive_scan_u32");
}
acpp_status acpp_exclusive_scan_u32(acpp_allocation_group *scratch, acpp_stream_t stream, const uint32_t *first, uint64_t n, uint32_t *d_first, uint32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  return scan_entry<uint32_t>(scratch, static_cast<hipStream_t>(stream), first, n, d_first, op, true, nullptr, init, config, "acpp_exclusive_scan_u32");
}

acpp_status acpp_exclusive_scan_i32_noalloc(acpp_stream_t stream, const int32_t *first, uint64_t n, int32_t *d_first, int32_t init, acpp_algorithm_op op, const acpp_launch_config *config) {
  static const char *name = "acpp_exclusive_scan_i32_noalloc";
  if (n > 0 && (!first || !d_first))
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": first or d_first is NULL");
  int dev = 0;
  ACPP_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= noalloc_max_devices)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, std::string(name) + ": device id out of range");
  std::lock_guard<std::mutex> lock{g_noalloc_mutex};
  noalloc_device *d = nullptr;
  acpp_status st = noalloc_setup(dev, &d);
  if (st != ACPP_SUCCESS) return st;
  const fixed_region region{d->ptr, noalloc_bytes, &d->rs, host_trip_flag_if_allocated(dev), d->compute_units};
  return with_op<int32_t, int32_t>(op, [&](auto opv) -> acpp_status {
    using Op = decltype(opv);
    if (n == 0) return clear_error();  // decoupled_lookback_scan.hpp:565-566
    acpp_status s = launch<int32_t, Op>(nullptr, static_cast<hipStream_t>(stream), first, n, d_first, true,
                                        true, init, config, name, 0, &region);
    return s == ACPP_SUCCESS ? clear_error() : s;
  });
}

}  // extern "C"
