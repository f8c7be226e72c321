// Host runtime of the library: error state, device queries and the scratch
// allocation cache.
//
// The cache mirrors the reference's util::allocation_cache /
// allocation_group (include/hipSYCL/algorithms/util/allocation_cache.hpp:42-181):
// a thread-safe best-fit pool per device, and a non-thread-safe group that
// returns everything it obtained to the pool when released.  It adds a second
// pool of "sync regions" (zeroed once, kept reusable by the kernels), which
// is what lets the scan and the reduction run as single launches.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace acpp_impl {

namespace {
thread_local std::string g_last_error;
}

acpp_status set_error(acpp_status code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

acpp_status clear_error() {
  g_last_error.clear();
  return ACPP_SUCCESS;
}

acpp_status hip_error(hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorName(e) + " (" +
                    hipGetErrorString(e) + ")";
  // The error is reported through the returned status; clear HIP's
  // per-thread "last error" so that the caller's next unrelated HIP call
  // does not inherit it.
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory) return set_error(ACPP_ERROR_OUT_OF_MEMORY, msg);
  return set_error(ACPP_ERROR_HIP, msg);
}

acpp_status check_launch(const char *kernel_name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, kernel_name);
  return ACPP_SUCCESS;
}

namespace {
struct device_info {
  int compute_units = 0;
};
std::mutex g_dev_mutex;
std::vector<device_info> g_devices;

acpp_status query_compute_units(int dev, int *out) {
  std::lock_guard<std::mutex> lock{g_dev_mutex};
  if (dev < 0) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "negative device id");
  if (static_cast<size_t>(dev) >= g_devices.size()) g_devices.resize(dev + 1);
  if (g_devices[dev].compute_units == 0) {
    int cu = 0;
    ACPP_CHECK_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    g_devices[dev].compute_units = cu;
  }
  *out = g_devices[dev].compute_units;
  return ACPP_SUCCESS;
}
// One pinned host word per device, raised by report_trip on the device.
// It is allocated once per process (mapped into every device's address
// space) and never freed.
constexpr int max_devices = 64;
std::mutex g_trip_mutex;
uint32_t *g_trip_flags = nullptr;
}  // namespace

uint32_t *host_trip_flag(int device) {
  if (device < 0 || device >= max_devices) return nullptr;
  std::lock_guard<std::mutex> lock{g_trip_mutex};
  if (!g_trip_flags) {
    void *p = nullptr;
    if (hipHostMalloc(&p, max_devices * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocPortable) !=
        hipSuccess) {
      (void)hipGetLastError();
      return nullptr;  // kernels then only count trips in their region
    }
    memset(p, 0, max_devices * sizeof(uint32_t));
    g_trip_flags = static_cast<uint32_t *>(p);
  }
  return g_trip_flags + device;
}

uint32_t *host_trip_flag_if_allocated(int device) {
  if (device < 0 || device >= max_devices) return nullptr;
  std::lock_guard<std::mutex> lock{g_trip_mutex};
  return g_trip_flags ? g_trip_flags + device : nullptr;
}

namespace {
// Take (read and clear) the trip flag of `device`.
bool take_trip_flag(int device) {
  if (device < 0 || device >= max_devices) return false;
  uint32_t *flags = nullptr;
  {
    std::lock_guard<std::mutex> lock{g_trip_mutex};
    flags = g_trip_flags;
  }
  if (!flags) return false;
  return __atomic_exchange_n(flags + device, 0u, __ATOMIC_ACQ_REL) != 0;
}

// Peek at the trip flag of `device` without clearing it.
bool trip_flag_raised(int device) {
  if (device < 0 || device >= max_devices) return false;
  std::lock_guard<std::mutex> lock{g_trip_mutex};
  return g_trip_flags && __atomic_load_n(g_trip_flags + device, __ATOMIC_ACQUIRE) != 0;
}

acpp_status report_trips(int device, const char *where, bool clear) {
  if (!(clear ? take_trip_flag(device) : trip_flag_raised(device))) return ACPP_SUCCESS;
  return set_error(ACPP_ERROR_WATCHDOG,
                   std::string(where) + ": a kernel on device " + std::to_string(device) +
                       " gave up waiting for another work-group (watchdog); results written "
                       "since the last check may be wrong");
}
}  // namespace

}  // namespace acpp_impl

using namespace acpp_impl;

namespace {
struct allocation {
  void *ptr = nullptr;
  size_t size = 0;
  int dev = 0;
  bool sync = false;
  std::shared_ptr<region_state> state;  // sync regions only
};
}  // namespace

struct acpp_allocation_cache {
  int type = ACPP_ALLOCATION_DEVICE;
  std::mutex mutex;
  std::vector<allocation> available;

  acpp_status allocate(size_t bytes, int dev, allocation &out) {
    int prev = 0;
    ACPP_CHECK_HIP(hipGetDevice(&prev));
    if (prev != dev) ACPP_CHECK_HIP(hipSetDevice(dev));
    void *p = nullptr;
    hipError_t e;
    if (type == ACPP_ALLOCATION_SHARED)
      e = hipMallocManaged(&p, bytes);
    else if (type == ACPP_ALLOCATION_HOST)
      e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    else
      e = hipMalloc(&p, bytes);
    if (prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess) return hip_error(e, "scratch allocation");
    out.ptr = p;
    out.size = bytes;
    out.dev = dev;
    return ACPP_SUCCESS;
  }

  void deallocate(const allocation &a) {
    if (type == ACPP_ALLOCATION_HOST)
      (void)hipHostFree(a.ptr);
    else
      (void)hipFree(a.ptr);
  }

  // Best fit: the smallest available allocation that is large enough and
  // suitably aligned (allocation_cache.hpp:82-114).
  bool find(size_t bytes, size_t alignment, int dev, bool sync, allocation &out) {
    std::lock_guard<std::mutex> lock{mutex};
    size_t best = available.size();
    for (size_t i = 0; i < available.size(); ++i) {
      const allocation &a = available[i];
      if (a.dev != dev || a.sync != sync || a.size < bytes) continue;
      if (reinterpret_cast<uintptr_t>(a.ptr) % alignment != 0) continue;
      if (best == available.size() || a.size < available[best].size) best = i;
    }
    if (best == available.size()) return false;
    out = available[best];
    available.erase(available.begin() + best);
    return true;
  }

  void give_back(const allocation &a) {
    std::lock_guard<std::mutex> lock{mutex};
    available.push_back(a);
  }

  void purge() {
    std::lock_guard<std::mutex> lock{mutex};
    for (const allocation &a : available) deallocate(a);
    available.clear();
  }
};

struct acpp_allocation_group {
  acpp_allocation_cache *parent = nullptr;
  int dev = 0;
  std::vector<allocation> held;
};

namespace {
// Round scratch requests up so that regions are reused across nearby sizes.
size_t round_request(size_t bytes) {
  size_t r = 256;
  while (r < bytes) r <<= 1;
  return r;
}
}  // namespace

namespace acpp_impl {

acpp_status group_obtain(acpp_allocation_group *g, size_t bytes, void **out) {
  allocation a;
  bytes = round_request(bytes);
  if (!g->parent->find(bytes, 256, g->dev, false, a)) {
    acpp_status s = g->parent->allocate(bytes, g->dev, a);
    if (s != ACPP_SUCCESS) return s;
  }
  g->held.push_back(a);
  *out = a.ptr;
  return ACPP_SUCCESS;
}

acpp_status group_obtain_sync(acpp_allocation_group *g, size_t bytes,
                              hipStream_t stream, void **out,
                              size_t *capacity_bytes, region_state **state) {
  if (g->parent->type != ACPP_ALLOCATION_DEVICE && g->parent->type != ACPP_ALLOCATION_SHARED)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT,
                     "kernels need device-accessible scratch; use an "
                     "ACPP_ALLOCATION_DEVICE or _SHARED cache");
  allocation a;
  bytes = round_request(bytes);
  if (g->parent->find(bytes, 256, g->dev, true, a)) {
    if (a.state && a.state->captured) {
      // Back from a group that held it for a graph: that graph is gone, so
      // one clear before the next eager launch suffices.
      a.state->captured = false;
      a.state->mode = region_state::graph;
    }
  } else {
    acpp_status s = g->parent->allocate(bytes, g->dev, a);
    if (s != ACPP_SUCCESS) return s;
    a.sync = true;
    a.state = std::make_shared<region_state>();
    if (stream_is_capturing(stream)) {
      // The captured launch clears the part it uses (begin_region_launch
      // under capture), and the first eager launch clears all of it: a
      // clear here would be a second one in every replay of the graph.
      a.state->mode = region_state::graph;
    } else {
      s = reset_sync_region(a.ptr, a.size, g->dev, stream);
      if (s != ACPP_SUCCESS) {
        g->parent->deallocate(a);
        return s;
      }
    }
  }
  g->held.push_back(a);
  *out = a.ptr;
  *capacity_bytes = a.size;
  if (state) *state = a.state.get();
  return ACPP_SUCCESS;
}

acpp_status begin_region_launch(acpp_allocation_group *g, void *region, size_t used, size_t capacity,
                                region_state *rs, hipStream_t stream, region_launch *out) {
  return begin_region_launch_flag(host_trip_flag(g->dev), region, used, capacity, rs, stream, out);
}

acpp_status begin_region_launch_flag(uint32_t *trip_flag, void *region, size_t used, size_t capacity,
                                     region_state *rs, hipStream_t stream, region_launch *out) {
  region_launch l;
  l.capturing = stream_is_capturing(stream);
  if (l.capturing) {
    acpp_status st = reset_sync_region_flag(region, used, trip_flag, stream);
    if (st != ACPP_SUCCESS) return st;
    rs->mode = region_state::graph;
    rs->captured = true;
  } else {
    if (rs->mode == region_state::graph || rs->captured || rs->epoch >= max_epoch) {
      acpp_status st = reset_sync_region_flag(region, capacity, trip_flag, stream);
      if (st != ACPP_SUCCESS) return st;
      rs->epoch = 0;
      rs->ticket_base = 0;
    }
    rs->mode = region_state::host_managed;
    l.epoch = rs->epoch + 1;
    l.ticket_base = rs->ticket_base;
  }
  *out = l;
  return ACPP_SUCCESS;
}

void end_region_launch(region_state *rs, const region_launch &l, uint32_t tickets, acpp_status launch_status) {
  if (launch_status != ACPP_SUCCESS) {
    rs->mode = region_state::graph;
  } else if (!l.capturing) {
    rs->epoch = l.epoch;
    rs->ticket_base = l.ticket_base + tickets;
  }
}

int group_device(const acpp_allocation_group *g) { return g->dev; }

acpp_status enter_device(acpp_allocation_group *g, int *compute_units) {
  int cur = 0;
  ACPP_CHECK_HIP(hipGetDevice(&cur));
  if (cur != g->dev) ACPP_CHECK_HIP(hipSetDevice(g->dev));
  return query_compute_units(g->dev, compute_units);
}

}  // namespace acpp_impl

extern "C" {

const char *acpp_last_error_message(void) { return g_last_error.c_str(); }

const char *acpp_version(void) { return "0.1.0"; }

acpp_status acpp_device_compute_units(int device, int *out) {
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "out is NULL");
  return query_compute_units(device, out);
}

acpp_status acpp_allocation_cache_create(int allocation_type, acpp_allocation_cache **out) {
  if (!out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "out is NULL");
  if (allocation_type < ACPP_ALLOCATION_DEVICE || allocation_type > ACPP_ALLOCATION_HOST)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "unknown allocation type");
  auto *c = new (std::nothrow) acpp_allocation_cache;
  if (!c) return set_error(ACPP_ERROR_OUT_OF_MEMORY, "host allocation failed");
  c->type = allocation_type;
  *out = c;
  return clear_error();
}

acpp_status acpp_allocation_cache_purge(acpp_allocation_cache *cache) {
  if (!cache) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "cache is NULL");
  cache->purge();
  return clear_error();
}

acpp_status acpp_allocation_cache_destroy(acpp_allocation_cache *cache) {
  if (!cache) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "cache is NULL");
  cache->purge();
  delete cache;
  return clear_error();
}

acpp_status acpp_allocation_group_create(acpp_allocation_cache *cache, int device,
                                         acpp_allocation_group **out) {
  if (!cache || !out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "cache or out is NULL");
  if (device < 0) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "negative device id");
  auto *g = new (std::nothrow) acpp_allocation_group;
  if (!g) return set_error(ACPP_ERROR_OUT_OF_MEMORY, "host allocation failed");
  g->parent = cache;
  g->dev = device;
  *out = g;
  return clear_error();
}

acpp_status acpp_allocation_group_obtain(acpp_allocation_group *group, size_t bytes,
                                         size_t alignment, void **out) {
  if (!group || !out) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "group or out is NULL");
  if (alignment == 0 || (alignment & (alignment - 1)) != 0 || alignment > 256)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "alignment must be a power of two <= 256");
  allocation a;
  size_t req = round_request(bytes);
  if (!group->parent->find(req, alignment, group->dev, false, a)) {
    acpp_status s = group->parent->allocate(req, group->dev, a);
    if (s != ACPP_SUCCESS) return s;
  }
  group->held.push_back(a);
  *out = a.ptr;
  return clear_error();
}

acpp_status acpp_allocation_group_release(acpp_allocation_group *group) {
  if (!group) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "group is NULL");
  for (const allocation &a : group->held) group->parent->give_back(a);
  group->held.clear();
  acpp_status s = report_trips(group->dev, "acpp_allocation_group_release", false);
  return s == ACPP_SUCCESS ? clear_error() : s;
}

acpp_status acpp_allocation_group_destroy(acpp_allocation_group *group) {
  if (!group) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "group is NULL");
  for (const allocation &a : group->held) group->parent->give_back(a);
  group->held.clear();
  const int dev = group->dev;
  delete group;
  acpp_status s = report_trips(dev, "acpp_allocation_group_destroy", false);
  return s == ACPP_SUCCESS ? clear_error() : s;
}

acpp_status acpp_check_asynchronous_errors(int device) {
  if (device < 0) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "negative device id");
  acpp_status s = report_trips(device, "acpp_check_asynchronous_errors", true);
  return s == ACPP_SUCCESS ? clear_error() : s;
}

acpp_status acpp_allocation_group_watchdog_trips(acpp_allocation_group *group, uint64_t *trips) {
  if (!group || !trips) return set_error(ACPP_ERROR_INVALID_ARGUMENT, "group or trips is NULL");
  uint64_t total = 0;
  for (const allocation &a : group->held) {
    if (!a.sync) continue;
    uint32_t v = 0;
    const char *src = static_cast<const char *>(a.ptr) + offsetof(sync_header, watchdog_trips);
    ACPP_CHECK_HIP(hipMemcpy(&v, src, sizeof(v), hipMemcpyDeviceToHost));
    total += v;
  }
  *trips = total;
  return clear_error();
}

}  // extern "C"
