// Baseline for the segmented forms: one plain call per segment, issued from
// native code, the way a caller without the segmented entry points would do
// it.  Each call is an algorithms::inclusive_scan / reduce
// (include/hipSYCL/algorithms/numeric.hpp:226-279) on one segment.  Scratch
// goes back to the cache after every call: all calls are ordered on one
// stream, which is the reference's rule for safe reuse
// (include/hipSYCL/algorithms/util/allocation_cache.hpp:129-141).
#include <stdint.h>

#include "acpp_mi355x.h"

extern "C" {

int acpp_bench_per_segment_inclusive_scan_i64(acpp_allocation_cache *cache, int device,
                                              acpp_stream_t stream, const int64_t *first,
                                              int64_t *d_first, const uint64_t *offsets_host,
                                              uint64_t num_segments) {
  acpp_allocation_group *g = nullptr;
  int st = acpp_allocation_group_create(cache, device, &g);
  if (st != ACPP_SUCCESS) return st;
  for (uint64_t s = 0; s < num_segments && st == ACPP_SUCCESS; ++s) {
    const uint64_t b = offsets_host[s], e = offsets_host[s + 1];
    if (e > b) st = acpp_inclusive_scan_i64(g, stream, first + b, e - b, d_first + b, ACPP_OP_PLUS, nullptr, nullptr);
    if (st == ACPP_SUCCESS) st = acpp_allocation_group_release(g);
  }
  acpp_allocation_group_destroy(g);
  return st;
}

int acpp_bench_per_segment_reduce_f32(acpp_allocation_cache *cache, int device, acpp_stream_t stream,
                                      const float *first, const uint64_t *offsets_host,
                                      uint64_t num_segments, float *out, float init) {
  acpp_allocation_group *g = nullptr;
  int st = acpp_allocation_group_create(cache, device, &g);
  if (st != ACPP_SUCCESS) return st;
  for (uint64_t s = 0; s < num_segments && st == ACPP_SUCCESS; ++s) {
    const uint64_t b = offsets_host[s], e = offsets_host[s + 1];
    // algorithms::reduce leaves *out untouched for an empty range
    // (numeric.hpp:165-169); the benchmark layouts have no empty segments.
    if (e > b) st = acpp_reduce_f32(g, stream, first + b, e - b, out + s, init, ACPP_OP_PLUS, nullptr);
    if (st == ACPP_SUCCESS) st = acpp_allocation_group_release(g);
  }
  acpp_allocation_group_destroy(g);
  return st;
}

}  // extern "C"
