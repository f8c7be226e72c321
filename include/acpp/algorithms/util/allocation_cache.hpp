// Scratch memory for the algorithms: allocation_cache / allocation_group.
//
// Same classes, members and rules as
// include/hipSYCL/algorithms/util/allocation_cache.hpp:30-181 (contract in
// doc/algorithms.md:48-120): the cache is a thread-safe best-fit pool that
// frees its memory on purge() or destruction; a group (not thread-safe)
// hands out allocations and returns them to the cache on release() or
// destruction.  The pool itself lives in the library (C-ABI
// acpp_allocation_cache_* / acpp_allocation_group_*), so the same cache can
// be shared with callers of the C-ABI.
#pragma once

#include <cstddef>

#include "acpp/queue.hpp"
#include "acpp_mi355x.h"

namespace acpp {
namespace algorithms {
namespace util {

enum class allocation_type { device = ACPP_ALLOCATION_DEVICE, shared = ACPP_ALLOCATION_SHARED, host = ACPP_ALLOCATION_HOST };

class allocation_cache {
 public:
  explicit allocation_cache(allocation_type t) {
    acpp::detail::check_status(acpp_allocation_cache_create(static_cast<int>(t), &cache_));
  }
  allocation_cache(const allocation_cache &) = delete;
  allocation_cache &operator=(const allocation_cache &) = delete;
  ~allocation_cache() {
    if (cache_) acpp_allocation_cache_destroy(cache_);
  }

  void purge() { acpp::detail::check_status(acpp_allocation_cache_purge(cache_)); }

  acpp_allocation_cache *native() const { return cache_; }

 private:
  acpp_allocation_cache *cache_ = nullptr;
};

class allocation_group {
 public:
  allocation_group(allocation_cache *parent, int device) {
    acpp::detail::check_status(acpp_allocation_group_create(parent->native(), device, &group_));
  }
  allocation_group(allocation_cache *parent, const queue &q) : allocation_group(parent, q.get_device()) {}
  allocation_group(const allocation_group &) = delete;
  allocation_group &operator=(const allocation_group &) = delete;
  ~allocation_group() {
    if (group_) acpp_allocation_group_destroy(group_);
  }

  void release() { acpp::detail::check_status(acpp_allocation_group_release(group_)); }

  template <class T>
  T *obtain(std::size_t count) {
    void *p = nullptr;
    acpp::detail::check_status(acpp_allocation_group_obtain(group_, count * sizeof(T), alignof(T), &p));
    return static_cast<T *>(p);
  }

  acpp_allocation_group *native() const { return group_; }

 private:
  acpp_allocation_group *group_ = nullptr;
};

}  // namespace util
}  // namespace algorithms
}  // namespace acpp
