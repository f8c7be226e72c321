// Minimal in-order queue and event for the algorithms facade.
//
// The reference's algorithms take a sycl::queue and return a sycl::event
// (include/hipSYCL/algorithms/numeric.hpp:170-327, doc/algorithms.md:31-40).
// On MI355X the natural equivalents are a HIP stream (in-order) and a HIP
// event recorded on it:
//   acpp::queue  -- owns (or wraps) one hipStream_t on one device;
//   acpp::event  -- shared handle to a hipEvent_t; a default-constructed
//                   event is complete (the reference returns one for n == 0,
//                   doc/algorithms.md:37);
//   acpp::exception -- thrown for errors.  Argument and launch errors are
//                   detected when a call is enqueued, so the facade throws
//                   at the call.  Errors that arise while kernels run (a
//                   wait for another work-group that gave up, the library's
//                   watchdog) are asynchronous, as in the reference
//                   (queue::wait_and_throw / throw_asynchronous,
//                   include/hipSYCL/sycl/queue.hpp:378-383): queue::wait()
//                   and event::wait() throw them (code ACPP_ERROR_WATCHDOG)
//                   once the work has completed.
#pragma once

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "acpp_mi355x.h"

namespace acpp {

class exception : public std::runtime_error {
 public:
  exception(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const noexcept { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw exception(ACPP_ERROR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
inline void check_status(acpp_status s) {
  if (s != ACPP_SUCCESS) throw exception(s, acpp_last_error_message());
}
}  // namespace detail

class event {
 public:
  event() = default;  // complete
  event(std::shared_ptr<hipEvent_t> e, int device) : ev_(std::move(e)), device_(device) {}

  // Waits, then throws the device's asynchronous errors.
  void wait() const {
    if (!ev_) return;
    detail::check_hip(hipEventSynchronize(*ev_), "hipEventSynchronize");
    detail::check_status(acpp_check_asynchronous_errors(device_));
  }
  bool is_complete() const {
    if (!ev_) return true;
    return hipEventQuery(*ev_) == hipSuccess;
  }
  // hipEvent_t, or nullptr for a complete default event.
  hipEvent_t native() const { return ev_ ? *ev_ : nullptr; }

 private:
  std::shared_ptr<hipEvent_t> ev_;
  int device_ = 0;
};

class queue {
 public:
  // A new in-order stream on `device` (default priority).
  explicit queue(int device = 0) : device_(device), owned_(true) {
    detail::check_hip(hipSetDevice(device), "hipSetDevice");
    detail::check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  }
  // Wrap an existing stream (not destroyed by the queue).
  queue(int device, hipStream_t stream) : device_(device), stream_(stream), owned_(false) {}

  queue(const queue &) = delete;
  queue &operator=(const queue &) = delete;
  ~queue() {
    if (owned_ && stream_) {
      (void)hipStreamSynchronize(stream_);
      (void)hipStreamDestroy(stream_);
    }
  }

  int get_device() const { return device_; }
  hipStream_t get_native() const { return stream_; }

  // Waits for everything enqueued so far, then throws the device's
  // asynchronous errors (queue::wait_and_throw).
  void wait() {
    detail::check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    detail::check_status(acpp_check_asynchronous_errors(device_));
  }

  // Make later work on this queue wait for `deps`.
  void depends_on(const std::vector<event> &deps) {
    for (const event &e : deps)
      if (e.native())
        detail::check_hip(hipStreamWaitEvent(stream_, e.native(), 0), "hipStreamWaitEvent");
  }

  // An event that completes when everything enqueued so far has finished.
  event record() {
    auto *raw = new hipEvent_t(nullptr);
    hipError_t e = hipEventCreateWithFlags(raw, hipEventDisableTiming);
    if (e != hipSuccess) {
      delete raw;
      detail::check_hip(e, "hipEventCreate");
    }
    std::shared_ptr<hipEvent_t> handle(raw, [](hipEvent_t *p) {
      (void)hipEventDestroy(*p);
      delete p;
    });
    detail::check_hip(hipEventRecord(*handle, stream_), "hipEventRecord");
    return event(std::move(handle), device_);
  }

 private:
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  bool owned_ = false;
};

}  // namespace acpp
