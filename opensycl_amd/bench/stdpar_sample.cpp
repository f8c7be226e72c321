// Run-time samples for the stdpar offload model: one stdpar call
// (acpp/stdpar/numeric.hpp) on pinned host memory, timed once with the call
// kept on the host (the sequential C++ standard library algorithm) and once
// offloaded, wall clock from the call to its completion (scans included:
// sync() is inside the timed region).  opensycl_amd/offload_model.py trains
// the offload model on these samples; the same calls are what the model then
// decides between.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <vector>

#include "acpp/stdpar/numeric.hpp"

namespace {

namespace sp = acpp::stdpar;

template <class T>
struct pinned {
  T *p = nullptr;
  explicit pinned(std::size_t n) {
    acpp::detail::check_hip(hipHostMalloc(reinterpret_cast<void **>(&p), (n ? n : 1) * sizeof(T), hipHostMallocDefault),
                            "hipHostMalloc");
  }
  ~pinned() { (void)hipHostFree(p); }
};

template <class F>
double median_us(F &&f, int reps) {
  std::vector<double> t;
  f();  // warm-up
  for (int r = 0; r < reps; ++r) {
    const auto a = std::chrono::steady_clock::now();
    f();
    const auto b = std::chrono::steady_clock::now();
    t.push_back(std::chrono::duration<double, std::micro>(b - a).count());
  }
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

template <class T>
double run(int op, uint64_t n, int reps, T *x, T *y, T *z) {
  switch (op) {
    case ACPP_STDPAR_OP_REDUCE:
      return median_us([&] { volatile T r = sp::reduce(sp::par_unseq, x, x + n, T(0)); (void)r; }, reps);
    case ACPP_STDPAR_OP_TRANSFORM_REDUCE:
      return median_us([&] {
        volatile T r = sp::transform_reduce(sp::par_unseq, x, x + n, T(0), acpp::plus<T>{}, acpp::square<T>{});
        (void)r;
      }, reps);
    case ACPP_STDPAR_OP_BINARY_TRANSFORM_REDUCE:
      return median_us([&] { volatile T r = sp::transform_reduce(sp::par_unseq, x, x + n, y, T(0)); (void)r; }, reps);
    case ACPP_STDPAR_OP_INCLUSIVE_SCAN:
      return median_us([&] {
        sp::inclusive_scan(sp::par_unseq, x, x + n, z);
        sp::sync();
      }, reps);
    default:
      return median_us([&] {
        sp::exclusive_scan(sp::par_unseq, x, x + n, z, T(0));
        sp::sync();
      }, reps);
  }
}

template <class T>
int sample(int op, uint64_t n, int reps, double *host_us, double *device_us) {
  pinned<T> x(n), y(n), z(n);
  for (uint64_t i = 0; i < n; ++i) {
    x.p[i] = static_cast<T>(static_cast<int64_t>(i % 7) - 3);
    y.p[i] = static_cast<T>(static_cast<int64_t>(i % 5) - 2);
  }
  sp::set_offload_policy(sp::offload_policy::never);
  *host_us = run<T>(op, n, reps, x.p, y.p, z.p);
  sp::set_offload_policy(sp::offload_policy::always);
  *device_us = run<T>(op, n, reps, x.p, y.p, z.p);
  sp::set_offload_policy(sp::offload_policy::automatic);
  return 0;
}

}  // namespace

// dtype: 0 int32, 1 int64, 2 float, 3 double.  Returns 0 on success.
extern "C" int acpp_bench_stdpar_sample(int op, int dtype, uint64_t n, int reps, double *host_us,
                                        double *device_us) {
  try {
    switch (dtype) {
      case 0: return sample<int32_t>(op, n, reps, host_us, device_us);
      case 1: return sample<int64_t>(op, n, reps, host_us, device_us);
      case 2: return sample<float>(op, n, reps, host_us, device_us);
      case 3: return sample<double>(op, n, reps, host_us, device_us);
      default: return -1;
    }
  } catch (const std::exception &) {
    return -2;
  }
}
