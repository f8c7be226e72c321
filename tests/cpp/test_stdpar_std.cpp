// std::reduce / transform_reduce / *_scan called with std::execution::par_unseq
// themselves, compiled with include/acpp/stdpar/std before the system
// headers (as the reference's stdpar compiler driver puts its own <numeric>
// first): the calls must go through acpp::stdpar, checked here by the
// offload policy (every call offloaded) and against the sequential standard
// library.  Exit status 0 = all passed.
#include <cstdio>
#include <execution>
#include <numeric>
#include <vector>

namespace sp = acpp::stdpar;

template <class T>
struct pinned_allocator {
  using value_type = T;
  pinned_allocator() = default;
  template <class U>
  pinned_allocator(const pinned_allocator<U> &) {}
  T *allocate(std::size_t n) {
    void *p = nullptr;
    acpp::detail::check_hip(hipHostMalloc(&p, (n ? n : 1) * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<T *>(p);
  }
  void deallocate(T *p, std::size_t) { (void)hipHostFree(p); }
  template <class U>
  bool operator==(const pinned_allocator<U> &) const { return true; }
  template <class U>
  bool operator!=(const pinned_allocator<U> &) const { return false; }
};

int main() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    std::printf("no GPU\n");
    return 2;
  }
  int failures = 0;
  const std::size_t n = 1 << 20;
  std::vector<long long, pinned_allocator<long long>> v(n), w(n), d(n);
  for (std::size_t i = 0; i < n; ++i) {
    v[i] = static_cast<long long>(i % 1000) - 500;
    w[i] = static_cast<long long>(i % 7);
  }
  for (auto policy : {sp::offload_policy::always, sp::offload_policy::never}) {
    sp::set_offload_policy(policy);
    const std::size_t before = sp::offloaded_calls();
    std::vector<long long> r(n);
    failures += std::reduce(std::execution::par_unseq, v.begin(), v.end()) != std::reduce(v.begin(), v.end());
    failures += std::reduce(std::execution::par_unseq, v.begin(), v.end(), 7ll) != std::reduce(v.begin(), v.end(), 7ll);
    failures += std::transform_reduce(std::execution::par_unseq, v.begin(), v.end(), w.begin(), 1ll) !=
                std::transform_reduce(v.begin(), v.end(), w.begin(), 1ll);
    std::inclusive_scan(std::execution::par_unseq, v.begin(), v.end(), d.begin());
    sp::sync();
    std::inclusive_scan(v.begin(), v.end(), r.begin());
    failures += !std::equal(d.begin(), d.end(), r.begin());
    std::exclusive_scan(std::execution::par_unseq, v.begin(), v.end(), d.begin(), 3ll, acpp::maximum<long long>{});
    sp::sync();
    std::exclusive_scan(v.begin(), v.end(), r.begin(), 3ll, acpp::maximum<long long>{});
    failures += !std::equal(d.begin(), d.end(), r.begin());
    // the five std:: calls went through acpp::stdpar: all on the GPU, or none
    const std::size_t ran = sp::offloaded_calls() - before;
    if (ran != (policy == sp::offload_policy::always ? 5u : 0u)) {
      std::printf("FAIL: %zu of the std:: calls were offloaded\n", ran);
      ++failures;
    }
  }
  // std::execution::seq is untouched: the standard library runs it
  failures += std::reduce(std::execution::seq, v.begin(), v.end()) != std::reduce(v.begin(), v.end());
  sp::set_offload_policy(sp::offload_policy::automatic);
  std::printf(failures ? "FAILED (%d)\n" : "ALL PASSED\n", failures);
  return failures ? 1 : 0;
}
