// stdpar entry points (include/acpp/stdpar/numeric.hpp) on std::vector data
// in pinned host memory, checked against the sequential C++ standard
// library, with the cases of the reference's pstl tests -- once with every
// call offloaded, once with every call on the host, once with the learned
// choice, so that both branches must agree bit for bit on integers:
//   tests/pstl/reduce.cpp:24-66 (sizes 0, 1, 127, 1000, 1000*1000; init 10 / 0 / 0ll),
//   tests/pstl/transform_reduce.cpp, inclusive_scan.cpp, exclusive_scan.cpp,
//   transform_inclusive_scan.cpp, transform_exclusive_scan.cpp.
// Exit status 0 = all passed.
#include <cstdio>
#include <numeric>
#include <vector>

#include "acpp/stdpar/numeric.hpp"

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
template <class T>
using pvec = std::vector<T, pinned_allocator<T>>;

static int failures = 0;
#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                \
      ++failures;                        \
    }                                    \
  } while (0)

template <class T>
void test_basic_reduction(T init, std::size_t size) {
  pvec<T> data(size);
  for (std::size_t i = 0; i < data.size(); ++i) data[i] = static_cast<T>(i);
  T ref = std::reduce(data.begin(), data.end(), init, std::plus<>{});
  T res = sp::reduce(sp::par_unseq, data.begin(), data.end(), init, std::plus<>{});
  CHECK(res == ref, "reduce size %zu", size);
  T res2 = sp::reduce(sp::par_unseq, data.begin(), data.end(), init);
  CHECK(res2 == res, "reduce(init) size %zu", size);
  T ref3 = std::reduce(data.begin(), data.end());
  T res3 = sp::reduce(sp::par_unseq, data.begin(), data.end());
  CHECK(res3 == ref3, "reduce() size %zu", size);
}

template <class T>
void test_transform_reduce(std::size_t size) {
  pvec<T> a(size), b(size);
  for (std::size_t i = 0; i < size; ++i) {
    a[i] = static_cast<T>(i % 13);
    b[i] = static_cast<T>(i % 7);
  }
  T ref = std::transform_reduce(a.begin(), a.end(), b.begin(), T(3));
  T res = sp::transform_reduce(sp::par_unseq, a.begin(), a.end(), b.begin(), T(3));
  CHECK(res == ref, "transform_reduce size %zu", size);
  T res2 = sp::transform_reduce(sp::par_unseq, a.begin(), a.end(), b.begin(), T(3), std::plus<T>{},
                                std::multiplies<T>{});
  CHECK(res2 == ref, "transform_reduce(ops) size %zu", size);
}

template <class T>
void test_scans(std::size_t size) {
  pvec<T> x(size), y(size), z(size);
  for (std::size_t i = 0; i < size; ++i) x[i] = static_cast<T>((i * 7) % 19) - 9;
  std::vector<T> ry(size), rz(size);
  std::inclusive_scan(x.begin(), x.end(), ry.begin());
  auto e1 = sp::inclusive_scan(sp::par_unseq, x.begin(), x.end(), y.begin());
  sp::sync();  // offloaded scans return without waiting
  CHECK(e1 == y.end(), "inclusive_scan end size %zu", size);
  CHECK(std::equal(y.begin(), y.end(), ry.begin()), "inclusive_scan size %zu", size);
  std::inclusive_scan(x.begin(), x.end(), ry.begin(), std::plus<>{}, T(5));
  sp::inclusive_scan(sp::par_unseq, x.begin(), x.end(), y.begin(), std::plus<>{}, T(5));
  sp::sync();
  CHECK(std::equal(y.begin(), y.end(), ry.begin()), "inclusive_scan(init) size %zu", size);
  std::exclusive_scan(x.begin(), x.end(), rz.begin(), T(11));
  auto e2 = sp::exclusive_scan(sp::par_unseq, x.begin(), x.end(), z.begin(), T(11));
  sp::sync();
  CHECK(e2 == z.end(), "exclusive_scan end size %zu", size);
  CHECK(std::equal(z.begin(), z.end(), rz.begin()), "exclusive_scan size %zu", size);
  std::exclusive_scan(x.begin(), x.end(), rz.begin(), T(0), acpp::maximum<T>{});
  sp::exclusive_scan(sp::par_unseq, x.begin(), x.end(), z.begin(), T(0), acpp::maximum<T>{});
  sp::sync();
  CHECK(std::equal(z.begin(), z.end(), rz.begin()), "exclusive_scan(max) size %zu", size);
}

// tests/pstl/transform_inclusive_scan.cpp:73-100 and
// transform_exclusive_scan.cpp:73-93: data[i] = i, unary_op = op(x, x),
// (plus, init 3) and (multiplies, init 3ull); here unary_op is
// acpp::self_op<Op>, the same function as a named type.
template <class T, class Op>
void test_transform_scans(T init, Op op, std::size_t size) {
  pvec<T> data(size), d0(size), d1(size), d2(size);
  for (std::size_t i = 0; i < size; ++i) data[i] = static_cast<T>(i);
  acpp::self_op<Op> unary{op};
  std::vector<T> r0(size), r1(size), r2(size);
  std::transform_inclusive_scan(data.begin(), data.end(), r0.begin(), op, unary);
  std::transform_inclusive_scan(data.begin(), data.end(), r1.begin(), op, unary, init);
  std::transform_exclusive_scan(data.begin(), data.end(), r2.begin(), init, op, unary);
  CHECK(sp::transform_inclusive_scan(sp::par_unseq, data.begin(), data.end(), d0.begin(), op, unary) ==
            d0.end(),
        "transform_inclusive_scan end size %zu", size);
  CHECK(sp::transform_inclusive_scan(sp::par_unseq, data.begin(), data.end(), d1.begin(), op, unary,
                                     init) == d1.end(),
        "transform_inclusive_scan(init) end size %zu", size);
  CHECK(sp::transform_exclusive_scan(sp::par_unseq, data.begin(), data.end(), d2.begin(), init, op,
                                     unary) == d2.end(),
        "transform_exclusive_scan end size %zu", size);
  sp::sync();
  CHECK(std::equal(d0.begin(), d0.end(), r0.begin()), "transform_inclusive_scan size %zu", size);
  CHECK(std::equal(d1.begin(), d1.end(), r1.begin()), "transform_inclusive_scan(init) size %zu", size);
  CHECK(std::equal(d2.begin(), d2.end(), r2.begin()), "transform_exclusive_scan size %zu", size);
}

// tests/pstl/transform_reduce.cpp:23-35: identity transform, plus, init 10 / 0 / 0ll;
// plus a negate / square transform on the same data.
template <class T>
void test_unary_transform_reduce(T init, std::size_t size) {
  pvec<T> data(size);
  for (std::size_t i = 0; i < size; ++i) data[i] = static_cast<T>(i);
  T ref = std::transform_reduce(data.begin(), data.end(), init, std::plus<>{}, [](auto x) { return x; });
  T res = sp::transform_reduce(sp::par_unseq, data.begin(), data.end(), init, std::plus<>{}, acpp::identity<>{});
  CHECK(res == ref, "transform_reduce(identity) size %zu", size);
  ref = std::transform_reduce(data.begin(), data.end(), init, std::plus<>{}, std::negate<T>{});
  res = sp::transform_reduce(sp::par_unseq, data.begin(), data.end(), init, std::plus<>{}, std::negate<T>{});
  CHECK(res == ref, "transform_reduce(negate) size %zu", size);
  ref = std::transform_reduce(data.begin(), data.end(), init, acpp::maximum<T>{}, acpp::square<T>{});
  res = sp::transform_reduce(sp::par_unseq, data.begin(), data.end(), init, acpp::maximum<T>{},
                             acpp::square<T>{});
  CHECK(res == ref, "transform_reduce(max, square) size %zu", size);
}

static void run_battery(const char *label) {
  test_basic_reduction(10, 0);
  test_basic_reduction(10, 1);
  test_basic_reduction(10, 127);
  test_basic_reduction(0, 1000);
  test_basic_reduction(0ll, 1000 * 1000);
  for (std::size_t n : {std::size_t(0), std::size_t(1), std::size_t(127), std::size_t(1000), std::size_t(1) << 20}) {
    test_transform_reduce<int>(n);
    test_transform_reduce<long long>(n);
    test_scans<int>(n);
    test_scans<long long>(n);
  }
  test_unary_transform_reduce(10, 0);
  test_unary_transform_reduce(10, 1);
  test_unary_transform_reduce(10, 127);
  test_unary_transform_reduce(0, 1000);
  test_unary_transform_reduce(0ll, 1000 * 1000);
  for (std::size_t n : {std::size_t(0), std::size_t(1), std::size_t(127), std::size_t(1000), std::size_t(1024) * 1024}) {
    test_transform_scans(3u, std::plus<>{}, n);
    test_transform_scans(3ull, std::multiplies<>{}, n);
    test_transform_scans(static_cast<std::size_t>(3), std::plus<>{}, n);
  }
  std::printf("%s: battery done (%d failures so far)\n", label, failures);
}

// The decision itself: pageable memory stays on the host, device memory is
// offloaded, pinned memory follows the learned crossover size.
static void test_decision() {
  const int dev = 0;
  int d = -1;
  std::vector<int> pageable(1 << 20, 1);
  const void *pp[] = {pageable.data()};
  acpp::detail::check_status(acpp_stdpar_should_offload(dev, ACPP_STDPAR_OP_REDUCE, 4, 0, pageable.size(), pp, 1, &d));
  CHECK(d == 0, "pageable memory stays on the host");
  CHECK(sp::reduce(sp::par_unseq, pageable.begin(), pageable.end(), 0) == (1 << 20), "reduce of pageable memory");
  int *dp = nullptr;
  acpp::detail::check_hip(hipMalloc(&dp, 64), "hipMalloc");
  const void *dpp[] = {dp};
  acpp::detail::check_status(acpp_stdpar_should_offload(dev, ACPP_STDPAR_OP_REDUCE, 4, 0, 16, dpp, 1, &d));
  CHECK(d == 1, "device memory is offloaded");
  const void *mixed[] = {dp, pageable.data()};
  CHECK(acpp_stdpar_should_offload(dev, ACPP_STDPAR_OP_REDUCE, 4, 0, 16, mixed, 2, &d) == ACPP_ERROR_INVALID_ARGUMENT,
        "device-only and host-only memory in one call is rejected");
  (void)hipFree(dp);
  pvec<int> pinned(16);
  const void *pin[] = {pinned.data()};
  for (int op = 0; op < 5; ++op) {
    uint64_t min_n = 0;
    acpp::detail::check_status(acpp_stdpar_offload_threshold(dev, op, 4, 0, &min_n));
    std::printf("op %d int32: offload from n = %llu\n", op, static_cast<unsigned long long>(min_n));
    if (min_n > 1 && min_n != UINT64_MAX) {
      acpp::detail::check_status(acpp_stdpar_should_offload(dev, op, 4, 0, min_n - 1, pin, 1, &d));
      CHECK(d == 0, "op %d below the crossover stays on the host", op);
      acpp::detail::check_status(acpp_stdpar_should_offload(dev, op, 4, 0, min_n, pin, 1, &d));
      CHECK(d == 1, "op %d at the crossover is offloaded", op);
    }
  }
}

int main() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    std::printf("no GPU\n");
    return 2;
  }
  sp::set_offload_policy(sp::offload_policy::always);
  run_battery("always offload");
  sp::set_offload_policy(sp::offload_policy::never);
  run_battery("host only");
  sp::set_offload_policy(sp::offload_policy::automatic);
  run_battery("learned choice");
  test_decision();
  std::printf(failures ? "FAILED (%d)\n" : "ALL PASSED\n", failures);
  return failures ? 1 : 0;
}
