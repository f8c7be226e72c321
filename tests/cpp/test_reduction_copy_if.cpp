// C++ facade test: sycl::reduction objects through acpp::parallel_for
// (include/hipSYCL/sycl/libkernel/reduction.hpp:165-236, handler.hpp:799-975;
// the cases of tests/sycl/reduction.cpp), acpp::algorithms::copy_if
// (algorithm.hpp:178-243; the cases of tests/pstl/copy_if.cpp), and the
// asynchronous error path of queue::wait.  Expected values come from the
// C++ standard library's sequential algorithms.
//
// Exit status 0 = every case passed.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

#include "acpp/algorithms/algorithm.hpp"
#include "acpp/reduction.hpp"

namespace alg = acpp::algorithms;

static int failures = 0;

#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
      ++failures;                        \
    }                                    \
  } while (0)

template <class T>
struct device_buffer {
  T *ptr = nullptr;
  size_t n = 0;
  explicit device_buffer(size_t count) : n(count) {
    acpp::detail::check_hip(hipMalloc(&ptr, (count ? count : 1) * sizeof(T)), "hipMalloc");
  }
  ~device_buffer() { (void)hipFree(ptr); }
  void upload(const std::vector<T> &h) {
    acpp::detail::check_hip(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "copy");
  }
  std::vector<T> download() const {
    std::vector<T> h(n);
    acpp::detail::check_hip(hipMemcpy(h.data(), ptr, n * sizeof(T), hipMemcpyDeviceToHost), "copy");
    return h;
  }
  T get(size_t i = 0) const {
    T v;
    acpp::detail::check_hip(hipMemcpy(&v, ptr + i, sizeof(T), hipMemcpyDeviceToHost), "copy");
    return v;
  }
  void set(T v, size_t i = 0) {
    acpp::detail::check_hip(hipMemcpy(ptr + i, &v, sizeof(T), hipMemcpyHostToDevice), "copy");
  }
};

static const acpp::property_list init_to_identity{acpp::property::reduction::initialize_to_identity{}};

// tests/sycl/reduction.cpp:38-45, 323-344: sums of 0..n-1 with and without
// initialize_to_identity; a given identity; empty ranges.
static void test_single_reductions(acpp::queue &q, alg::util::allocation_cache &cache) {
  alg::util::allocation_group g{&cache, q};
  for (size_t n : {size_t(0), size_t(1), size_t(1024), size_t(1000003)}) {
    std::vector<long> x(n);
    std::iota(x.begin(), x.end(), 0L);
    device_buffer<long> xd(n), out(1);
    if (n) xd.upload(x);
    const long sum = std::accumulate(x.begin(), x.end(), 0L);
    out.set(12345);
    acpp::parallel_for(q, g, n, acpp::combine_each(xd.ptr, acpp::reduction(out.ptr, acpp::plus<long>{}, init_to_identity)));
    q.wait();
    CHECK(out.get() == sum, "init_to_identity sum n=%zu", n);
    out.set(100);
    acpp::parallel_for(q, g, n, acpp::combine_each(xd.ptr, acpp::reduction(out.ptr, std::plus<long>{})));
    acpp::parallel_for(q, g, n, acpp::combine_each(xd.ptr, acpp::reduction(out.ptr, std::plus<long>{})));
    q.wait();
    CHECK(out.get() == 100 + 2 * sum, "accumulating sum n=%zu", n);
    out.set(7);
    acpp::parallel_for(q, g, n, acpp::combine_each(xd.ptr, acpp::reduction(out.ptr, -5L, acpp::maximum<long>{}, init_to_identity)));
    q.wait();
    CHECK(out.get() == (n ? std::max(-5L, static_cast<long>(n) - 1) : -5L), "given identity max n=%zu", n);
  }
  // float sum: a given +0.0 identity is what an empty reduction yields, and
  // it absorbs a -0.0 fold; the library's own identity is -0.0.
  device_buffer<float> fx(4), fo(1);
  fx.upload({-0.0f, -0.0f, -0.0f, -0.0f});
  acpp::parallel_for(q, g, 0, acpp::combine_each(fx.ptr, acpp::reduction(fo.ptr, 0.0f, acpp::plus<float>{}, init_to_identity)));
  q.wait();
  CHECK(fo.get() == 0.0f && !std::signbit(fo.get()), "empty float reduction yields the given +0.0");
  acpp::parallel_for(q, g, 4, acpp::combine_each(fx.ptr, acpp::reduction(fo.ptr, acpp::plus<float>{}, init_to_identity)));
  q.wait();
  CHECK(fo.get() == 0.0f && std::signbit(fo.get()), "sum of -0.0 with the library identity stays -0.0");
  acpp::parallel_for(q, g, 4, acpp::combine_each(fx.ptr, acpp::reduction(fo.ptr, 0.0f, acpp::plus<float>{}, init_to_identity)));
  q.wait();
  CHECK(fo.get() == 0.0f && !std::signbit(fo.get()), "sum of -0.0 with a given +0.0 identity is +0.0");
  // narrow integers: wrapping product
  std::vector<short> s(777);
  for (size_t i = 0; i < s.size(); ++i) s[i] = static_cast<short>(i % 5 == 0 ? 3 : 1);
  device_buffer<short> sd(s.size()), so(1);
  sd.upload(s);
  acpp::parallel_for(q, g, s.size(), acpp::combine_each(sd.ptr, acpp::reduction(so.ptr, acpp::multiplies<short>{}, init_to_identity)));
  q.wait();
  unsigned prod = 1;
  for (short v : s) prod = (prod * static_cast<unsigned>(v)) & 0xffffu;
  CHECK(so.get() == static_cast<short>(prod), "int16 product");
}

// tests/sycl/reduction.cpp:116-208: several reductions in one parallel_for.
static void test_fused_reductions(acpp::queue &q, alg::util::allocation_cache &cache) {
  alg::util::allocation_group g{&cache, q};
  const size_t n = 3000001;
  std::vector<int> x(n), y(n);
  std::mt19937 rng(5);
  for (size_t i = 0; i < n; ++i) {
    x[i] = static_cast<int>(rng() % 2001) - 1000;
    y[i] = static_cast<int>(rng() % 3);
  }
  device_buffer<int> xd(n), yd(n), o(4);
  xd.upload(x);
  yd.upload(y);
  o.upload({10, 20, 30, 40});
  acpp::parallel_for(q, g, n, acpp::combine_each(xd.ptr, acpp::reduction(o.ptr + 0, acpp::plus<int>{})),
                     acpp::combine_each(xd.ptr, acpp::reduction(o.ptr + 1, acpp::minimum<int>{}, init_to_identity)),
                     acpp::combine_each(yd.ptr, acpp::reduction(o.ptr + 2, 0, acpp::bit_or<int>{}, init_to_identity)),
                     acpp::combine_each(yd.ptr, acpp::reduction(o.ptr + 3, acpp::maximum<int>{})));
  q.wait();
  const auto got = o.download();
  long long sum = 10;
  for (int v : x) sum += v;
  int mn = x[0], orv = 0, mx = 40;
  for (size_t i = 0; i < n; ++i) {
    mn = std::min(mn, x[i]);
    orv |= y[i];
    mx = std::max(mx, y[i]);
  }
  CHECK(got[0] == static_cast<int>(sum), "fused: accumulating sum");
  CHECK(got[1] == mn, "fused: minimum from the identity");
  CHECK(got[2] == orv, "fused: bit_or with a given identity");
  CHECK(got[3] == mx, "fused: accumulating maximum");
  // empty range: given identities and the operators' own
  o.upload({1, 2, 3, 4});
  acpp::parallel_for(q, g, 0, acpp::combine_each(xd.ptr, acpp::reduction(o.ptr + 0, 77, acpp::plus<int>{}, init_to_identity)),
                     acpp::combine_each(xd.ptr, acpp::reduction(o.ptr + 1, acpp::maximum<int>{}, init_to_identity)),
                     acpp::combine_each(xd.ptr, acpp::reduction(o.ptr + 2, acpp::plus<int>{})));
  q.wait();
  const auto e = o.download();
  CHECK(e[0] == 77 && e[1] == std::numeric_limits<int>::lowest() && e[2] == 3 && e[3] == 4,
        "fused: empty range (%d %d %d %d)", e[0], e[1], e[2], e[3]);
}

// tests/pstl/copy_if.cpp: sizes 0, 1, 10, 1000, and large; predicates of
// each kind; count and contents against std::copy_if.
template <class T, class P>
static void copy_if_case(acpp::queue &q, alg::util::allocation_group &g, const std::vector<T> &x, P pred,
                         const char *what) {
  std::vector<T> want;
  std::copy_if(x.begin(), x.end(), std::back_inserter(want), pred);
  device_buffer<T> xd(x.size()), d(x.size());
  device_buffer<std::size_t> cnt(1);
  if (!x.empty()) xd.upload(x);
  cnt.set(999);
  alg::copy_if(q, g, xd.ptr, xd.ptr + x.size(), d.ptr, pred, cnt.ptr);
  q.wait();
  const std::size_t c = cnt.get();
  CHECK(c == want.size(), "copy_if %s n=%zu: count %zu != %zu", what, x.size(), c, want.size());
  if (c == want.size() && c) {
    auto got = d.download();
    got.resize(c);
    CHECK(got == want, "copy_if %s n=%zu: contents", what, x.size());
  }
}

static void test_copy_if(acpp::queue &q, alg::util::allocation_cache &cache) {
  alg::util::allocation_group g{&cache, q};
  for (size_t n : {size_t(0), size_t(1), size_t(10), size_t(1000), size_t(1 << 20) + 17}) {
    std::vector<int> x(n);
    std::mt19937 rng(static_cast<unsigned>(n));
    for (auto &v : x) v = static_cast<int>(rng() % 100) - 50;
    copy_if_case(q, g, x, acpp::is_greater_than(0), "x > 0");
    copy_if_case(q, g, x, acpp::is_equal_to(7), "x == 7");
    copy_if_case(q, g, x, acpp::has_bits(1, 0), "even");
    copy_if_case(q, g, x, acpp::is_nonzero<>{}, "x != 0");
    std::vector<double> fx(x.begin(), x.end());
    copy_if_case(q, g, fx, acpp::is_less_equal(-10.0), "x <= -10.0");
    std::vector<unsigned long> ux(x.begin(), x.end());
    copy_if_case(q, g, ux, acpp::is_not_equal_to(0ul), "unsigned long x != 0");
    std::vector<signed char> bx(x.begin(), x.end());
    copy_if_case(q, g, bx, acpp::is_less_than(static_cast<signed char>(-5)), "int8 x < -5");
    std::vector<unsigned short> hx(x.begin(), x.end());
    copy_if_case(q, g, hx, acpp::has_bits(static_cast<unsigned short>(3), static_cast<unsigned short>(1)),
                 "uint16 (x & 3) == 1");
  }
  // empty range without a count: a complete event
  device_buffer<int> e(1);
  acpp::event ev = alg::copy_if(q, g, e.ptr, e.ptr, e.ptr, acpp::is_zero<>{});
  CHECK(ev.native() == nullptr, "empty copy_if without a count enqueues nothing");
}

// Binary transform_reduce with pair transforms (numeric.hpp:170-191)
// against std::transform_reduce's serial result.
static void test_binary_transform_reduce(acpp::queue &q, alg::util::allocation_cache &cache) {
  alg::util::allocation_group g{&cache, q};
  const size_t n = 1000003;
  std::vector<long> a(n), b(n);
  std::mt19937_64 rng(9);
  for (size_t i = 0; i < n; ++i) {
    a[i] = static_cast<long>(rng() % 2001) - 1000;
    b[i] = static_cast<long>(rng() % 2001) - 1000;
  }
  device_buffer<long> ad(n), bd(n), out(1);
  ad.upload(a);
  bd.upload(b);
  alg::transform_reduce(q, g, ad.ptr, ad.ptr + n, bd.ptr, out.ptr, 0L, std::plus<long>{},
                        acpp::absolute_difference<long>{});
  q.wait();
  CHECK(out.get() == std::transform_reduce(a.begin(), a.end(), b.begin(), 0L, std::plus<long>{},
                                           acpp::absolute_difference<long>{}),
        "sum of absolute differences");
  alg::transform_reduce(q, g, ad.ptr, ad.ptr + n, bd.ptr, out.ptr, -1L, acpp::maximum<long>{},
                        acpp::squared_difference<long>{});
  q.wait();
  CHECK(out.get() == std::transform_reduce(a.begin(), a.end(), b.begin(), -1L, acpp::maximum<long>{},
                                           acpp::squared_difference<long>{}),
        "maximum of squared differences");
  alg::transform_reduce(q, g, ad.ptr, ad.ptr + n, bd.ptr, out.ptr, 5L, acpp::minimum<long>{}, std::minus<long>{});
  q.wait();
  CHECK(out.get() == std::transform_reduce(a.begin(), a.end(), b.begin(), 5L, acpp::minimum<long>{},
                                           std::minus<long>{}),
        "minimum of differences");
}

// queue::wait throws a give-up of a kernel's wait (forced from the host
// with ACPP_VARIANT_FORCE_GIVE_UP), once.
static void test_async_error(acpp::queue &q) {
  alg::util::allocation_cache cache{alg::util::allocation_type::device};
  alg::util::allocation_group g{&cache, q};
  const size_t n = 3 * 8192;
  std::vector<int> x(n, 1);
  device_buffer<int> xd(n), yd(n);
  xd.upload(x);
  acpp_launch_config cfg{0, 0, 0, ACPP_VARIANT_FORCE_GIVE_UP};
  acpp::detail::check_status(acpp_inclusive_scan_i32(g.native(), q.get_native(), xd.ptr, n, yd.ptr, 0, nullptr, &cfg));
  int code = -1;
  try {
    q.wait();
  } catch (const acpp::exception &e) {
    code = e.code();
  }
  CHECK(code == ACPP_ERROR_WATCHDOG, "queue::wait throws the forced give-up (code %d)", code);
  bool again = false;
  try {
    q.wait();
  } catch (const acpp::exception &) {
    again = true;
  }
  CHECK(!again, "the give-up is thrown once");
}

int main() {
  try {
    acpp::queue q{0};
    alg::util::allocation_cache cache{alg::util::allocation_type::device};
    test_single_reductions(q, cache);
    test_fused_reductions(q, cache);
    test_copy_if(q, cache);
    test_binary_transform_reduce(q, cache);
    test_async_error(q);
  } catch (const std::exception &e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 2;
  }
  if (failures) {
    std::printf("%d FAILED\n", failures);
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
