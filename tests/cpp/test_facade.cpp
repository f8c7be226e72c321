// C++ facade test: the acpp::algorithms templates with the reference's
// parameter lists (include/hipSYCL/algorithms/numeric.hpp:170-297), checked
// against the C++ standard library's sequential algorithms, as the
// reference's own algorithm tests do (tests/pstl/reduce.cpp,
// inclusive_scan.cpp, exclusive_scan.cpp: std:: result vs offloaded result).
//
// Exit status 0 = every case passed.  Prints one line per case.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

#include "acpp/algorithms/numeric.hpp"

namespace alg = acpp::algorithms;

static int failures = 0;

#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      std::printf("FAIL: " __VA_ARGS__);  \
      std::printf("\n");                  \
      ++failures;                         \
    }                                     \
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
};

template <class T>
static std::vector<T> random_ints(size_t n, uint64_t seed, int64_t lo, int64_t hi) {
  std::mt19937_64 g(seed);
  std::uniform_int_distribution<int64_t> d(lo, hi);
  std::vector<T> v(n);
  for (auto &x : v) x = static_cast<T>(d(g));
  return v;
}

// Wrapping serial references (unsigned arithmetic, no signed overflow).
template <class T>
static T wrap_add(T a, T b) {
  using U = std::make_unsigned_t<T>;
  return static_cast<T>(static_cast<U>(a) + static_cast<U>(b));
}

static void test_int_reduce_and_scans(acpp::queue &q, alg::util::allocation_cache &cache) {
  // The sizes of the reference's pstl tests: 0, 1, 127, 1000, 2^20
  // (tests/pstl/reduce.cpp:30-43, inclusive_scan.cpp:77-110).
  for (size_t n : {size_t(0), size_t(1), size_t(127), size_t(1000), size_t(1) << 20}) {
    auto h = random_ints<int>(n, n + 1, -1000, 1000);
    device_buffer<int> x(n), y(n), z(n), out(1);
    x.upload(h);
    std::vector<int> sentinel{-77};
    out.upload(sentinel);
    alg::util::allocation_group scratch{&cache, q};
    alg::reduce(q, scratch, x.ptr, x.ptr + n, out.ptr, 5);
    alg::inclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr);
    auto e = alg::exclusive_scan(q, scratch, x.ptr, x.ptr + n, z.ptr, 3);
    e.wait();
    int expect = n ? 5 : -77;  // empty: *out untouched (numeric.hpp:165-169)
    for (size_t i = 0; i < n; ++i) expect = wrap_add(expect, h[i]);
    CHECK(out.download()[0] == expect, "reduce int n=%zu", n);
    if (n) {
      std::vector<int> ri(n), re(n);
      int acc = 0;
      for (size_t i = 0; i < n; ++i) { acc = wrap_add(acc, h[i]); ri[i] = acc; }
      acc = 3;
      for (size_t i = 0; i < n; ++i) { re[i] = acc; acc = wrap_add(acc, h[i]); }
      CHECK(y.download() == ri, "inclusive_scan int n=%zu", n);
      CHECK(z.download() == re, "exclusive_scan int n=%zu", n);
    }
    std::printf("int reduce / inclusive_scan / exclusive_scan n=%zu done\n", n);
  }
}

static void test_ops_and_init(acpp::queue &q, alg::util::allocation_cache &cache) {
  const size_t n = 100003;
  auto h = random_ints<long>(n, 7, -(1 << 20), 1 << 20);
  device_buffer<long> x(n), y(n), out(1);
  x.upload(h);
  alg::util::allocation_group scratch{&cache, q};
  // maximum with init, std::vector-free pointer form
  alg::reduce(q, scratch, x.ptr, x.ptr + n, out.ptr, long(-5), acpp::maximum<long>{}).wait();
  long mx = -5;
  for (long v : h) mx = std::max(mx, v);
  CHECK(out.download()[0] == mx, "reduce max long");
  // inclusive scan with op and init (decoupled_lookback_scan.hpp:264-271)
  alg::inclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr, acpp::minimum<long>{}, long(0)).wait();
  std::vector<long> r(n);
  long acc = 0;
  for (size_t i = 0; i < n; ++i) { acc = std::min(acc, h[i]); r[i] = acc; }
  CHECK(y.download() == r, "inclusive_scan min long with init");
  // bit_xor exclusive scan with a std:: functor
  alg::exclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr, long(0x55), std::bit_xor<long>{}).wait();
  acc = 0x55;
  for (size_t i = 0; i < n; ++i) { r[i] = acc; acc ^= h[i]; }
  CHECK(y.download() == r, "exclusive_scan xor long");
  std::printf("operators and init done\n");
}

static void test_float_and_dot(acpp::queue &q, alg::util::allocation_cache &cache) {
  const size_t n = (1 << 22) + 9;
  std::mt19937_64 g(11);
  std::normal_distribution<float> d(0.f, 1.f);
  std::vector<float> a(n), b(n);
  for (auto &v : a) v = d(g);
  for (auto &v : b) v = d(g);
  device_buffer<float> da(n), db(n), out(2);
  da.upload(a);
  db.upload(b);
  alg::util::allocation_group scratch{&cache, q};
  auto e1 = alg::reduce(q, scratch, da.ptr, da.ptr + n, out.ptr, 0.0f);
  auto e2 = alg::transform_reduce(q, scratch, da.ptr, da.ptr + n, db.ptr, out.ptr + 1, 1.0f, {e1});
  e2.wait();
  double s = 0, sa = 0, dot = 1, dabs = 1;
  for (size_t i = 0; i < n; ++i) {
    s += a[i];
    sa += std::fabs(a[i]);
    dot += double(a[i]) * b[i];
    dabs += std::fabs(double(a[i]) * b[i]);
  }
  auto got = out.download();
  // Tree-ordered fp32 sums: error <= (depth) * eps * sum |x|, depth ~ 64.
  const double eps = 1.0 / (1 << 23);
  CHECK(std::fabs(got[0] - s) <= 64 * eps * sa, "reduce float: %g vs %g", got[0], s);
  CHECK(std::fabs(got[1] - dot) <= 64 * eps * dabs, "dot float: %g vs %g", got[1], dot);
  std::printf("float reduce / dot done\n");
}

static void test_segmented(acpp::queue &q, alg::util::allocation_cache &cache) {
  std::vector<uint64_t> off{0};
  std::mt19937_64 g(3);
  for (int s = 0; s < 3000; ++s) off.push_back(off.back() + (g() % 2000));
  const size_t n = off.back();
  auto h = random_ints<long>(n, 5, -1000, 1000);
  device_buffer<long> x(n), y(n), red(off.size() - 1);
  device_buffer<uint64_t> doff(off.size());
  x.upload(h);
  doff.upload(off);
  alg::util::allocation_group scratch{&cache, q};
  alg::segmented_inclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr, doff.ptr, off.size() - 1);
  alg::segmented_reduce(q, scratch, x.ptr, x.ptr + n, doff.ptr, off.size() - 1, red.ptr, long(2)).wait();
  std::vector<long> ry(n), rr(off.size() - 1);
  for (size_t s = 0; s + 1 < off.size(); ++s) {
    long acc = 0;
    for (uint64_t i = off[s]; i < off[s + 1]; ++i) { acc += h[i]; ry[i] = acc; }
    rr[s] = 2 + acc;
  }
  CHECK(y.download() == ry, "segmented_inclusive_scan long");
  CHECK(red.download() == rr, "segmented_reduce long");
  std::printf("segmented done\n");
}

// Unary transform_reduce and the transform scans (numeric.hpp:193-212,
// 299-327) against std::transform_reduce / std::transform_*_scan.
static void test_transforms(acpp::queue &q, alg::util::allocation_cache &cache) {
  for (size_t n : {size_t(1), size_t(1000), size_t(123457)}) {
    alg::util::allocation_group scratch{&cache, q};
    auto h = random_ints<int>(n, 7 + n, -1000, 1000);
    device_buffer<int> x(n), y(n), out(1);
    x.upload(h);
    alg::transform_reduce(q, scratch, x.ptr, x.ptr + n, out.ptr, 5, std::plus<int>{}, acpp::square<int>{});
    q.wait();
    int ref = std::transform_reduce(h.begin(), h.end(), 5, std::plus<>{}, [](int v) { return v * v; });
    CHECK(out.download()[0] == ref, "transform_reduce(plus, square) n=%zu", n);
    alg::transform_reduce(q, scratch, x.ptr, x.ptr + n, out.ptr, 0, acpp::minimum<int>{}, acpp::absolute<int>{});
    q.wait();
    ref = std::transform_reduce(h.begin(), h.end(), 0, acpp::minimum<int>{}, [](int v) { return v < 0 ? -v : v; });
    CHECK(out.download()[0] == ref, "transform_reduce(min, abs) n=%zu", n);
    std::vector<int> r(n);
    alg::transform_inclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr, std::plus<>{}, std::negate<int>{});
    q.wait();
    std::transform_inclusive_scan(h.begin(), h.end(), r.begin(), std::plus<>{}, std::negate<int>{});
    CHECK(y.download() == r, "transform_inclusive_scan(negate) n=%zu", n);
    alg::transform_inclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr, acpp::maximum<int>{}, acpp::twice<int>{}, -7);
    q.wait();
    std::transform_inclusive_scan(h.begin(), h.end(), r.begin(), acpp::maximum<int>{}, [](int v) { return v + v; }, -7);
    CHECK(y.download() == r, "transform_inclusive_scan(max, twice, init) n=%zu", n);
    alg::transform_exclusive_scan(q, scratch, x.ptr, x.ptr + n, y.ptr, 3, std::plus<>{}, acpp::is_zero<int>{});
    q.wait();
    std::transform_exclusive_scan(h.begin(), h.end(), r.begin(), 3, std::plus<>{}, [](int v) { return v == 0 ? 1 : 0; });
    CHECK(y.download() == r, "transform_exclusive_scan(is_zero) n=%zu", n);
  }
  std::printf("transforms done\n");
}

static void test_errors(acpp::queue &q, alg::util::allocation_cache &cache) {
  // Misuse reported by the library surfaces as acpp::exception.
  alg::util::allocation_group scratch{&cache, q};
  device_buffer<int> x(10), y(10);
  bool threw = false;
  try {
    alg::segmented_inclusive_scan(q, scratch, x.ptr, x.ptr + 4, y.ptr, static_cast<const uint64_t *>(nullptr), 2);
  } catch (const acpp::exception &e) {
    threw = (e.code() == ACPP_ERROR_INVALID_ARGUMENT);
  }
  CHECK(threw, "null offsets must raise acpp::exception(INVALID_ARGUMENT)");
  std::printf("errors done\n");
}

int main() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    std::printf("no GPU\n");
    return 2;
  }
  acpp::queue q{0};
  alg::util::allocation_cache cache{alg::util::allocation_type::device};
  test_int_reduce_and_scans(q, cache);
  test_ops_and_init(q, cache);
  test_float_and_dot(q, cache);
  test_segmented(q, cache);
  test_transforms(q, cache);
  test_errors(q, cache);
  q.wait();
  std::printf(failures ? "FAILED (%d)\n" : "ALL PASSED\n", failures);
  return failures ? 1 : 0;
}
