"""C++ facade (include/acpp/algorithms/numeric.hpp): builds against the
library, rejects what the library does not provide at compile time, and
(on a GPU) runs tests/cpp/test_facade.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
BIN = os.path.join(ROOT, "tests", "lib", "test_facade")


def _syntax(code, tmp_path):
    src = tmp_path / "snippet.cpp"
    src.write_text(code)
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{INCLUDE}", "-I/opt/rocm/include",
                           "-D__HIP_PLATFORM_AMD__", str(src)], capture_output=True, text=True)


PRELUDE = """
#include "acpp/algorithms/numeric.hpp"
namespace alg = acpp::algorithms;
void f(acpp::queue& q, alg::util::allocation_group& g, int* x, int* y, float* fx, long* lx) {
"""


def test_facade_signatures_compile(tmp_path):
    # The reference's call forms (numeric.hpp:170-297) must all compile.
    r = _syntax(PRELUDE + """
  std::vector<acpp::event> deps;
  alg::reduce(q, g, x, x + 10, y);
  alg::reduce(q, g, x, x + 10, y, 3);
  alg::reduce(q, g, x, x + 10, y, 3, std::plus<int>{});
  alg::reduce(q, g, x, x + 10, y, 3, acpp::maximum<int>{}, deps);
  alg::reduce(q, g, fx, fx + 10, fx, 0.f, acpp::minimum<float>{});
  alg::transform_reduce(q, g, fx, fx + 10, fx, fx, 0.f);
  alg::transform_reduce(q, g, fx, fx + 10, fx, fx, 0.f, std::plus<float>{}, std::multiplies<float>{});
  alg::transform_reduce(q, g, x, x + 10, x, y, 0, acpp::maximum<int>{}, acpp::absolute_difference<>{});
  alg::transform_reduce(q, g, fx, fx + 10, fx, fx, 0.f, std::plus<float>{}, acpp::squared_difference<float>{});
  alg::transform_reduce(q, g, lx, lx + 10, lx, lx, 0L, acpp::minimum<long>{}, std::minus<long>{}, deps);
  alg::inclusive_scan(q, g, x, x + 10, y);
  alg::inclusive_scan(q, g, x, x + 10, y, std::bit_xor<int>{});
  alg::inclusive_scan(q, g, x, x + 10, y, acpp::plus<int>{}, 7, deps);
  alg::exclusive_scan(q, g, x, x + 10, y, 0);
  alg::exclusive_scan(q, g, lx, lx + 10, lx, 0L, acpp::maximum<long>{});
  alg::transform_reduce(q, g, x, x + 10, y, 0, std::plus<int>{}, acpp::square<int>{});
  alg::transform_reduce(q, g, fx, fx + 10, fx, 0.f, acpp::maximum<float>{}, acpp::absolute<>{}, deps);
  alg::transform_inclusive_scan(q, g, x, x + 10, y, std::plus<>{}, std::negate<int>{});
  alg::transform_inclusive_scan(q, g, x, x + 10, y, std::plus<>{}, acpp::self_op<std::plus<>>{}, 3);
  alg::transform_exclusive_scan(q, g, lx, lx + 10, lx, 3L, acpp::bit_or<long>{}, std::bit_not<long>{});
  alg::transform_exclusive_scan(q, g, x, x + 10, y, 0, std::plus<>{}, acpp::is_nonzero<>{}, deps);
}
""", tmp_path)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("call,message", [
    ("alg::reduce(q, g, x, x + 10, y, 0, [](int a, int b) { return a + b; });", "operator must be"),
    ("alg::reduce(q, g, fx, fx + 10, fx, 0.f, std::bit_and<float>{});", "floating-point types take"),
    ("alg::reduce(q, g, fx, fx + 10, fx, 0.f, acpp::bit_or<float>{});", "floating-point types take"),
    ("alg::transform_reduce(q, g, x, x + 10, x, y, 0, acpp::bit_and<int>{}, std::multiplies<int>{});",
     "reduces with plus, multiplies"),
    ("alg::transform_reduce(q, g, x, x + 10, x, y, 0, std::plus<int>{}, [](int a, int b) { return a - b; });",
     "pair transform must be"),
    ("alg::transform_reduce(q, g, fx, fx + 10, fx, fx, 0.f, std::plus<float>{}, std::bit_xor<float>{});",
     "bitwise pair transforms need"),
    ("alg::reduce(q, g, x, x + 10, lx, 0L);", "input value type must equal T"),
    ("alg::transform_reduce(q, g, x, x + 10, y, 0, std::plus<int>{}, [](int v) { return v; });",
     "element transform must be"),
    ("alg::transform_inclusive_scan(q, g, fx, fx + 10, fx, std::plus<>{}, acpp::bit_not<>{});",
     "bit_not needs an integer"),
    ("alg::transform_exclusive_scan(q, g, x, x + 10, y, 0, std::plus<>{}, acpp::self_op<std::bit_xor<>>{});",
     "element transform must be"),
])
def test_facade_rejects_at_compile_time(tmp_path, call, message):
    # No generic-functor fallback: what the library lacks fails to compile.
    r = _syntax(PRELUDE + call + "\n}\n", tmp_path)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr[-2000:]


def test_facade_binary_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_facade_on_gpu():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


STDPAR_BIN = os.path.join(ROOT, "tests", "lib", "test_stdpar")


def test_stdpar_binary_built():
    assert os.path.exists(STDPAR_BIN), "run __graft_entry__.build() first"


def test_stdpar_std_vector_with_custom_allocator_compiles(tmp_path):
    # std::vector iterators with any allocator are contiguous iterators
    r = _syntax("""
#include "acpp/stdpar/numeric.hpp"
template <class T> struct A : std::allocator<T> {};
long f(std::vector<long, A<long>>& v, std::vector<long, A<long>>& w) {
  acpp::stdpar::inclusive_scan(acpp::stdpar::par_unseq, v.begin(), v.end(), w.begin());
  return acpp::stdpar::reduce(acpp::stdpar::par_unseq, v.cbegin(), v.cend(), 0L);
}
""", tmp_path)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_stdpar_on_gpu():
    r = subprocess.run([STDPAR_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


RED_BIN = os.path.join(ROOT, "tests", "lib", "test_reduction_copy_if")

RED_PRELUDE = """
#include "acpp/reduction.hpp"
#include "acpp/algorithms/algorithm.hpp"
namespace alg = acpp::algorithms;
void f(acpp::queue& q, alg::util::allocation_group& g, int* x, int* y, float* fx, long* lx,
       std::size_t* cnt) {
  const acpp::property_list init{acpp::property::reduction::initialize_to_identity{}};
"""


def test_reduction_and_copy_if_signatures_compile(tmp_path):
    # reduction.hpp:165-236 factories, handler.hpp:799-975 fused form,
    # algorithm.hpp:178-187 copy_if
    r = _syntax(RED_PRELUDE + """
  acpp::parallel_for(q, g, 10, acpp::combine_each(x, acpp::reduction(y, acpp::plus<int>{})));
  acpp::parallel_for(q, g, 10, acpp::combine_each(x, acpp::reduction(y, std::plus<int>{}, init)));
  acpp::parallel_for(q, g, 10, acpp::combine_each(x, acpp::reduction(y, 0, acpp::plus<int>{}, init)), {});
  acpp::parallel_for(q, g, 10, acpp::combine_each(fx, acpp::reduction(fx, 1.f, acpp::multiplies<float>{})));
  acpp::parallel_for(q, g, 10, acpp::combine_each(x, acpp::reduction(y, acpp::plus<int>{})),
                     acpp::combine_each(x, acpp::reduction(y + 1, 5, acpp::maximum<int>{}, init)));
  alg::copy_if(q, g, x, x + 10, y, acpp::is_greater_than(3));
  alg::copy_if(q, g, x, x + 10, y, acpp::has_bits(1, 0), cnt);
  alg::copy_if(q, g, fx, fx + 10, fx, acpp::is_nonzero<>{}, cnt, {});
  alg::copy_if(q, g, lx, lx + 10, lx, acpp::is_less_equal(7L), nullptr);
}
""", tmp_path)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("call,message", [
    ("alg::copy_if(q, g, x, x + 10, y, [](int v) { return v > 0; });", "the predicate must be"),
    ("alg::copy_if(q, g, x, x + 10, y, acpp::is_greater_than(3L));", "the predicate must be"),
    ("alg::copy_if(q, g, fx, fx + 10, fx, acpp::has_bits(1.f, 0.f));", "has_bits needs an integer"),
    ("acpp::parallel_for(q, g, 10, acpp::combine_each(x, acpp::reduction(y, [](int a, int b) { return a + b; })));",
     "operator must be"),
    ("acpp::parallel_for(q, g, 10, acpp::combine_each(fx, acpp::reduction(fx, acpp::bit_or<float>{})));",
     "floating-point types take"),
])
def test_reduction_and_copy_if_reject_at_compile_time(tmp_path, call, message):
    r = _syntax(RED_PRELUDE + call + "\n}\n", tmp_path)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr[-2000:]


def test_reduction_copy_if_binary_built():
    assert os.path.exists(RED_BIN), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_reduction_and_copy_if_on_gpu():
    r = subprocess.run([RED_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


STDPAR_STD_BIN = os.path.join(ROOT, "tests", "lib", "test_stdpar_std")


def test_stdpar_std_overloads_chosen_over_the_standard_library(tmp_path):
    # with include/acpp/stdpar/std first, std::reduce(par_unseq, ...) must
    # resolve to the forwarding overload (libstdc++'s own templates on any
    # execution policy would make the call ambiguous or bypass the library)
    src = tmp_path / "s.cpp"
    src.write_text("""
#include <numeric>
#include <execution>
#include <vector>
long f(std::vector<long>& v, std::vector<long>& w) {
  std::inclusive_scan(std::execution::par_unseq, v.begin(), v.end(), w.begin());
  std::exclusive_scan(std::execution::par_unseq, v.begin(), v.end(), w.begin(), 3L);
  return std::reduce(std::execution::par_unseq, v.begin(), v.end()) +
         std::transform_reduce(std::execution::par_unseq, v.begin(), v.end(), w.begin(), 0L) +
         std::reduce(std::execution::seq, v.begin(), v.end());
}
""")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{os.path.join(INCLUDE, 'acpp', 'stdpar', 'std')}",
                        f"-I{INCLUDE}", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_stdpar_std_on_gpu():
    r = subprocess.run([STDPAR_STD_BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
