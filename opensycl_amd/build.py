"""In-tree build of the HIP library for gfx950.

Produces ``opensycl_amd/lib/libacpp_mi355x.so`` from ``opensycl_amd/csrc``.
Every HIP translation unit is compiled with hipcc for ``--offload-arch=gfx950``
with packed-FP32 instructions disabled; objects go to ``opensycl_amd/_build``.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_NAME = "libacpp_mi355x.so"
LIB_PATH = os.path.join(LIB_DIR, LIB_NAME)
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    # Packed FP32 VALU instructions are kept out of every code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _sources() -> list[str]:
    names = sorted(os.listdir(CSRC))
    return [os.path.join(CSRC, n) for n in names if n.endswith((".hip", ".cpp"))]


def _headers_mtime() -> float:
    latest = 0.0
    for d in (CSRC, os.path.join(INCLUDE), os.path.join(INCLUDE, "acpp", "device")):
        for n in os.listdir(d):
            if n.endswith((".hpp", ".h")):
                latest = max(latest, os.path.getmtime(os.path.join(d, n)))
    return latest


def _compile(src: str, obj: str) -> tuple[str, int, str]:
    cmd = [HIPCC, *COMMON_FLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall",
               "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the GPU target feature; that note is noise.
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    return src, r.returncode, err


def build(force: bool = True, jobs: int | None = None, verbose: bool = False) -> str:
    """Compile and link the library; returns the path of the shared object."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    hdr_time = _headers_mtime()
    todo = []
    objs = []
    for src in _sources():
        obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
        objs.append(obj)
        stale = (not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(src)
                 or os.path.getmtime(obj) < hdr_time)
        if force or stale:
            todo.append((src, obj))
    jobs = jobs or min(8, os.cpu_count() or 4)
    failures = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, err in ex.map(lambda a: _compile(*a), todo):
            if verbose or (rc != 0):
                print(f"[build] {os.path.basename(src)}: rc={rc}", file=sys.stderr)
            if err.strip() and (verbose or rc != 0):
                print(err, file=sys.stderr)
            if rc != 0:
                failures.append(src)
    if failures:
        raise RuntimeError("compilation failed: " + ", ".join(os.path.basename(f) for f in failures))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_test_kernels() -> str:
    """Build tests/lib/libacpp_test_kernels.so: kernels that exercise the
    device headers (group collectives, atomic_ref) for the GPU tests."""
    src_dir = os.path.join(ROOT, "tests", "csrc")
    out_dir = os.path.join(ROOT, "tests", "lib")
    os.makedirs(out_dir, exist_ok=True)
    srcs = sorted(os.path.join(src_dir, n) for n in os.listdir(src_dir) if n.endswith(".hip"))
    out = os.path.join(out_dir, "libacpp_test_kernels.so")
    cmd = [HIPCC, *COMMON_FLAGS, "-shared", "-o", out + ".tmp", *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("test kernel build failed:\n" + r.stderr)
    os.replace(out + ".tmp", out)
    return out


def build_bench_helpers() -> str:
    """Build opensycl_amd/lib/libacpp_bench.so: native per-segment call loops
    used as the baseline of the segmented benchmarks."""
    src = os.path.join(PKG_DIR, "bench", "per_segment.cpp")
    out = os.path.join(LIB_DIR, "libacpp_bench.so")
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{INCLUDE}", "-o", out + ".tmp", src,
           f"-L{LIB_DIR}", "-lacpp_mi355x", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("bench helper build failed:\n" + r.stderr)
    os.replace(out + ".tmp", out)
    return out


def build_cpp_tests() -> list[str]:
    """Build tests/lib/<name> for every tests/cpp/<name>.cpp: C++ tests of the
    facade and the stdpar entry points (host code only, compiled with g++
    against the library and the HIP runtime)."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(ROOT, "tests", "lib")
    os.makedirs(out_dir, exist_ok=True)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    outs = []
    for name in sorted(os.listdir(src_dir)):
        if not name.endswith(".cpp"):
            continue
        src = os.path.join(src_dir, name)
        out = os.path.join(out_dir, name[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{INCLUDE}", f"-I{rocm}/include",
               "-D__HIP_PLATFORM_AMD__", src, "-o", out + ".tmp", f"-L{LIB_DIR}", "-lacpp_mi355x",
               f"-L{rocm}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../opensycl_amd/lib",
               f"-Wl,-rpath,{rocm}/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test {name} build failed:\n" + r.stderr)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_oracle() -> str:
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    d = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", d, "-B"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return os.path.join(d, "liboracle_acpp.so")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose="-v" in sys.argv))
