"""CPU-side checks of the C-ABI boundary (no GPU needed, no compute calls).

- the HIP library loads and exports every function include/*.h declares;
- argument validation answers with status codes, never a crash;
- the allocation cache / group life cycle works without touching a device.
"""
import ctypes
import os
import re

import pytest

import opensycl_amd as acpp
from opensycl_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(acpp_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_binding_signature():
    names = declared_functions()
    missing = [n for n in names if n not in _abi.SIGNATURES]
    assert not missing, missing


def test_version_and_errors():
    lib = _abi.load()
    assert lib.acpp_version().decode().count(".") == 2


def test_null_arguments_are_rejected():
    lib = _abi.load()
    out = ctypes.c_int32()
    st = lib.acpp_reduce_i32(None, None, None, 10, ctypes.addressof(out), 0, 0, None)
    assert st == _abi.ERROR_INVALID_ARGUMENT
    assert b"scratch" in lib.acpp_last_error_message()
    cache = acpp.AllocationCache()
    g = acpp.AllocationGroup(cache, 0)
    st = lib.acpp_reduce_i32(g.handle, None, None, 10, ctypes.addressof(out), 0, 0, None)
    assert st == _abi.ERROR_INVALID_ARGUMENT
    st = lib.acpp_inclusive_scan_f32(g.handle, None, None, 10, None, 0, None, None)
    assert st == _abi.ERROR_INVALID_ARGUMENT
    # bitwise operators on floating point: unsupported combination
    st = lib.acpp_exclusive_scan_f32(g.handle, None, 8, 0, 16, ctypes.c_float(0), 4, None)
    assert st == _abi.ERROR_UNSUPPORTED
    st = lib.acpp_reduce_f64(g.handle, None, 8, 0, ctypes.addressof(out), 0.0, 6, None)
    assert st == _abi.ERROR_UNSUPPORTED
    # unknown operator code
    st = lib.acpp_reduce_i32(g.handle, None, 8, 0, ctypes.addressof(out), 0, 42, None)
    assert st == _abi.ERROR_INVALID_ARGUMENT
    g.close()
    cache.close()


def test_empty_input_touches_nothing():
    # numeric.hpp:165-169: n == 0 returns without writing *out; no device call
    lib = _abi.load()
    cache = acpp.AllocationCache()
    g = acpp.AllocationGroup(cache, 0)
    out = ctypes.c_int32(1234)
    assert lib.acpp_reduce_i32(g.handle, None, None, 0, ctypes.addressof(out), 5, 0, None) == 0
    assert out.value == 1234
    assert lib.acpp_inclusive_scan_i64(g.handle, None, None, 0, None, 0, None, None) == 0
    assert lib.acpp_exclusive_scan_u32(g.handle, None, None, 0, None, 9, 0, None) == 0
    assert lib.acpp_transform_reduce_dot_f32(g.handle, None, None, 0, None,
                                             ctypes.addressof(out), 0.0, None) == 0
    g.close()
    cache.close()


def test_allocation_cache_lifecycle_without_device_work():
    lib = _abi.load()
    h = ctypes.c_void_p()
    assert lib.acpp_allocation_cache_create(7, ctypes.byref(h)) == _abi.ERROR_INVALID_ARGUMENT
    for kind in (acpp.ALLOCATION_DEVICE, acpp.ALLOCATION_SHARED, acpp.ALLOCATION_HOST):
        c = acpp.AllocationCache(kind)
        g = acpp.AllocationGroup(c, 0)
        g.release()
        g.close()
        c.purge()
        c.close()
    c = acpp.AllocationCache()
    g = ctypes.c_void_p()
    assert lib.acpp_allocation_group_create(c.handle, -1, ctypes.byref(g)) == \
        _abi.ERROR_INVALID_ARGUMENT
    c.close()


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_abi, "_lib", None)
    monkeypatch.setattr(_abi, "LIB_PATH", "/nonexistent/libacpp_mi355x.so")
    with pytest.raises(ImportError):
        _abi.load()


def test_copy_if_rejects_bf16_clearly():
    # copy_if exists for the integer types, f32 and f64; the Python mirror
    # says so before any library lookup (no device work: CPU tensors suffice).
    torch = pytest.importorskip("torch")
    x = torch.zeros(8, dtype=torch.bfloat16)
    for call in (lambda: acpp.copy_if(None, None, x, x, acpp.PRED_GT, 0),
                 lambda: acpp.copy_if_stencil(None, None, x, torch.zeros(8, dtype=torch.uint8), x)):
        with pytest.raises(TypeError, match="integer, f32 or f64"):
            call()
