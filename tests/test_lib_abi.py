"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/htamd.h declares, with the argument counts the Python binding uses."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "htamd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(htamd_\w+)\s*\(([^;]*?)\);", text, flags=re.M | re.S)


@pytest.fixture(scope="module")
def lib():
    from holotomocupy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from holotomocupy_amd.build import build
        build()
    return _lib.load()


def test_header_parses():
    names = [n for n, _ in declared()]
    assert len(names) >= 15 and "htamd_prop_apply" in names and "htamd_last_error" in names


def test_every_declared_symbol_is_exported(lib):
    for name, _ in declared():
        assert hasattr(lib, name), name


def test_binding_arities_match_header(lib):
    from holotomocupy_amd import _lib
    for name, args in declared():
        args = args.strip()
        n = 0 if args in ("", "void") else args.count(",") + 1
        if name in _lib.SIGNATURES:
            assert len(_lib.SIGNATURES[name]) == n, name


def test_abi_version_and_error_path(lib):
    assert lib.htamd_abi_version() == 1
    plan = ctypes.c_void_p()
    import numpy as np
    d = np.array([1.0])
    from holotomocupy_amd import _lib
    rc = lib.htamd_prop_plan_create(13, 8, 1, 1.0, 1.0, d.ctypes.data_as(_lib.c_dp), ctypes.byref(plan))
    assert rc == -2 and b"even sizes" in lib.htamd_last_error()
    rc = lib.htamd_prop_plan_create(16, 8200, 1, 1.0, 1.0, d.ctypes.data_as(_lib.c_dp), ctypes.byref(plan))
    assert rc == -2
    theta = np.linspace(0, np.pi, 8, endpoint=False)
    rc = lib.htamd_tomo_plan_create(13, 8, theta.ctypes.data_as(_lib.c_dp), None, ctypes.byref(plan))
    assert rc == -2 and b"even" in lib.htamd_last_error()
    rc = lib.htamd_tomo_plan_create(4098, 8, theta.ctypes.data_as(_lib.c_dp), None, ctypes.byref(plan))
    assert rc == -2
    rc = lib.htamd_shift_plan_create(0, 8, 8, 8, 1, ctypes.byref(plan))
    assert rc == -1


def test_noalloc_d1_adj_rejects_batches_beyond_its_scratch(lib):
    """The allocation-free dcurlySadjc refuses, on the host and before it looks
    up its device scratch, a batch whose partial sums exceed the fixed maximum
    (no GPU is touched: a 3000-wide plan uploads no FFT tables and the buffer
    arguments are never dereferenced)."""
    from holotomocupy_amd import _lib
    cap = ctypes.c_int64()
    assert lib.htamd_bspline_d1_adj_noalloc_max_bytes(ctypes.byref(cap)) == 0 and cap.value == 4 << 20
    plan = ctypes.c_void_p()
    assert lib.htamd_shift_plan_create(3000, 3000, 3000, 3000, 1, ctypes.byref(plan)) == 0
    try:
        per = ctypes.c_int64()
        assert lib.htamd_bspline_d1_adj_workspace_size(plan, 1, ctypes.byref(per)) == 0
        B = cap.value // per.value + 1
        fake = _lib.c_vp(256)
        rc = lib.htamd_bspline_d1_adj_noalloc(plan, fake, fake, fake, 2, fake, 1, fake, fake, B, None)
        assert rc == -2 and b"fixed maximum" in lib.htamd_last_error()
    finally:
        lib.htamd_shift_plan_destroy(plan)
