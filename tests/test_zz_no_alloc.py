"""Allocation-free dcurlySadjc (htamd_bspline_d1_adj_noalloc) against the
workspace form htamd_bspline_d1_adj, both called through the C ABI.

The two share their kernels, so the bar is bit-for-bit equality.  Two calls in
a row, the second with fewer angles and another geometry, check that the
shared device scratch keeps no state that leaks from one call into the next;
the result is also checked against the FP64 oracle (relative L2 <= 1e-5 for
out1, 1e-4 for the out2 reductions).
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import shift as osh

pytestmark = pytest.mark.gpu


def _run(lib, plan, c, g, r, m, B, noalloc):
    from holotomocupy_amd import _lib
    out1 = torch.empty_like(c)
    out2 = torch.empty((B, 2), dtype=torch.float32, device=c.device)
    st = _lib.stream_ptr()
    if noalloc:
        _lib.check(lib.htamd_bspline_d1_adj_noalloc(plan, _lib.ptr(c), _lib.ptr(g), _lib.ptr(r), 2, _lib.ptr(m), 1,
                                                    _lib.ptr(out1), _lib.ptr(out2), B, st),
                   "htamd_bspline_d1_adj_noalloc")
    else:
        nbytes = ctypes.c_int64()
        _lib.call("htamd_bspline_d1_adj_workspace_size", plan, B, ctypes.byref(nbytes))
        ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=c.device)
        _lib.check(lib.htamd_bspline_d1_adj(plan, _lib.ptr(c), _lib.ptr(g), _lib.ptr(r), 2, _lib.ptr(m), 1,
                                            _lib.ptr(out1), _lib.ptr(out2), _lib.ptr(ws), nbytes.value, B, st),
                   "htamd_bspline_d1_adj")
    torch.cuda.synchronize()
    return out1.cpu().numpy(), out2.cpu().numpy()


def _case(lib, geom, seed):
    from holotomocupy_amd import _lib
    n, npsi, nz, nzpsi, B = geom
    rng = np.random.default_rng(seed)
    plan = ctypes.c_void_p()
    _lib.call("htamd_shift_plan_create", n, npsi, nz, nzpsi, 1, ctypes.byref(plan))
    try:
        c = (rng.standard_normal((B, nzpsi, npsi)) + 1j * rng.standard_normal((B, nzpsi, npsi))).astype(np.complex64)
        g = (rng.standard_normal((B, nz, n)) + 1j * rng.standard_normal((B, nz, n))).astype(np.complex64)
        r = rng.uniform(-4, 4, (B, 2)).astype(np.float32)
        m = rng.uniform(0.95, 1.3, B).astype(np.float32)
        dev = [torch.from_numpy(a).cuda() for a in (c, g, r, m)]
        ref1, ref2 = _run(lib, plan, *dev, B, noalloc=False)
        got1, got2 = _run(lib, plan, *dev, B, noalloc=True)
    finally:
        lib.htamd_shift_plan_destroy(plan)
    assert np.array_equal(got1, ref1) and np.array_equal(got2, ref2), geom
    o1, o2 = osh.Shift(n, npsi, nz, nzpsi, "complex64").dcurlySadjc(c, r, m, g)
    assert np.linalg.norm(got1 - o1) / np.linalg.norm(o1) < 1e-5
    assert np.linalg.norm(got2 - o2) / np.linalg.norm(o2) < 1e-4
    return got2


def test_noalloc_matches_workspace_form_twice(gpu):
    from holotomocupy_amd import _lib
    lib = _lib.load()
    torch.cuda.set_device(gpu)
    a = _case(lib, (96, 80, 64, 72, 6), 11)
    b = _case(lib, (48, 40, 40, 36, 2), 12)  # second call: smaller input
    assert a.shape == (6, 2) and b.shape == (2, 2)
