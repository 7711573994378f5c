"""le2d_diff_norms_noalloc: the norms without any device allocation (partials
in the library's fixed device array) give the bits of le2d_diff_norms, on
two calls in a row, the second on a smaller grid, and with no allocation
made through the device allocator hook."""
import ctypes

import numpy as np
import pytest

import le2d_amd

pytestmark = pytest.mark.gpu


def _views(nx, ny, seed):
    rng = np.random.default_rng(seed)
    g0 = rng.standard_normal((ny, nx, 5))
    a = le2d_amd.Task(nx, ny, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5))
    b = le2d_amd.Task(nx, ny, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), dtype=np.float32)
    a.load(g0)
    b.load(g0.astype(np.float32))
    a.step_n(3)
    b.step_n(3)
    return a, b


def _norms(fn, va, vb):
    lib = le2d_amd.load(np.float64)
    out = (ctypes.c_double * 4)()
    rc = getattr(lib, fn)(ctypes.byref(va), ctypes.byref(vb), out)
    assert rc == 0, (fn, rc, lib.le2d_last_error().decode())
    return np.array(out[:])


def test_noalloc_norms_equal_the_allocating_entry_point():
    calls = []
    alloc = le2d_amd.ALLOC_FN(lambda p, n, c: calls.append(n) or 1)  # refuses, and records
    free = le2d_amd.FREE_FN(lambda p, c: None)
    for nx, ny, seed in ((700, 523, 1), (97, 41, 2)):  # the second call is smaller
        a, b = _views(nx, ny, seed)
        try:
            va, vb = a.device_view(), b.device_view()
            ref = _norms("le2d_diff_norms", vb, va)
            le2d_amd.set_device_allocator(alloc, free)
            try:
                got = _norms("le2d_diff_norms_noalloc", vb, va)
                same = _norms("le2d_diff_norms_noalloc", va, va)
            finally:
                le2d_amd.set_device_allocator(None, None)
            assert not calls  # nothing was allocated
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (nx, got, ref)
            assert same[0] == 0.0 and same[2] == 0.0 and same[1] == ref[1]
            assert 0 < got[0] / got[1] < 1e-5
        finally:
            a.close()
            b.close()


def test_noalloc_rejects_mismatched_views():
    a, b = _views(40, 30, 3)
    c, d = _views(41, 30, 4)
    try:
        lib = le2d_amd.load(np.float64)
        out = (ctypes.c_double * 4)()
        va, vc = a.device_view(), c.device_view()
        assert lib.le2d_diff_norms_noalloc(ctypes.byref(va), ctypes.byref(vc), out) == \
            le2d_amd.ERR_ARG
    finally:
        for t in (a, b, c, d):
            t.close()
