"""The allocation-free gravity entry point against the allocating one, through the C
ABI: bit-identical wrenches on two calls in a row (the second smaller, so stale
scratch contents from the first must not leak into it), both precisions, and the
size limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from exgame_amd import _lib

pytestmark = pytest.mark.gpu

PREC = {"f64": 0, "f32c": 1}


def charges(n, seed):
    rng = np.random.default_rng(seed)
    c = np.zeros((n, 8))
    c[:, :3] = rng.uniform(-1, 1, (n, 3))
    c[:, 3:6] = rng.uniform(-0.2, 0.2, (n, 3))
    c[:, 6] = rng.lognormal(0.0, 0.5, n) / n
    c[:, 7] = (c[:, 6] / 4.0) ** (1.0 / 3.0)
    return c


def allocating(c, G, prec):
    L = _lib.lib()
    out = np.zeros((c.shape[0], 6))
    _lib.check(L.exg_gravity_accumulate_precision(c.shape[0], _lib.fptr(np.ascontiguousarray(c)), float(G),
                                                  PREC[prec], _lib.fptr(out)))
    return out


def hip():
    """The HIP runtime the library itself runs on (already loaded with it)."""
    _lib.lib()
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    return h


def no_alloc(c, G, prec):
    """Device buffers from the caller (the test), then the allocation-free call."""
    h = hip()
    L = _lib.lib()
    f = L.exg_gravity_accumulate_noalloc
    f.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    n = c.shape[0]
    src = np.ascontiguousarray(c, dtype=np.float64)
    out = np.zeros((n, 6))
    dc, dw = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(dc), src.nbytes) == 0 and h.hipMalloc(C.byref(dw), out.nbytes) == 0
    try:
        assert h.hipMemcpy(dc, src.ctypes.data, src.nbytes, 1) == 0          # host → device
        assert h.hipMemset(dw, 0xFF, out.nbytes) == 0                          # NaN until written
        _lib.check(f(n, dc, float(G), PREC[prec], dw, None))
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(out.ctypes.data, dw, out.nbytes, 2) == 0          # device → host
    finally:
        h.hipFree(dc)
        h.hipFree(dw)
    return out


@pytest.mark.parametrize("prec", ["f64", "f32c"])
def test_no_alloc_gravity_equals_the_allocating_entry_point(prec):
    for n, seed in ((5003, 1), (700, 2)):       # two calls in a row, the second smaller
        c = charges(n, seed)
        got, want = no_alloc(c, 1.3, prec), allocating(c, 1.3, prec)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want), (n, prec, np.abs(got - want).max())


def test_no_alloc_gravity_at_its_size_limit_and_beyond():
    c = charges(65536, 3)
    assert np.array_equal(no_alloc(c, 1.0, "f64"), allocating(c, 1.0, "f64"))
    L = _lib.lib()
    f = L.exg_gravity_accumulate_noalloc
    f.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    # refused on the host before anything runs (the pointers are never touched)
    assert f(65537, 16, 1.0, 0, 16, None) == 6
    assert f(0, 16, 1.0, 0, 16, None) == 6
