"""espec_fcorrel_noalloc: FCORREL (fcorrel.f:1-68) on device buffers with
no allocation of its own, against espec_fcorrel on the same data.

Both run dot_kernel with the same block count, so the sums must be
bit-identical.  Two calls in a row, the second smaller, check that the
scratch array's ticket and partials carry nothing from one call into the
next; a size above the stated maximum is refused before anything runs.
Device buffers come from hipMalloc of the HIP runtime that
libespec_hip.so itself loaded (libamdhip64.so.7)."""
import ctypes as C

import numpy as np
import pytest

import espec_amd as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(gpu):
    E.lib()  # loads libamdhip64.so.7 as a dependency; dlopen returns that copy
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def dev(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
    if a.nbytes:
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def test_fcorrel_noalloc_matches_fcorrel(hip):
    L = E.lib()
    f = L.espec_fcorrel_noalloc
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(11)
    for n in (300_001, 1_000):  # the second call is smaller
        print(f"case: fcorrel_noalloc n={n}", flush=True)
        u1, u2 = rng.standard_normal(n), rng.standard_normal(n)
        d1, d2, do = dev(hip, u1), dev(hip, u2), dev(hip, np.zeros(1))
        try:
            assert f(n, d1, d2, do, None) == 0
            assert hip.hipDeviceSynchronize() == 0
            got = np.zeros(1)
            assert hip.hipMemcpy(got.ctypes.data, do, 8, 2) == 0  # device to host
        finally:
            for p in (d1, d2, do):
                hip.hipFree(p)
        ref = E.fcorrel(u1, u2)
        assert got[0] == ref, (n, got[0], ref)
        assert abs(got[0] - float(np.dot(u1, u2))) <= 1e-12 * np.abs(u1 * u2).sum()


def test_fcorrel_noalloc_refuses_oversize(hip):
    L = E.lib()
    f = L.espec_fcorrel_noalloc
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    # refused on the size check, before the pointers are used
    assert f((1 << 30) + 1, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), None) == 2  # ESPEC_ERR_ARG
