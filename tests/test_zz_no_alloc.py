"""bemx_linsolve_device_noalloc, the FP64 LINSOLVE entry point that
allocates nothing (its workspace is a static device array), against
bemx_linsolve (GPU): two calls in a row, the second smaller, each
bit-identical to bemx_linsolve's FP64 solution of the same system (they run
the same kernels in the same order) and within 1e-12 of LAPACK; orders
beyond the array's maximum and a short leading dimension are refused.
Device memory for the system comes from the HIP runtime through ctypes."""
import ctypes as C
import os

import numpy as np
import pytest

from bemx import api

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    for path in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            lib = C.CDLL(path)
            break
        except OSError:
            continue
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipSetDevice.argtypes = [C.c_int]
    return lib


def _system(n, seed):
    rng = np.random.default_rng(seed)
    a = 1e-3 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    a[rng.permutation(n), np.arange(n)] += 10.0 + 5j  # every column pivots
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    return np.asfortranarray(a), b


def test_noalloc_solve_matches_linsolve_on_two_calls():
    hip = _hip()
    assert hip.hipSetDevice(0) == 0
    ctx = api.Context(0)
    lib = api.lib()
    try:
        for n, seed in ((700, 1), (333, 2)):
            print("noalloc solve n=%d" % n, flush=True)
            a, b = _system(n, seed)
            da, db = C.c_void_p(), C.c_void_p()
            assert hip.hipMalloc(C.byref(da), a.nbytes) == 0
            assert hip.hipMalloc(C.byref(db), b.nbytes) == 0
            try:
                assert hip.hipMemcpy(da, a.ctypes.data, a.nbytes, H2D) == 0
                assert hip.hipMemcpy(db, b.ctypes.data, b.nbytes, H2D) == 0
                info = C.c_longlong(-1)
                code = lib.bemx_linsolve_device_noalloc(ctx.h, n, da, n, db, C.byref(info))
                assert code == 0, lib.bemx_last_error(ctx.h)
                assert info.value == 0
                x = np.empty_like(b)
                assert hip.hipMemcpy(x.ctypes.data, db, x.nbytes, D2H) == 0
            finally:
                hip.hipFree(da)
                hip.hipFree(db)
            ref = ctx.linsolve(a, b, api.MODE_FP64)
            np.testing.assert_array_equal(x, ref)
            xr = np.linalg.solve(a, b)
            assert np.max(np.abs(x - xr)) / np.max(np.abs(xr)) < 1e-12
        info = C.c_longlong(0)
        dummy = C.c_void_p(16)  # never dereferenced: the sizes are refused first
        assert lib.bemx_linsolve_device_noalloc(ctx.h, 32769, dummy, 32769, dummy,
                                                C.byref(info)) == 3
        assert lib.bemx_linsolve_device_noalloc(ctx.h, 100, dummy, 99, dummy,
                                                C.byref(info)) == 3
    finally:
        ctx.close()
