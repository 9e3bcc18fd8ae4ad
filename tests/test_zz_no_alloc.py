"""The no-allocation pair-table entry point (csx_debye_pair_tables_noalloc)
against csx_debye_pair_tables, through the C ABI.

Device buffers come from the HIP runtime the library itself links
(libamdhip64 from ROCm), so the test and the library share one runtime.
Bar: bitwise equal tables (the same kernel sums in the same order), on two
calls in a row, the second with a smaller structure and q grid.
"""
import ctypes as C

import numpy as np
import pytest

import crabsaxs_amd as csx
from crabsaxs_amd import _lib

pytestmark = pytest.mark.gpu


def hip():
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    L.hipStreamDestroy.argtypes = [C.c_void_p]
    L.hipStreamSynchronize.argtypes = [C.c_void_p]
    return L


D2H = 2


def noalloc_tables(ctx, s, q, ntypes, stream):
    H = hip()
    nb = ntypes * (ntypes + 1) // 2
    nbytes = nb * q.size * 8
    dev = C.c_void_p()
    assert H.hipMalloc(C.byref(dev), nbytes) == 0
    try:
        lib = _lib.load()
        st = lib.csx_debye_pair_tables_noalloc(s.handle(ctx), _lib.dptr(q), q.size, dev, stream)
        if st != 0:
            raise csx.InvalidInput(lib.csx_last_error().decode())
        assert H.hipStreamSynchronize(stream) == 0
        out = np.empty(nb * q.size)
        assert H.hipMemcpy(out.ctypes.data_as(C.c_void_p), dev, nbytes, D2H) == 0
        return out.reshape(nb, q.size)
    finally:
        H.hipFree(dev)


def test_noalloc_matches_pair_tables_twice(gpu_ctx):
    H = hip()
    stream = C.c_void_p()
    assert H.hipStreamCreate(C.byref(stream)) == 0
    try:
        cases = [(csx.generate_globule(6000, 24.0, 1.0, seed=91), np.linspace(0.0, 1.0, 300)),
                 (csx.generate_globule(1500, 14.0, 1.0, seed=92), np.linspace(0.0, 0.5, 64))]
        for s, q in cases:
            want, _ = csx.debye_pair_tables(s, q)
            T = int(round((np.sqrt(8 * want.shape[0] + 1) - 1) / 2))
            got = noalloc_tables(gpu_ctx, s, q, T, stream)
            print("noalloc", len(s), q.size, np.max(np.abs(got - want)), flush=True)
            assert got.shape == want.shape
            assert np.array_equal(got, want)
        # the null stream
        s, q = cases[1]
        want, _ = csx.debye_pair_tables(s, q)
        T = int(round((np.sqrt(8 * want.shape[0] + 1) - 1) / 2))
        assert np.array_equal(noalloc_tables(gpu_ctx, s, q, T, None), want)
    finally:
        H.hipStreamDestroy(stream)


def test_noalloc_rejects_what_the_scratch_cannot_hold(gpu_ctx):
    big = csx.generate_globule(90000, 60.0, 1.0, seed=93)
    q = np.linspace(0.0, 0.5, 500)
    with pytest.raises(csx.InvalidInput, match="scratch"):
        noalloc_tables(gpu_ctx, big, q, 4, None)
    s = csx.generate_globule(800, 12.0, 1.0, seed=94)
    s.handle(gpu_ctx)
    try:
        gpu_ctx.set_tuning(0, 2, 1024)  # no work items for this j chunk yet
        with pytest.raises(csx.InvalidInput, match="j chunk"):
            noalloc_tables(gpu_ctx, s, q[:50], 4, None)
        want, _ = csx.debye_pair_tables(s, q[:50])  # makes them
        T = int(round((np.sqrt(8 * want.shape[0] + 1) - 1) / 2))
        assert np.array_equal(noalloc_tables(gpu_ctx, s, q[:50], T, None), want)
    finally:
        gpu_ctx.set_tuning()
