"""fgpu_reduce_loads_noalloc (reduced load vectors with every work buffer in
the library's static scratch array, nothing allocated) through the C ABI
against fgpu_reduce_loads on the same resident reduction: bitwise equal on
two calls in a row, the second with fewer load cases, and a load-case count
beyond the scratch size refused."""
import ctypes

import numpy as np
import pytest

from fedempy_amd import gpu, parts

pytestmark = pytest.mark.gpu


def _hip():
    """the HIP runtime the library itself links (already loaded with it)"""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


class DevArray:
    H2D, D2H = 1, 2

    def __init__(self, hip, nbytes):
        self.hip, self.p, self.n = hip, ctypes.c_void_p(), nbytes
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.asfortranarray(a, np.float64)
        assert a.nbytes == self.n and self.hip.hipMemcpy(self.p, a.ctypes.data, self.n, self.H2D) == 0

    def get(self, shape):
        a = np.zeros(shape, order="F")
        assert a.nbytes == self.n and self.hip.hipMemcpy(a.ctypes.data, self.p, self.n, self.D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_reduced_loads_match_the_allocating_entry_point():
    part = parts.hex_block(8, 6, 5, 0.08, 0.06, 0.05, seed=3, n_ext=6)
    R = gpu.GpuReducer()
    R.set_part(part)
    sizes = R.analyze(ngen=4)
    R.reduce()
    lib = R.lib
    hip = _hip()
    lib.fgpu_reduce_loads.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.fgpu_reduce_loads_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    neq, nd = sizes["neq"], sizes["ndim"]
    rng = np.random.default_rng(4)
    for nrhs in (7, 3):  # the second call smaller: stale scratch contents must not matter
        print("noalloc reduced loads: %d load cases" % nrhs, flush=True)
        F = np.asfortranarray(rng.normal(0.0, 1e3, (neq, nrhs)))
        ref = np.zeros((nd, nrhs), order="F")
        assert lib.fgpu_reduce_loads(R.h, nrhs, F.ctypes.data, ref.ctypes.data) == 0
        Fd, Fr = DevArray(hip, F.nbytes), DevArray(hip, ref.nbytes)
        Fd.put(F)
        Fr.put(np.full((nd, nrhs), np.nan))
        assert lib.fgpu_reduce_loads_noalloc(R.h, nrhs, Fd.p, Fr.p) == 0
        got = Fr.get((nd, nrhs))
        Fd.free()
        Fr.free()
        assert np.array_equal(got, ref)
    huge = 4096
    Fd, Fr = DevArray(hip, 8 * huge * neq), DevArray(hip, 8 * huge * nd)
    assert lib.fgpu_reduce_loads_noalloc(R.h, huge, Fd.p, Fr.p) == -1
    assert b"scratch" in lib.fgpu_last_error(R.h)
    Fd.free()
    Fr.free()
