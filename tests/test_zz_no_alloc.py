"""The allocation-free surrogate forward pass (sur_forward_noalloc,
include/dem_surrogate.h) against sur_forward: outputs equal bit for bit on
two calls in a row, the second with fewer rows, and a row count beyond the
limit refused.  The device buffers of the inputs and outputs are the
caller's, allocated here through the HIP runtime the library links to."""
import ctypes as C

import numpy as np
import pytest

import gpucode_amd as G
from gpucode_amd.surrogate import MLP

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DevBuf:
    """Device memory of the caller, through libamdhip64 directly."""
    _hip = None

    def __init__(self, nbytes):
        if DevBuf._hip is None:
            L = C.CDLL("libamdhip64.so.7")
            L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            L.hipFree.argtypes = [C.c_void_p]
            L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            DevBuf._hip = L
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        assert DevBuf._hip.hipMalloc(C.byref(self.ptr), nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert DevBuf._hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, shape, dtype=np.float32):
        a = np.empty(shape, dtype)
        assert DevBuf._hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.ptr:
            DevBuf._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def test_noalloc_forward_equals_forward():
    m = MLP(seed=3)
    rng = np.random.default_rng(0)
    # parameters away from the initial ones, so every layer matters
    m.set_params(m.params() + rng.normal(size=MLP.nparams()).astype(np.float32) * 0.05)
    nmax = 5000
    dx, dy = DevBuf(nmax * 8 * 4), DevBuf(nmax * 2 * 4)
    try:
        for n in (nmax, 1234):
            x = rng.normal(size=(n, 8)).astype(np.float32)
            y_ref = m.forward(x)
            dx.upload(x)
            dy.upload(np.full((nmax, 2), np.nan, np.float32))  # every output must be written
            m.forward_noalloc(n, dx.ptr.value, dy.ptr.value)
            y = dy.download((n, 2))
            assert np.array_equal(y, y_ref), (n, np.abs(y - y_ref).max())
            rest = dy.download((nmax, 2))[n:]
            assert np.isnan(rest).all()  # nothing written past n rows
        with pytest.raises(G.DemError) as e:
            m.forward_noalloc(8193, dx.ptr.value, dy.ptr.value)
        assert e.value.code == -1
    finally:
        dx.free()
        dy.free()
        m.close()
