"""lumagpu_moments_noalloc (the allocation-free variant of lumagpu_moments) through the C ABI:
its output equals lumagpu_moments' bit for bit on two calls in a row, the second on a smaller
grid (the static scratch keeps the first call's partials), and grids beyond its scratch are
refused.  The output buffer comes from the HIP runtime the library itself links, so the test
touches no second runtime."""
import ctypes as C

import numpy as np
import pytest

from luma_amd import _abi, setups
from luma_amd._abi import EINVAL, LumaGpuError, call
from luma_amd.grid import GpuGrid

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    _abi.load()                       # brings in the library's libamdhip64
    return C.CDLL("libamdhip64.so.7")


class DeviceBuffer:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), H2D) == 0

    def download(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


def reference_moments(g):
    out = np.empty(5)
    call("lumagpu_moments", g.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def test_moments_noalloc_matches_moments_on_two_calls_second_smaller():
    buf = DeviceBuffer(hip_runtime(), 5)
    try:
        for shape, prec in (((64, 40, 36), "fp64"), ((20, 12, 10), "fp32s")):
            print(f"[noalloc] {shape} {prec}", flush=True)
            buf.upload(np.full(5, np.nan))
            case = setups.taylor_green(shape, u0=0.05, tau=0.6, dims=3)
            with GpuGrid.from_case(case, prec) as g:
                g.run(3)
                ref = reference_moments(g)
                call("lumagpu_moments_noalloc", g.h, buf.ptr)
                g.sync()
                got = buf.download()
            assert np.array_equal(got, ref), (shape, got, ref)
    finally:
        buf.free()


def test_moments_noalloc_refuses_grids_beyond_its_scratch():
    # 3 x 1024 x 257: 263,168 row blocks of 256 cells, beyond the 2^18 the scratch holds
    buf = DeviceBuffer(hip_runtime(), 5)
    try:
        buf.upload(np.zeros(5))
        case = setups.taylor_green((3, 1024, 257), u0=0.0, tau=0.6, dims=3)
        with GpuGrid.from_case(case, "fp32s") as g:
            with pytest.raises(LumaGpuError) as e:
                call("lumagpu_moments_noalloc", g.h, buf.ptr)
            assert e.value.code == EINVAL and "too large" in str(e.value)
        assert not buf.download().any()
    finally:
        buf.free()
