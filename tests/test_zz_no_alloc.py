"""The allocation-free entry point les_zero_padding_noalloc against the
existing les_zero_padding (108_01_padd.f90:73-132), plane by plane, through
the C-ABI: two calls in a row on device buffers, the second with fewer
planes, each equal to the existing entry point's result bit for bit (the
same kernels run); a request beyond the fixed scratch is refused with
LES_E_SIZE.  Device buffers come from the HIP runtime the library itself
links (libamdhip64), so that one runtime owns every pointer."""
import ctypes as C

import numpy as np
import pytest

from overlappingbls_amd import config

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime already loaded."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        for f in (self.rt.hipMalloc, self.rt.hipFree, self.rt.hipMemcpy, self.rt.hipDeviceSynchronize):
            f.restype = C.c_int

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dev, host):
        assert self.rt.hipMemcpy(dev, host.ctypes.data, host.nbytes, H2D) == 0

    def get(self, host, dev):
        assert self.rt.hipMemcpy(host.ctypes.data, dev, host.nbytes, D2H) == 0

    def free(self, p):
        assert self.rt.hipFree(p) == 0


def pad_noalloc(g, hip, planes, ld):
    """planes: (n, ny, nx) float64 -> (n, 3ny/2, 3nx/2) via the device entry point."""
    n, ny, nx = planes.shape
    host = np.zeros((n, ny, ld))
    host[:, :, :nx] = planes
    host[:, :, nx:] = 7.0e30          # padding columns are never read
    out = np.full((n, 3 * ny // 2, 3 * nx // 2), np.nan)
    d_var, d_big = hip.alloc(host.nbytes), hip.alloc(out.nbytes)
    try:
        hip.put(d_var, host)
        hip.put(d_big, out)
        st = g._f["zero_padding_noalloc"](g._h, d_var, ld, d_big, n)
        if st == 0:
            hip.get(out, d_big)
    finally:
        hip.free(d_var)
        hip.free(d_big)
    return st, out


@pytest.mark.parametrize("case", ["tiny_ragged", "wide"])
def test_noalloc_padding_matches_the_existing_entry_point(case, gpu_lib):
    from overlappingbls_amd.abi import LES_E_SIZE, LES_OK
    nl = config.workload(case)
    g = gpu_lib.create(nl.params())
    hip = Hip()
    ny, nx = g.ny, g.nx
    ld = nx + 2                        # the reference's ldx
    for call, nplanes in enumerate((5, 2)):
        planes = config.uniform(40 + call, 1, (nplanes, ny, nx))
        st, big = pad_noalloc(g, hip, planes, ld)
        assert st == LES_OK, st
        for k in range(nplanes):
            ref = g.zero_padding(planes[k])
            assert np.array_equal(big[k], ref), (call, k)
    # beyond the 48 MiB scratch: refused before anything is launched
    too_many = (6 << 20) // (ny * (3 * nx // 2)) + 1
    st = g._f["zero_padding_noalloc"](g._h, C.c_void_p(16), ld, C.c_void_p(16), too_many)
    assert st == LES_E_SIZE
