"""vt_sensor_raw_device, the raw sensor stage that allocates nothing (glare levels in the library's static device
array), against the allocating vt_sensor_test on the same inputs: DN bit-identical on two calls in a row, the
second smaller than the first, with glare on; oversize glare levels and host memory are refused on the host.

The test's own device buffers come from the HIP runtime libvt.so links (the caller's allocations, not the entry
point's); torch is not used, since its wheel carries a HIP runtime of its own that does not start once this one
has."""
import ctypes as C

import numpy as np
import pytest

from vtwin import lib, scenes

pytestmark = pytest.mark.gpu


def _case(w, h, seed, level):
    r = scenes.Rng(seed)
    rad = np.zeros((h, w, 4), np.float32)
    rad[:, :, 0:3] = r.uniform(0.0, level, w * h * 3).reshape(h, w, 3)
    rad[:, :, 3] = 1.0
    ri = r.uniform(0.3, 1.0, w * h).reshape(h, w).astype(np.float32)
    return rad, ri


def _params(w, h, glare_level, cfa=0, frame=3):
    p = scenes.default_sensor(cfa=cfa, exposure=2e-3, noise=True, frame=frame, seed=1006)
    p.update(glare_fraction=0.1, glare_sigma=0.25 * float(np.hypot(w, h)) / glare_level, glare_level=glare_level)
    return p


class DeviceBuffers:
    """hipMalloc'd buffers of the runtime libvt.so uses, freed on close"""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0        # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0    # device to host
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def test_sensor_raw_device_matches_the_allocating_stage():
    ctx = lib.Context(0)
    dev = DeviceBuffers()
    try:
        for (w, h, F, cfa, seed) in [(203, 157, 4, 0, 11), (96, 64, 2, 1, 12)]:
            print("sensor_raw_device case %dx%d, glare level 1/%d, cfa %d" % (w, h, F, cfa), flush=True)
            rad, ri = _case(w, h, seed, 3.0)
            p = _params(w, h, F, cfa)
            want = ctx.sensor_test(p, rad, ri)
            trad, tri = dev.upload(rad), dev.upload(ri)
            tdn = dev.upload(np.full((h, w), 0xFFFF, np.uint16))
            ctx.sensor_raw_device(p, w, h, trad, tri, tdn)
            ctx.sync()
            got = dev.download(tdn, (h, w), np.uint16)
            assert got.shape == want.shape
            diff = np.count_nonzero(got != want)
            assert diff == 0, "%d of %d DNs differ" % (diff, got.size)
            assert want.max() > 0
        # refused on the host, before any launch
        with pytest.raises(lib.VTError) as e:
            ctx.sensor_raw_device(_params(4096, 4096, 1), 4096, 4096, trad, tri, tdn)
        assert e.value.code == lib.VT_ERR_SIZE and "glare level" in str(e.value)
        with pytest.raises(lib.VTError) as e:
            ctx.sensor_raw_device(p, w, h, rad.ctypes.data, tri, tdn)
        assert e.value.code == lib.VT_ERR_ARG and "not device memory" in str(e.value)
        # the context still works after the refusals
        ctx.sensor_raw_device(p, w, h, trad, tri, tdn)
        ctx.sync()
        assert np.array_equal(dev.download(tdn, (h, w), np.uint16), want)
    finally:
        dev.close()
        ctx.close()
