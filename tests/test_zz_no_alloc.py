"""btg_raster_device_noalloc: rasterization of device polygons into a device
image with no allocation (the scan scratch is a static device array). Its
images must equal btg_raster's bit for bit, on two calls in a row (the second
smaller, over the scratch the first left behind), for ids and for masks, and
a polygon count above the static maximum is refused.

Device buffers for the inputs and outputs are allocated by the test through
the HIP runtime the library itself is linked against (libamdhip64.so.7: a
second runtime in the process, such as torch's own, cannot share its device
pointers); the library is called through its C ABI."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_scene

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    from beamtrace2007_amd import gpu
    gpu.lib()  # loads libbtg.so and with it the runtime below
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


class DevBuf:
    def __init__(self, L, nbytes):
        self.L, self.n = L, max(1, nbytes)
        self.p = C.c_void_p()
        assert L.hipMalloc(C.byref(self.p), self.n) == 0

    def put(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.L.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.L.hipFree(self.p)


def test_noalloc_raster_equals_btg_raster(require_gpu):
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("room")
    r = GpuRenderer(s)
    L = hip()
    W, H = 320, 240
    cam = camera(sp["view"], W, H)
    prim, _ = r.primary(cam, W, H, tile=8)
    shadow, _ = r.point_shadows(cam, W, H, prim, sp["point_light"])
    assert len(prim) > 3 * 1024 and len(shadow) > 0  # several scan tiles
    cases = [(cam, W, H, prim, False), (cam, W, H, shadow, True)]
    W2, H2 = 96, 64
    cam2 = camera(sp["view"], W2, H2)
    prim2, _ = r.primary(cam2, W2, H2, tile=16)
    cases.append((cam2, W2, H2, prim2[:100], False))  # smaller input after larger ones
    for c, w, h, polys, shade in cases:
        print(f"noalloc raster {w}x{h}, {len(polys)} polygons, shadow {shade}", flush=True)
        want = r.raster(c, w, h, polys, shadow=shade)
        dtype = np.uint8 if shade else np.int32
        dev_polys, out = DevBuf(L, polys.nbytes), DevBuf(L, h * w * np.dtype(dtype).itemsize)
        dev_polys.put(polys)
        out.put(np.full(h * w, 7, dtype))  # every pixel must be written
        r.raster_device_noalloc(c, w, h, dev_polys.p.value, len(polys), out.p.value, shadow=shade)
        got = out.get(dtype, h * w).reshape(h, w)
        dev_polys.free()
        out.free()
        assert np.array_equal(got, want), (w, h, len(polys), shade)
    dev_polys, out = DevBuf(L, prim.nbytes), DevBuf(L, W * H * 4)
    with pytest.raises(RuntimeError, match="at most 1048576 polygons"):
        r.raster_device_noalloc(cam, W, H, dev_polys.p.value, (1 << 20) + 1, out.p.value)
    dev_polys.free()
    out.free()
    r.close()
