"""birp_render_views_noalloc (the entry point that allocates nothing; its fast launches keep their
descriptors and work counters in the library's fixed device array) against birp_render_async, the
entry point it is a variant of: every image bitwise equal, on two calls in a row -- the second with
fewer and smaller views, so the array's contents from the first call must not leak -- for FP32
storage (exact and fast views mixed) and F16 storage, and a clear refusal beyond 64 views. Device
images for the no-alloc calls come from the HIP runtime the library itself links (hipMalloc through
ctypes), so the test owns its outputs as such a caller would."""
import ctypes as C

import numpy as np
import pytest

from birp_cpp_amd import capi
from birp_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def hip():
    L = C.CDLL("libamdhip64.so.7")  # the runtime libbirp_hip.so is linked against (already loaded)
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return L


def views_set(sizes, steps, maths, seed):
    pos = capi.synth_orbit(len(sizes), seed)
    return [W.view(w, h, st, pos=tuple(float(c) for c in p), math=m) for (w, h), st, m, p in zip(sizes, steps, maths, pos)]


def via_async(vol, v):
    img = capi.Image(vol.dev, v.width, v.height)
    try:
        capi.render_async(vol, v, img)
        return img.download()
    finally:
        img.free()


def via_noalloc(H, vol, views):
    bufs = []
    try:
        for v in views:
            p = C.c_void_p()
            assert H.hipMalloc(C.byref(p), v.width * v.height * 4) == 0
            assert H.hipMemset(p, 0xFF, v.width * v.height * 4) == 0  # NaN bits: every pixel must be written
            bufs.append(p)
        arr = (capi.View * len(views))(*views)
        ptrs = (C.c_void_p * len(views))(*[b.value for b in bufs])
        capi.check(capi.lib().birp_render_views_noalloc(vol.h, arr, len(views), ptrs), "birp_render_views_noalloc")
        vol.dev.synchronize()
        out = []
        for v, b in zip(views, bufs):
            host = np.empty((v.height, v.width), np.float32)
            assert H.hipMemcpy(host.ctypes.data, b, v.width * v.height * 4, 2) == 0  # device to host
            out.append(host)
        return out
    finally:
        for b in bufs:
            H.hipFree(b)


@pytest.mark.parametrize("dtype", [capi.DTYPE_F32, capi.DTYPE_F16])
def test_noalloc_views_equal_render_async(device, dtype):
    H = hip()
    vals, org, spc, _ = W.cube_128()
    vol = capi.Volume.upload(device, vals, org, spc, dtype=dtype)
    try:
        F, E = capi.MATH_FAST, capi.MATH_EXACT
        mixed = [F, E, F, F, E, F, F, F, E, F] if dtype == capi.DTYPE_F32 else [F] * 10
        first = views_set([(576, 336), (144, 84), (288, 168), (144, 84), (288, 168), (576, 336), (64, 48), (144, 84), (96, 64), (288, 168)],
                          [0.1, 0.25, 0.25, 0.5, 0.1, 0.25, 0.1, 0.1, 0.5, 0.25], mixed, seed=61)
        second = views_set([(64, 48), (40, 30), (96, 64)], [0.5, 0.25, 0.1], [F, F, E] if dtype == capi.DTYPE_F32 else [F] * 3, seed=62)
        for views in (first, second):  # two calls in a row, the second smaller
            got = via_noalloc(H, vol, views)
            for k, v in enumerate(views):
                want = via_async(vol, v)
                assert np.array_equal(got[k].view(np.uint32), want.view(np.uint32)), (len(views), k)
                assert np.isfinite(got[k]).all()
        too_many = [second[0]] * 65
        arr = (capi.View * 65)(*too_many)
        ptrs = (C.c_void_p * 65)(*([1] * 65))
        assert capi.lib().birp_render_views_noalloc(vol.h, arr, 65, ptrs) == capi.ERR_UNSUPPORTED
        assert b"64 views" in capi.lib().birp_last_error()
    finally:
        vol.free()
