"""npb_blur_dev_noalloc (the entry point that allocates nothing) against
npb_blur_dev on the same device buffers, through the C ABI: bit-identical
output on two calls in a row, the second with a smaller image, in both colour
spaces; a size above the fixed scratch is rejected before any launch.

Device buffers are the caller's: allocated here with the HIP runtime the
library itself is linked against."""
import ctypes as C

import numpy as np
import pytest

from nprush_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p


def test_noalloc_blur_equals_blur_dev():
    lib = _lib.load()
    hip = _hip()
    b = C.c_void_p()
    assert lib.npb_create(0, C.byref(b)) == 0
    rng = np.random.default_rng(3)
    # largest first, then a smaller one: the second call reuses the scratch
    # that the first left filled
    cases = [(1920, 1080, 16, 8.0, 1), (455, 256, 16, 8.0, 0), (37, 19, 3, 1.5, 1)]
    cap = 1920 * 1080 * 16
    src, dst_a, dst_b = _dev(hip, cap), _dev(hip, cap), _dev(hip, cap)
    try:
        for (W, H, radius, sigma, srgb) in cases:
            print(f"case {W}x{H} radius {radius} srgb {srgb}", flush=True)
            img = (rng.integers(0, 256, (H, W, 4)) / 255.0).astype(np.float32)
            nb = img.nbytes
            assert hip.hipMemcpy(src, img.ctypes.data_as(C.c_void_p), nb, H2D) == 0
            assert lib.npb_blur_dev(b, src, dst_a, W, H, radius, sigma, srgb) == 0
            assert lib.npb_blur_dev_noalloc(b, src, dst_b, W, H, radius, sigma, srgb) == 0, \
                lib.npb_last_error(b)
            a = np.empty_like(img)
            c = np.empty_like(img)
            assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), dst_a, nb, D2H) == 0
            assert hip.hipMemcpy(c.ctypes.data_as(C.c_void_p), dst_b, nb, D2H) == 0
            assert np.array_equal(a, c), (W, H)
            assert np.all(a[..., 3] == 1.0) and a[..., :3].std() > 0
        # above the fixed scratch: rejected on the host, nothing launched
        rc = lib.npb_blur_dev_noalloc(b, src, dst_b, 2048, 2048, 16, 8.0, 1)
        assert rc == -1   # NPS_ERR_ARG
        assert b"exceed" in lib.npb_last_error(b)
    finally:
        for p in (src, dst_a, dst_b):
            hip.hipFree(p)
        lib.npb_destroy(b)
