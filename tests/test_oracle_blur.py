"""Pins of the blur oracle (GpuBlurPipelineManager.blur, blur.wgsl:13-48)
against closed forms computed independently in FP64.  Values on either side
of an 8-bit rounding boundary are compared only when the FP64 value is at
least 1e-4 (in units of 1/255) away from it."""
import numpy as np
import pytest


def _srgb_to_lin(c):
    c = np.asarray(c, np.float64)
    return np.where(c > 0.04045, ((c + 0.055) / 1.055) ** 2.4, c / 12.92)


def _lin_to_srgb(c):
    c = np.maximum(np.asarray(c, np.float64), 0.0)
    return np.where(c > 0.0031308, 1.055 * c ** (1 / 2.4) - 0.055, c * 12.92)


def _q8(x):
    return np.rint(np.clip(x, 0.0, 1.0) * 255.0) / 255.0


def _safe(x):
    """True where the FP64 value is not within 1e-4 LSB of a rounding boundary."""
    f = np.clip(x, 0.0, 1.0) * 255.0
    return np.abs(f - np.floor(f) - 0.5) > 1e-4


def _weights(radius, sigma):
    i = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-(i * i) / (2.0 * sigma * sigma))
    return w / w.sum()


@pytest.mark.parametrize("srgb", [False, True])
def test_constant_images(oracle, srgb):
    k = np.arange(256, dtype=np.float64)
    out = []
    for kk in range(256):
        im = np.full((2, 3, 4), kk / 255.0, np.float32)
        out.append(oracle.blur_f32(im, 3, 2, radius=4, is_srgb=srgb)[0, 0, 0])
    got = np.array(out, np.float64)
    if srgb:
        lin = _srgb_to_lin(k / 255.0)
        mid = _q8(lin)
        exp = _q8(_lin_to_srgb(mid))
        ok = _safe(lin) & _safe(_lin_to_srgb(mid))
    else:
        exp = k / 255.0
        ok = np.ones(256, bool)
    assert ok.sum() > 240
    assert np.array_equal(got[ok].astype(np.float32), exp[ok].astype(np.float32))
    if srgb:   # the temp is 8-bit linear: dark values do not survive the round trip
        assert got[10] != np.float32(10 / 255.0)


def test_impulse_response_is_separable_gaussian(oracle):
    W = H = 21
    r, s = 4, 2.0
    img = np.zeros((H, W, 4), np.float32)
    img[10, 10, :3] = 1.0
    out = oracle.blur_f32(img, W, H, radius=r, sigma=s, is_srgb=False)
    w = _weights(r, s)
    row = np.zeros(W)
    row[10 - r:10 + r + 1] = w
    temp = _q8(row)
    col = np.zeros(H)
    col[10 - r:10 + r + 1] = w
    exp = np.outer(col, temp)
    ok = _safe(row)[None, :] & _safe(exp)
    got = out[..., 0].astype(np.float64)
    assert np.array_equal(got[ok], _q8(exp)[ok].astype(np.float32).astype(np.float64))
    assert np.all(out[..., 3] == 1.0)
    assert np.array_equal(out[..., 0], out[..., 1])


def test_clamp_to_edge(oracle):
    W, H = 16, 4
    r, s = 3, 1.5
    img = np.zeros((H, W, 4), np.float32)
    img[:, 0, :3] = 1.0
    out = oracle.blur_f32(img, W, H, radius=r, sigma=s, is_srgb=False)
    w = _weights(r, s)   # index 0 is i = -r
    # pixel x: taps x+i clamped; white where x+i <= 0
    exp = np.array([w[np.arange(-r, r + 1) + x <= 0].sum() for x in range(W)])
    got = out[0, :, 0].astype(np.float64)
    ok = _safe(exp)
    assert np.array_equal(got[ok], _q8(exp)[ok].astype(np.float32).astype(np.float64))
    assert np.all(out[1:] == out[0:1])


def test_mirror_symmetry(oracle):
    rng = np.random.default_rng(0)
    W, H = 33, 17
    img = (rng.integers(0, 256, (H, W, 4)) / 255.0).astype(np.float32)
    a = oracle.blur_f32(img, W, H, radius=5, is_srgb=True)
    b = oracle.blur_f32(np.ascontiguousarray(img[:, ::-1]), W, H, radius=5, is_srgb=True)
    assert np.array_equal(a[:, ::-1], b)
