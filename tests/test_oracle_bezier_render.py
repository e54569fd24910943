"""Pins of the CPU oracle's display raster of the curves (bezier_forward.wgsl
with the blend state of GpuBezierForwardPipelineManager.ts, restated as
ora_bz_render_f32) by closed forms on straight strokes: distance, pressure
taper, smoothstep edge, the brush's v coordinate (signed cross distance) and
repeat addressing in u, draw order, the near cull, the quad's +-4 clamp, load
versus clear and the per-blend 8-bit rounding."""
import numpy as np
import pytest

from test_oracle_bezier import curve, eye_vp, persp_w_is_z, straight


def px_centres(W, H):
    aspect = np.float64(W) / H
    x = ((np.arange(W) + 0.5) / W * 2.0 - 1.0) * aspect
    y = -((np.arange(H) + 0.5) / H * 2.0 - 1.0)
    return np.meshgrid(x, y)  # X[py, px], Y[py, px]


def smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def brush_const(v=255, bw=8, bh=8):
    b = np.zeros((bh, bw, 4), np.uint8)
    b[..., 0] = v
    b[..., 3] = 255
    return b


def line_alpha(X, Y, width, soft, op, brush=lambda t, bv: np.ones_like(t), raw=False):
    """Closed form of a horizontal stroke from x = -0.5 to 0.5 at y = 0 (uniform
    parameter t = x + 0.5), for pixels with |x| < 0.5; 0 where the fragment is
    discarded (alpha < 0.001)."""
    t = X + 0.5
    p = 1.0 - 4.0 * (t - 0.5) ** 2
    lw, ls = width * p, soft * p
    d = np.abs(Y)
    a_geom = 1.0 - smoothstep(lw - ls, lw + ls, d)
    bv = np.clip(Y / (lw + ls) * 0.5 + 0.5, 0.0, 1.0)
    a = np.clip(a_geom * brush(t, bv) * op * p, 0.0, 0.999)
    return a if raw else np.where(a < 0.001, 0.0, a)


def interior(X, a_raw, margin=0.04):
    # inside the stroke's extent, away from the 0.001 discard threshold
    return (np.abs(X) < 0.5 - margin) & (np.abs(a_raw - 0.001) > 1e-4)


W = H = 64


def test_straight_stroke_closed_form(oracle):
    c = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.1, 0.05, (0.2, 0.4, 0.6), 0.8)
    img = oracle.bz_render_f32(c[None], eye_vp(), W, H, brush_const(), quantize=False)
    X, Y = px_centres(W, H)
    a = line_alpha(X, Y, 0.1, 0.05, 0.8)
    m = interior(X, line_alpha(X, Y, 0.1, 0.05, 0.8, raw=True))
    exp = np.stack([0.2 * a, 0.4 * a, 0.6 * a, a], -1)
    assert np.abs(img[m] - exp[m]).max() < 2e-5
    assert (a[m] > 0.3).sum() > 50  # the stroke is really drawn
    # nothing beyond width + softness of the stroke
    far = (np.abs(Y) > 0.16) | (np.abs(X) > 0.66)
    assert np.all(img[far] == 0.0)


def test_draw_order_index_zero_on_top(oracle):
    c0 = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.1, 0.05, (1.0, 0.0, 0.0), 0.7)
    c1 = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.1, 0.05, (0.0, 0.0, 1.0), 0.7)
    img = oracle.bz_render_f32(np.stack([c0, c1]), eye_vp(), W, H, brush_const(), quantize=False)
    X, Y = px_centres(W, H)
    a = line_alpha(X, Y, 0.1, 0.05, 0.7)
    m = interior(X, a) & (a > 0.01)
    # curve 1 is drawn first, curve 0 over it
    assert np.abs(img[..., 0][m] - a[m]).max() < 2e-5
    assert np.abs(img[..., 2][m] - (a * (1.0 - a))[m]).max() < 2e-5
    assert np.abs(img[..., 3][m] - (a + a * (1.0 - a))[m]).max() < 2e-5


def test_brush_v_follows_signed_cross_distance(oracle):
    # rows of increasing value: the brush is linear in v inside [0.5/bh, 1 - 0.5/bh]
    bh, bw = 64, 16
    b = np.zeros((bh, bw, 4), np.uint8)
    b[..., 0] = (4 * np.arange(bh))[:, None]
    c = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.2, 0.05, (1.0, 1.0, 1.0), 1.0)
    img = oracle.bz_render_f32(c[None], eye_vp(), W, H, b, quantize=False)
    X, Y = px_centres(W, H)

    def ramp(t, bv):
        return np.clip(4.0 * (bv * bh - 0.5), 0.0, 4.0 * (bh - 1)) / 255.0

    a = line_alpha(X, Y, 0.2, 0.05, 1.0, ramp)
    m = interior(X, line_alpha(X, Y, 0.2, 0.05, 1.0, ramp, raw=True))
    assert np.abs(img[..., 3][m] - a[m]).max() < 2e-5
    # above the stroke (y > 0, left of its +x direction) the brush is brighter
    up, down = (Y > 0.03) & (np.abs(X) < 0.3), (Y < -0.03) & (np.abs(X) < 0.3)
    assert img[..., 3][up].mean() > 2.0 * img[..., 3][down].mean()


def test_brush_u_repeats(oracle):
    bw, bh = 4, 4
    b = np.zeros((bh, bw, 4), np.uint8)
    b[:, 0, 0] = 255  # only column 0 is set
    c = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.3, 0.01, (1.0, 1.0, 1.0), 1.0)
    img = oracle.bz_render_f32(c[None], eye_vp(), W, H, b, quantize=False)
    X, Y = px_centres(W, H)

    def col0(t, bv):
        x = 4.0 * t - 0.5
        return np.where(x < 0.0, x + 1.0, np.where(x < 1.0, 1.0 - x, np.where(x < 3.0, 0.0, x - 3.0)))

    a = line_alpha(X, Y, 0.3, 0.01, 1.0, col0)
    ar = line_alpha(X, Y, 0.3, 0.01, 1.0, col0, raw=True)
    m = (np.abs(X) < 0.49) & (np.abs(Y) < 0.05) & (np.abs(ar - 0.001) > 1e-4)
    assert np.abs(img[..., 3][m] - a[m]).max() < 2e-5
    # near t = 1 the filter wraps to column 0 (clamp-to-edge would give 0 there)
    end = m & (X > 0.4)
    assert img[..., 3][end].max() > 0.05


def test_near_cull_and_load(oracle):
    vp = persp_w_is_z()
    base = np.random.default_rng(0).random((H, W, 4)).astype(np.float32)
    behind = curve((-0.5, 0.0, 1.0), (0.0, 0.3, 1.0), (0.2, -0.3, -1.0), (0.5, 0.0, 1.0), 0.1, 0.05,
                   (1.0, 1.0, 1.0), 1.0)
    out = oracle.bz_render_f32(behind[None], vp, W, H, brush_const(), rgba=base, clear=False, quantize=False)
    assert np.array_equal(out, base)
    cleared = oracle.bz_render_f32(behind[None], vp, W, H, brush_const(), rgba=base, clear=True, quantize=False)
    assert np.all(cleared == 0.0)
    # in front of the camera the same curve draws, over the loaded image
    front = behind.copy()
    front[10] = 1.0
    out = oracle.bz_render_f32(front[None], vp, W, H, brush_const(), rgba=base, clear=False, quantize=False)
    assert (out != base).any(axis=-1).sum() > 50


def test_quad_clamp_and_quantize(oracle):
    # aspect 8: screen x spans [-8, 8]; a stroke across it draws only where the
    # quad, clamped to [-4, 4], covers the pixel centre
    Wd, Hd = 512, 64
    c = straight((-0.9, 0.0, 0.5), (0.9, 0.0, 0.5), 0.3, 0.05, (0.3, 0.6, 0.9), 1.0)
    img = oracle.bz_render_f32(c[None], eye_vp(), Wd, Hd, brush_const(), quantize=True)
    X, Y = px_centres(Wd, Hd)
    on = img[..., 3] > 0.0
    assert on.any()
    assert np.all(np.abs(X[on]) <= 4.0)
    assert on[(np.abs(X) < 3.9) & (np.abs(Y) < 0.05)].all()
    # every channel on the 8-bit grid
    k = img * 255.0
    assert np.abs(k - np.rint(k)).max() < 1e-4


def test_quantized_blend_rounds_each_step(oracle):
    c0 = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.1, 0.05, (1.0, 0.0, 0.0), 0.7)
    c1 = straight((-0.5, 0.0, 0.5), (0.5, 0.0, 0.5), 0.1, 0.05, (0.0, 0.0, 1.0), 0.7)
    img = oracle.bz_render_f32(np.stack([c0, c1]), eye_vp(), W, H, brush_const(), quantize=True)
    X, Y = px_centres(W, H)
    a = line_alpha(X, Y, 0.1, 0.05, 0.7)
    q = lambda x: np.rint(np.clip(x, 0.0, 1.0) * 255.0) / 255.0  # noqa: E731
    exp_b = q(q(a) * (1.0 - a))  # blue: written by curve 1, then scaled by curve 0
    exp_a = q(a + q(a) * (1.0 - a))
    m = interior(X, a) & (a > 0.01)
    # away from rounding ties
    m &= np.abs((q(a) * (1.0 - a)) * 255.0 % 1.0 - 0.5) > 1e-3
    m &= np.abs((a + q(a) * (1.0 - a)) * 255.0 % 1.0 - 0.5) > 1e-3
    assert np.array_equal(img[..., 2][m], exp_b[m].astype(np.float32))
    assert np.array_equal(img[..., 3][m], exp_a[m].astype(np.float32))


def test_sampler_repeat_and_clamp(oracle):
    b = np.zeros((2, 4, 4), np.uint8)
    b[0, :, 0] = [0, 40, 80, 120]
    b[1, :, 0] = [200, 200, 200, 200]
    v = lambda r, c: b[r, c, 0] / 255.0  # noqa: E731
    # texel centres are exact
    assert oracle.bz_brush_sample(b, 1.5 / 4, 0.25) == pytest.approx(v(0, 1), abs=1e-7)
    # half way between column 3 and column 0 (repeat), row 0
    assert oracle.bz_brush_sample(b, 0.0, 0.25) == pytest.approx(0.5 * (v(0, 3) + v(0, 0)), abs=1e-7)
    assert oracle.bz_brush_sample(b, 1.0, 0.25) == pytest.approx(0.5 * (v(0, 3) + v(0, 0)), abs=1e-7)
    # v below the first texel centre clamps to row 0, above the last to row 1
    assert oracle.bz_brush_sample(b, 1.5 / 4, 0.0) == pytest.approx(v(0, 1), abs=1e-7)
    assert oracle.bz_brush_sample(b, 1.5 / 4, 1.0) == pytest.approx(v(1, 1), abs=1e-7)
    # bilinear in between
    exp = 0.5 * (0.25 * v(0, 1) + 0.75 * v(0, 2)) + 0.5 * (0.25 * v(1, 1) + 0.75 * v(1, 2))
    assert oracle.bz_brush_sample(b, (1.75 + 0.5) / 4, 0.5) == pytest.approx(exp, abs=1e-6)
