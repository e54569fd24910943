"""Pins of the CPU oracle (no GPU).

The reference ships no tests or fixtures, so the oracle is pinned by what the
mathematics fixes: closed forms, invariants, FP64 finite differences and
hand-built cases whose outcome is worked out in the test.
"""
import math

import numpy as np
import pytest

from nprush_amd import camera, scenes

ULP = np.float32(2.0 ** -23)


def _ulps(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    scale = np.maximum(np.abs(want), np.finfo(np.float32).tiny)
    return np.abs(got - want) / (scale * float(ULP))


# ---------------------------------------------------------------- transcendentals
def test_exp2_log2_accuracy(oracle):
    rng = np.random.default_rng(0)
    x = rng.uniform(-120, 120, 4000).astype(np.float32)
    assert _ulps(oracle.math_f32(0, x), np.exp2(x.astype(np.float64))).max() < 3.0
    y = np.exp(rng.uniform(-60, 60, 4000)).astype(np.float32)
    ref = np.log2(y.astype(np.float64))
    got = oracle.math_f32(1, y)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    assert err.max() < 4 * float(ULP)


def test_pow_exp_log_sin_accuracy(oracle):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-4, 6, 3000)).astype(np.float32)
    y = rng.uniform(-3, 8, 3000).astype(np.float32)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    bound = 4.0 + 1.5 * np.abs(y * np.log2(x.astype(np.float64)))
    assert (_ulps(oracle.math_f32(2, x, y), ref) < bound).all()
    e = rng.uniform(-15, 15, 3000).astype(np.float32)
    assert _ulps(oracle.math_f32(3, e), np.exp(e.astype(np.float64))).max() < 12
    s = rng.uniform(-1e6, 1e6, 3000).astype(np.float32)
    sref = np.sin(s.astype(np.float64))
    assert np.abs(oracle.math_f32(5, s) - sref).max() < 1e-7
    cref = np.cos(s.astype(np.float64))
    assert np.abs(oracle.math_f32(7, s) - cref).max() < 1e-7
    q = np.float32(np.pi / 2) * np.arange(-8, 9, dtype=np.float32)
    assert np.abs(oracle.math_f32(7, q) - np.cos(q.astype(np.float64))).max() < 1e-7


def test_math_special_values(oracle):
    assert oracle.math_f32(0, [0.0])[0] == 1.0
    assert oracle.math_f32(0, [3.0])[0] == 8.0
    assert oracle.math_f32(0, [200.0])[0] == np.inf
    assert oracle.math_f32(0, [-200.0])[0] == 0.0
    assert oracle.math_f32(1, [1.0])[0] == 0.0 and oracle.math_f32(1, [8.0])[0] == 3.0
    assert oracle.math_f32(1, [0.0])[0] == -np.inf and np.isnan(oracle.math_f32(1, [-1.0])[0])
    assert oracle.math_f32(2, [2.0], [10.0])[0] == 1024.0


def test_f2i_saturates_like_wgsl(oracle):
    lib = oracle.load()
    assert lib.ora_f2i(3e9) == 2**31 - 1 and lib.ora_f2i(-3e9) == -(2**31)
    assert lib.ora_f2i(float("nan")) == 0
    assert lib.ora_f2i(-1.7) == -1 and lib.ora_f2i(1.9999) == 1


# ---------------------------------------------------------------- camera
def test_perspective_closed_form():
    p = camera.perspective(math.pi / 3, 1.5, 0.01, 100.0)
    f = 1.0 / math.tan(math.pi / 6)
    assert p[0] == np.float32(f / 1.5) and p[5] == np.float32(f) and p[11] == -1.0 and p[15] == 0.0
    assert p[10] == np.float32(100.0 / (0.01 - 100.0)) and p[14] == np.float32(100.0 * 0.01 / (0.01 - 100.0))


def test_view_proj_inverse_and_centre():
    vp, vi = scenes.config0_camera()
    M = vp.reshape(4, 4).T.astype(np.float64)
    Mi = vi.reshape(4, 4).T.astype(np.float64)
    assert np.abs(M @ Mi - np.eye(4)).max() < 1e-4  # float32 entries up to ~300
    clip = M @ np.array([0, 0, 0, 1.0])
    assert abs(clip[0]) < 1e-7 and abs(clip[1]) < 1e-7 and abs(clip[3] - 3.0) < 1e-6
    # a point above the origin appears above the centre (y up)
    assert (M @ np.array([0, 0.1, 0, 1.0]))[1] > 0


# ---------------------------------------------------------------- tiny scenes
def _splat(x=0.0, y=0.0, z=0.0, s=0.05, rgb=(1.0, 0.0, 0.0), op=0.8, q=(1.0, 0.0, 0.0, 0.0), a=2.0, b=0.5):
    out = np.zeros(16, np.float32)
    out[0:3] = (x, y, z)
    out[3] = out[12] = out[15] = s
    out[4:7] = rgb
    out[7] = op
    out[8:12] = q
    out[13] = a
    out[14] = b
    return out


def test_single_splat_closed_form(oracle):
    """One isotropic splat on the optical axis: alpha = op*exp(-b r^2) with
    r^2 = |d|^2 / (sigma^2 + filter^2), sigma = s*f/w (f = 1/tan(fov/2))."""
    W = H = 64
    s = _splat(s=0.08, rgb=(0.9, 0.2, 0.4), op=0.7)[None]
    vp, vi = scenes.config0_camera()
    tgt = np.zeros((H, W, 4), np.float32)
    dep = np.ones((H, W), np.float32)
    _, c, d, _, _ = oracle.backward_f32(s, vp, vi, tgt, dep, W, H)
    f = 1.0 / math.tan(math.pi / 6)
    sig2 = (0.08 * f / 3.0) ** 2 + (0.3 * 2.0 / H) ** 2
    yy, xx = np.mgrid[0:H, 0:W]
    px = ((xx + 0.5) / W) * 2 - 1
    py = -(((yy + 0.5) / H) * 2 - 1)
    r = np.sqrt(np.maximum((px ** 2 + py ** 2) / sig2, 1e-4))
    a = np.clip(0.7 * np.exp(-0.5 * r ** 2), 0, 0.999)
    a = np.where(-0.5 * r ** 2 > -15, a, 0.0)
    want = a[..., None] * np.array([0.9, 0.2, 0.4]) + (1 - a)[..., None] * 0.05
    assert np.abs(c - want).max() < 2e-5
    enc = 1 - 0.1 / 3.0
    assert np.abs(d - (a * enc + (1 - a))).max() < 2e-5


def test_index_order_is_depth_order(oracle):
    """Index 0 is composited in front regardless of depth (BW:199-210)."""
    W = H = 16
    front_far = _splat(z=-0.5, s=0.2, rgb=(1, 0, 0), op=0.9)
    back_near = _splat(z=0.5, s=0.2, rgb=(0, 0, 1), op=0.9)
    vp, vi = scenes.config0_camera()
    tgt = np.zeros((H, W, 4), np.float32)
    dep = np.ones((H, W), np.float32)
    _, c, _, _, _ = oracle.backward_f32(np.stack([front_far, back_near]), vp, vi, tgt, dep, W, H)
    f = 1.0 / math.tan(math.pi / 6)
    d2 = 2 * (0.5 / 16 * 2) ** 2  # pixel (8, 8) centre is (1/16, -1/16) from the axis

    def alpha(w):
        sig2 = (0.2 * f / w) ** 2 + (0.3 * 2.0 / H) ** 2
        return 0.9 * math.exp(-0.5 * d2 / sig2)

    a1, a2 = alpha(3.5), alpha(2.5)
    cx = c[8, 8]
    bg = (1 - a1) * (1 - a2) * 0.05
    assert abs(cx[0] - (a1 + bg)) < 1e-5 and abs(cx[2] - ((1 - a1) * a2 + bg)) < 1e-5


def test_transmittance_partition(oracle):
    """With every colour 1: C = sum T_i a_i + 0.05 T_n = 1 - 0.95 T_n, and the
    depth image with every splat at the same depth: D = enc + (1-enc) T_n."""
    W = H = 32
    rng = np.random.default_rng(3)
    s = np.stack([_splat(x=rng.uniform(-.3, .3), y=rng.uniform(-.3, .3), z=0.0, s=rng.uniform(.02, .1),
                         rgb=(1, 1, 1), op=rng.uniform(.1, .95)) for _ in range(40)])
    vp, vi = scenes.config0_camera()
    tgt = np.zeros((H, W, 4), np.float32)
    dep = np.ones((H, W), np.float32)
    lists, counts, _ = oracle.bin_tiles(s, vp, W, H)
    tiles = np.arange(len(counts), dtype=np.int32)
    _, c, d, _ = oracle.backward_f64(s.astype(np.float64), vp.astype(np.float64), vi.astype(np.float64), tgt, dep,
                                     W, H, tiles, lists, counts, want_grads=False)
    Tn = (1.0 - c[..., 0]) / 0.95
    enc = 1 - 0.1 / 3.0
    assert np.abs(d - (enc + (1 - enc) * Tn)).max() < 1e-9
    assert (Tn >= -1e-12).all() and (Tn <= 1 + 1e-12).all()


def test_project_axis_is_derivative_of_projection():
    """Restated project_axis (BW:52-58) equals the FP64 directional derivative of
    the projected centre along the axis."""
    vp, _ = camera.view_proj((0.7, 0.4, 2.6), aspect=1.5)
    M = vp.reshape(4, 4).T.astype(np.float64)
    aspect = 1.5

    def proj(p):
        c = M @ np.append(p, 1.0)
        return np.array([c[0] / c[3] * aspect, c[1] / c[3]])

    p = np.array([0.2, -0.1, 0.3])
    ax = np.array([0.3, 0.8, -0.52])
    c = M @ np.append(p, 1.0)
    ac = M @ np.append(ax, 0.0)
    w = c[3]
    analytic = np.array([(ac[0] * w - c[0] * ac[3]) / (w * w) * aspect, (ac[1] * w - c[1] * ac[3]) / (w * w)])
    h = 1e-6
    fd = (proj(p + h * ax) - proj(p - h * ax)) / (2 * h)
    assert np.abs(analytic - fd).max() < 1e-7


# ---------------------------------------------------------------- binning
def test_binning_properties(oracle):
    W, H = 96, 80
    rng = np.random.default_rng(4)
    s = scenes.make_splats(3000, 4)
    s[:, 0:3] *= 0.7
    s[rng.random(3000) < 0.1, 7] = 0.004  # culled by opacity
    s[rng.random(3000) < 0.05, 2] = 10.0  # behind the camera (w < 0)
    vp, _ = camera.view_proj((0.1, 0.2, 2.8), aspect=W / H)
    ids, cnt, full = oracle.bin_tiles(s, vp, W, H, cap=64)
    M = vp.reshape(4, 4).T.astype(np.float64)
    aspect = W / H
    ntx = (W + 15) // 16
    # independent FP64 bounding boxes (square, R*sqrt(max(cov00,cov11)))
    clip = (M @ np.c_[s[:, :3], np.ones(len(s))].T).T
    w = clip[:, 3]
    scr = np.c_[clip[:, 0] / w * aspect, clip[:, 1] / w]
    q = s[:, 8:12].astype(np.float64)
    qw, qx, qy, qz = q.T
    Rm = np.stack([
        np.stack([1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy + qw * qz), 2 * (qx * qz - qw * qy)], 1),
        np.stack([2 * (qx * qy - qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz + qw * qx)], 1),
        np.stack([2 * (qx * qz + qw * qy), 2 * (qy * qz - qw * qx), 1 - 2 * (qx * qx + qy * qy)], 1)], 1)
    sc = np.maximum(s[:, [3, 12, 15]].astype(np.float64), 1e-4)
    cov = np.zeros((len(s), 2))
    for k in range(3):
        ac = (M[:, :3] @ Rm[:, k, :].T).T
        ax = np.c_[(ac[:, 0] * w - clip[:, 0] * ac[:, 3]) / w ** 2 * aspect, (ac[:, 1] * w - clip[:, 1] * ac[:, 3]) / w ** 2]
        cov += (ax * sc[:, k:k + 1]) ** 2
    cov += (0.3 * 2 / H) ** 2
    R = (15.0 / np.maximum(s[:, 14], 1e-4)) ** (1.0 / s[:, 13])
    maxr = R * np.sqrt(np.maximum(cov.max(1), 1e-9))
    alive = (s[:, 7] >= 0.005) & (w >= 0)
    for t in range(len(cnt)):
        lst = ids[t, : cnt[t]]
        assert (np.diff(lst.astype(np.int64)) > 0).all()
        assert cnt[t] == min(full[t], 64)
        assert alive[lst].all()
        tx, ty = t % ntx, t // ntx
        x0 = ((16 * tx + 0.5) / W * 2 - 1) * aspect
        x1 = ((min(16 * tx + 16, W) + 0.5) / W * 2 - 1) * aspect
        y1 = -((16 * ty + 0.5) / H * 2 - 1)
        y0 = -((min(16 * ty + 16, H) + 0.5) / H * 2 - 1)
        gap = np.maximum.reduce([scr[:, 0] - maxr - x1, x0 - (scr[:, 0] + maxr), scr[:, 1] - maxr - y1,
                                 y0 - (scr[:, 1] + maxr)])
        inside = alive & (gap < -1e-5)
        outside = ~alive | (gap > 1e-5)
        want = np.nonzero(inside)[0]
        if full[t] <= 64:
            assert set(want).issubset(set(lst.tolist()))
        assert not outside[lst].any()


# ---------------------------------------------------------------- gradients vs FD
def _fd_scene():
    rng = np.random.default_rng(8)
    n = 12
    s = np.stack([_splat(x=rng.uniform(-.25, .25), y=rng.uniform(-.25, .25), z=rng.uniform(-.3, .3),
                         s=rng.uniform(.04, .09), rgb=rng.uniform(.1, .9, 3), op=rng.uniform(.3, .8),
                         q=tuple(np.r_[1.0, rng.normal(0, .3, 3)] / np.linalg.norm(np.r_[1.0, rng.normal(0, .3, 3)])),
                         a=rng.uniform(1.6, 2.6), b=rng.uniform(.4, .7)) for _ in range(n)])
    s[:, 12] *= 1.4
    s[:, 15] *= 0.7
    return s.astype(np.float64)


@pytest.mark.parametrize("fg", [False, True])
def test_backward_matches_finite_differences(oracle, fg):
    """FP64 analytic gradients (no regulariser, no 3D pull, no alpha<0.001 skip) against central
    differences (h = 1e-4) of the loss the reference seeds:
    sum |C - tgt|^2 + fg (D - tgt_depth)^2.  Position gradients omit the
    conic's dependence on position (BW:327-348), so they are compared with a
    looser bound and the deviation is reported."""
    W = H = 32
    s = _fd_scene()
    vp, vi = scenes.config0_camera()
    rng = np.random.default_rng(9)
    tgt = np.zeros((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3))
    dep = np.ones((H, W), np.float32)
    if fg:
        dep[8:24, 8:24] = 0.965
    lists, counts, _ = oracle.bin_tiles(s.astype(np.float32), vp, W, H)
    tiles = np.arange(len(counts), dtype=np.int32)
    args = (vp.astype(np.float64), vi.astype(np.float64), tgt, dep, W, H, tiles, lists, counts)
    g, _, _, _ = oracle.backward_f64(s, *args, flags=oracle.NO_REG | oracle.NO_PULL | oracle.NO_SKIP)
    g_skip, _, _, _ = oracle.backward_f64(s, *args, flags=oracle.NO_REG | oracle.NO_PULL)
    h = 1e-4
    worst = {}
    for k in range(16):
        for i in range(len(s)):
            sp, sm = s.copy(), s.copy()
            sp[i, k] += h
            sm[i, k] -= h
            lp = oracle.backward_f64(sp, *args, want_grads=False)[3][0]
            lm = oracle.backward_f64(sm, *args, want_grads=False)[3][0]
            fd = (lp - lm) / (2 * h)
            err = abs(fd - g[i, k]) / max(1e-3, abs(fd))
            worst[k] = max(worst.get(k, 0.0), err)
    for k in [3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]:
        assert worst[k] < 2e-3, f"param {k}: relative FD error {worst[k]:.3g}"
    # the reference's alpha < 0.001 skip (BW:263) drops the faint tail; reported
    rel_skip = np.abs(g_skip - g).max(0) / np.maximum(np.abs(g).max(0), 1e-12)
    print("relative gradient change from the alpha<0.001 skip per param:", np.round(rel_skip, 4))
    assert rel_skip[4:8].max() < 0.01
    # position: screen-space part only (the conic term is missing by design)
    print("position FD deviation (reported):", {k: round(worst[k], 4) for k in (0, 1, 2)})
    assert worst[0] < 0.5 and worst[1] < 0.5


def test_fixed_point_grads_track_fp64(oracle):
    """The i32 path equals the FP64 gradients up to truncation: per pair at most
    one unit of 1e-4 (1e-5 for colour/opacity) is lost."""
    W = H = 32
    s = _fd_scene().astype(np.float32)
    vp, vi = scenes.config0_camera()
    rng = np.random.default_rng(10)
    tgt = np.zeros((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3))
    dep = np.ones((H, W), np.float32)
    gi, _, _, _, _ = oracle.backward_f32(s, vp, vi, tgt, dep, W, H)
    lists, counts, _ = oracle.bin_tiles(s, vp, W, H)
    tiles = np.arange(len(counts), dtype=np.int32)
    g64, _, _, _ = oracle.backward_f64(s.astype(np.float64), vp.astype(np.float64), vi.astype(np.float64), tgt, dep,
                                       W, H, tiles, lists, counts)
    fp = np.array([1e4] * 4 + [1e5] * 4 + [1e4] * 8)
    pairs = float(W * H * len(s))
    err = np.abs(gi / fp - g64) * fp
    assert (err <= pairs + 1).all()


# ---------------------------------------------------------------- Adam
def test_adam_first_step_closed_form(oracle):
    """Step 1 with gradient g: m_hat = g, v_hat = g^2, update = clamp(lr g/(|g|+eps), +-mu)."""
    n = 4
    s = np.tile(_splat(s=0.5, rgb=(0.5, 0.5, 0.5), op=0.5, a=3.0, b=1.0), (n, 1))
    raw = np.zeros((n, 16), np.int32)
    raw[0, :] = 16384 * 10          # grad = 10/fp per param
    raw[1, :] = -16384 * 3
    raw[2, 4] = 1                   # tiny colour gradient
    z = np.zeros((n, 16), np.float32)
    out, g_out, m, v, acc, t = oracle.step_f32(s, raw, z, z, np.zeros(n, np.float32), 0.0)
    lr = np.array([0.0005] * 3 + [0.01] + [0.02] * 3 + [0.01] + [0.005] * 4 + [0.01] * 4)
    mu = np.array([0.005] * 4 + [0.001] * 3 + [0.0005] + [0.005] * 5 + [0.05, 0.05, 0.005])
    fp = np.array([1e4] * 4 + [1e5] * 4 + [1e4] * 8)
    for row, sign in [(0, 1.0), (1, -1.0)]:
        g = sign * np.abs(raw[row] / fp / 16384.0)
        upd = np.clip(lr * g / (np.abs(g) + 1e-8), -mu, mu)
        want = s[row].astype(np.float64) - upd
        want[8:12] /= np.linalg.norm(want[8:12])
        assert np.abs(out[row] - want).max() < 1e-6
    assert not g_out.any() and t == 1.0
    assert abs(acc[0] - math.sqrt(3) * 10 / 1e4) < 1e-8


def test_adam_kill_and_bounds(oracle):
    s = np.stack([_splat(op=0.049), _splat(s=0.004), _splat(op=0.5, rgb=(2.0, -1.0, 0.5))])
    z = np.zeros((3, 16), np.float32)
    out, *_ = oracle.step_f32(s, np.zeros((3, 16), np.int32), z, z, np.zeros(3, np.float32), 5.0)
    assert out[0, 7] == 0.0          # opacity < 0.05 killed
    assert out[1, 7] == 0.0          # volume 6.4e-8 < 1e-7 killed
    assert out[2, 4] == 1.0 and out[2, 5] == 0.05   # colour clamped to [0.05, 1]


# ---------------------------------------------------------------- ADC
def test_adc_hand_built(oracle):
    """Serial scan worked by hand (splat_adc.wgsl:27-87), threshold 25*5e-5;
    dead slots 1, 4, 5, 6:
    0 alive hot, large -> split into slot 1 (first dead);
    2 alive cold;
    3 alive hot, small -> clone into slot 4;
    4 is reached alive (child of 3) with a stale hot accum -> clones into 5;
    5 is reached alive with accum 0 -> nothing;
    6 dead when reached -> skipped (its stale accum is not reset);
    7 alive hot, large -> split into slot 6 (the cursor is at 6): that child is
      never visited, so slot 6 keeps its stale accum."""
    hot, cold = 25 * 5e-5 * 4, 25 * 5e-5 * 0.5
    s = np.stack([_splat(s=0.05) if i in (0, 2) else _splat(s=0.003) for i in range(8)])
    for i in (1, 4, 5, 6):
        s[i, 7] = 0.0
    s[7] = _splat(s=0.05, op=0.6)
    acc = np.array([hot, 0, cold, hot, hot, 0, hot, hot], np.float32)
    m = np.ones((8, 16), np.float32)
    v = np.ones((8, 16), np.float32)
    s2, m2, v2, a2, st = oracle.adc_f32(s, m, v, acc, 100.0)
    assert list(st) == [4, 4, 2, 2]   # qualifiers 0,3,4,7; children 1,4,5,6; splits 0,7; clones 3,4
    sx = np.float32(0.05) * np.float32(0.625)
    assert s2[0, 3] == sx and s2[1, 3] == sx and s2[6, 3] == sx and s2[7, 3] == sx
    assert np.isclose(s2[0, 0], -sx * 0.05) and np.isclose(s2[1, 0], sx * 0.05)
    assert s2[6, 7] == np.float32(0.6) and s2[1, 7] == s[0, 7]
    assert s2[4, 7] == s[3, 7] and 0 < abs(s2[4, 0] - s[3, 0]) <= 0.001 and s2[4, 3] == s[3, 3]
    assert s2[5, 7] == s[3, 7] and abs(s2[5, 0] - s[3, 0]) <= 0.002
    for i in (0, 1, 3, 4, 5, 6, 7):
        assert not m2[i].any() and not v2[i].any()
    assert m2[2].all() and v2[2].all()
    assert list(a2) == [0, 0, 0, 0, 0, 0, np.float32(hot), 0]


def test_adc_child_before_cursor_not_visited(oracle):
    """Parent 3 takes dead slot 1 (< 3): the child is never visited, so its
    stale accum is neither tested nor reset."""
    hot = 25 * 5e-5 * 4
    s = np.stack([_splat(s=0.003) for _ in range(4)])
    s[1, 7] = 0.0
    acc = np.array([0, hot, 0, hot], np.float32)
    z = np.zeros((4, 16), np.float32)
    s2, _, _, a2, st = oracle.adc_f32(s, z, z, acc, 7.0)
    assert list(st) == [1, 1, 0, 1]
    assert a2[1] == np.float32(hot) and a2[3] == 0.0
    assert s2[1, 7] == s[3, 7]


# ---------------------------------------------------------------- render
def test_render_single_splat_closed_form(oracle):
    W = H = 32
    s = _splat(s=0.1, rgb=(0.2, 0.8, 0.4), op=0.6)[None]
    vp, _ = scenes.config0_camera()
    c, d = oracle.render_f32(s, vp, W, H, clear=True, quantize=False)
    # pixel (16,16): d = (1/32, -1/32); alpha = op exp(-0.5 |d|^2 / sigma^2)
    f = 1.0 / math.tan(math.pi / 6)
    sig2 = (0.1 * f / 3.0) ** 2 + (0.3 * 2.0 / H) ** 2
    a = 0.6 * math.exp(-0.5 * 2 * (1 / 32) ** 2 / sig2)
    assert abs(c[16, 16, 3] - 1.0) < 1e-6  # alpha channel: a + 1*(1-a)
    assert abs(c[16, 16, 1] - (0.8 * a + 0.05 * (1 - a))) < 1e-5
    assert abs(d[16, 16, 0] - ((1 - 0.1 / 3.0) * a + (1 - a))) < 1e-5
    # corners untouched
    assert np.allclose(c[0, 0], [0.05, 0.05, 0.05, 1.0]) and np.allclose(d[0, 0], 1.0)
    cq, _ = oracle.render_f32(s, vp, W, H, clear=True, quantize=True)
    assert np.allclose(cq * 255, np.rint(cq * 255), atol=1e-4)


# ---------------------------------------------------------------- cut radius
def test_cut_radius_gives_exact_zero_alpha(oracle):
    """The GPU skips pow/exp for pairs with r^2 > rcut2 = max((1.001 R)^2, 1e-4),
    R = pow(15/max(b,1e-4), 1/a), when a in [0.1, 64] and b >= 1e-4.  Check with
    the deterministic FP32 functions that every such pair has
    power = -b exp2(a log2(r2)/2) <= -15, i.e. alpha is exactly 0 in the
    reference's formula (BW:236-237)."""
    lib = oracle.load()
    rng = np.random.default_rng(12)
    f32 = np.float32
    for _ in range(20000):
        a = f32(np.exp(rng.uniform(np.log(0.1), np.log(64.0))))
        b = f32(np.exp(rng.uniform(np.log(1e-4), np.log(1e4))))
        R = f32(lib.ora_powf(float(f32(15.0) / max(b, f32(1e-4))), float(f32(1.0) / a)))
        if not (0 < R <= 1e15):
            continue
        rc = f32(R * f32(1.001))
        rcut2 = max(f32(rc * rc), f32(1e-4))
        for r2 in (np.nextafter(rcut2, f32(np.inf)), f32(rcut2 * f32(1.0 + rng.uniform(0, 0.01)))):
            lr = f32(f32(0.5) * f32(lib.ora_log2f(float(max(r2, f32(1e-4))))))
            power = f32(-b * f32(lib.ora_exp2f(float(f32(a * lr)))))
            assert not (power > f32(-15.0)), (a, b, r2, power)
