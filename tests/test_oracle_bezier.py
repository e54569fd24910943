"""Pins of the CPU Bezier oracle (bezier_backward.wgsl, bezier_step.wgsl,
bezier_adc.wgsl restated in oracle/oracle_bezier.c) by closed forms,
finite differences and invariants that do not restate its formulas."""
import numpy as np
import pytest

LUMA = np.array([0.2126, 0.7152, 0.0722], np.float32).astype(np.float64)


def eye_vp():
    return np.eye(4, dtype=np.float32).reshape(16)


def persp_w_is_z():
    """clip = (x, y, z, z): clip.w equals the control point's z."""
    m = np.zeros((4, 4), np.float32)  # m[col][row], flattened column-major
    m[0, 0] = 1.0
    m[1, 1] = 1.0
    m[2, 2] = 1.0
    m[2, 3] = 1.0
    return m.reshape(16)


def curve(p0, p1, p2, p3, width, soft, rgb, a):
    b = np.zeros(20, np.float32)
    for k, p in enumerate((p0, p1, p2, p3)):
        b[4 * k:4 * k + 3] = p
    b[3], b[7] = width, soft
    b[16:19] = rgb
    b[19] = a
    return b


def straight(A, B, width, soft, rgb, a):
    A, B = np.asarray(A, np.float64), np.asarray(B, np.float64)
    pts = [A + (B - A) * (k / 3.0) for k in range(4)]
    return curve(*pts, width, soft, rgb, a)


def textures(W, H, tgt_rgb=(0.2, 0.3, 0.4), tdepth=1.0, bg=(0.0, 0.0, 0.0), bgd=1.0):
    t = np.zeros((H, W, 4), np.float32)
    t[..., :3] = tgt_rgb
    t[..., 3] = 1.0
    d = np.full((H, W), tdepth, np.float32)
    b = np.zeros((H, W, 4), np.float32)
    b[..., :3] = bg
    bd = np.full((H, W), bgd, np.float32)
    nrm = np.zeros((H, W, 4), np.float32)
    return t, d, b, bd, nrm


def screen(px, py, W, H):
    aspect = W / H
    return (((px + 0.5) / W) * 2 - 1) * aspect, -(((py + 0.5) / H) * 2 - 1)


def smoothstep(e0, e1, x):
    if e1 == e0:  # zero-width band at the curve ends (pressure 0): a step
        return 1.0 if x > e0 else 0.0
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3 - 2 * t)


def stroke_alpha(A2, B2, width, soft, op, p):
    """Alpha of a straight stroke between screen points A2, B2 at screen point p."""
    ab = B2 - A2
    t = np.clip(np.dot(p - A2, ab) / np.dot(ab, ab), 0.0, 1.0)
    d = np.linalg.norm(p - (A2 + t * ab))
    pr = 1 - 4 * (t - 0.5) ** 2
    a = (1 - smoothstep((width - soft) * pr, (width + soft) * pr, d)) * op * pr
    return min(max(a, 0.0), 0.999), t


def surrogate_loss(out, tgt, tdepth, W, H, mode):
    """The loss whose gradient the reverse pass seeds (bezier_backward.wgsl:271-314):
    contrast-weighted |e|^2 + 2 (luma . e)^2, plus (D - t)^2 / 2 on foreground pixels
    in colour mode, plus 2 (coverage - edge)^2 in edge mode."""
    C = out["C"].astype(np.float64)
    e = C - tgt[..., :3].astype(np.float64)
    luma = (tgt[..., :3].astype(np.float64) * LUMA).sum(-1)
    pad = np.pad(luma, 1, mode="edge")
    gx = pad[1:-1, 2:] - pad[1:-1, :-2]
    gy = pad[2:, 1:-1] - pad[:-2, 1:-1]
    cw = 1 + np.clip(np.sqrt(gx * gx + gy * gy) * 8, 0, 3)
    L = (cw * ((e * e).sum(-1) + 2 * ((e * LUMA).sum(-1)) ** 2)).sum()
    if mode > 0.5:
        fg = tdepth < 0.99
        L += (0.5 * (out["D"].astype(np.float64) - tdepth) ** 2)[fg].sum()
    else:
        cov = 1 - out["T"].astype(np.float64)
        L += (2 * (cov - tgt[..., 0]) ** 2).sum()
    return L


# ------------------------------------------------------------------ forward
def test_straight_stroke_image_closed_form(oracle):
    W, H = 48, 32
    A, B = np.array([-0.6, -0.3, 0.5]), np.array([0.7, 0.4, 0.5])
    b = straight(A, B, 0.06, 0.03, (0.9, 0.2, 0.1), 0.8)
    uni = oracle.bz_uniforms(eye_vp(), eye_vp(), mode=1.0)
    tgt, td, bg, bgd, nrm = textures(W, H, bg=(0.1, 0.5, 0.3), bgd=0.97)
    out = oracle.bz_backward_f32(b[None], uni, tgt, td, bg, bgd, nrm, W, H)
    aspect = W / H
    A2, B2 = np.array([A[0] * aspect, A[1]]), np.array([B[0] * aspect, B[1]])
    dval = 1 - 0.1 / 1.0
    n_cov = 0
    for py in range(H):
        for px in range(W):
            a, _ = stroke_alpha(A2, B2, 0.06, 0.03, 0.8, np.array(screen(px, py, W, H)))
            n_cov += a > 0.01
            want = a * np.array([0.9, 0.2, 0.1]) + (1 - a) * np.array([0.1, 0.5, 0.3])
            assert np.abs(out["C"][py, px] - want).max() < 2e-5, (px, py)
            assert abs(out["T"][py, px] - (1 - a)) < 2e-5
            assert abs(out["D"][py, px] - (a * dval + (1 - a) * 0.97)) < 2e-5
    assert n_cov > 20


def test_composite_order_is_ascending_id(oracle):
    W, H = 16, 16
    b0 = straight([-0.8, 0.0, 0.5], [0.8, 0.0, 0.5], 0.2, 0.05, (1.0, 0.0, 0.0), 0.7)
    b1 = straight([0.0, -0.8, 0.5], [0.0, 0.8, 0.5], 0.2, 0.05, (0.0, 0.0, 1.0), 0.6)
    uni = oracle.bz_uniforms(eye_vp(), eye_vp(), mode=0.0)
    tgt, td, bg, bgd, nrm = textures(W, H)
    out = oracle.bz_backward_f32(np.stack([b0, b1]), uni, tgt, td, bg, bgd, nrm, W, H)
    for py in range(H):
        for px in range(W):
            p = np.array(screen(px, py, W, H))
            a0, _ = stroke_alpha(np.array([-0.8, 0.0]), np.array([0.8, 0.0]), 0.2, 0.05, 0.7, p)
            a1, _ = stroke_alpha(np.array([0.0, -0.8]), np.array([0.0, 0.8]), 0.2, 0.05, 0.6, p)
            want = a0 * np.array([1, 0, 0]) + (1 - a0) * a1 * np.array([0, 0, 1])
            assert np.abs(out["C"][py, px] - want).max() < 2e-5


def test_binning_is_conservative_and_capped(oracle):
    """Lists hold the first 1024 ascending ids of the curves whose padded hull box
    meets the tile; the image from the binned lists equals the image from lists of
    every live curve, so no curve outside its box contributes."""
    rng = np.random.default_rng(3)
    W, H, n = 40, 24, 60
    bz = np.zeros((n, 20), np.float32)
    for i in range(n):
        c = rng.uniform(-1.2, 1.2, 3)
        pts = [c + rng.normal(0, 0.15, 3) for _ in range(4)]
        for p in pts:
            p[2] = 0.5
        bz[i] = curve(*pts, rng.uniform(0.005, 0.05), rng.uniform(0.002, 0.02), rng.uniform(0, 1, 3),
                      rng.uniform(0.2, 0.9))
    bz[5, 19] = 0.004  # below the kill threshold: never listed
    uni = oracle.bz_uniforms(eye_vp(), eye_vp(), mode=1.0)
    tgt, td, bg, bgd, nrm = textures(W, H, bg=(0.3, 0.3, 0.3))
    lists, cnt, full = oracle.bz_bin(bz, uni, W, H)
    assert (cnt == full).all() and cnt.max() < 1024
    for k in range(len(cnt)):
        ids = lists[k, :cnt[k]]
        assert (np.diff(ids.astype(np.int64)) > 0).all()
        assert 5 not in ids
    binned = oracle.bz_backward_f32(bz, uni, tgt, td, bg, bgd, nrm, W, H, lists=lists, counts=cnt)
    live = np.array([i for i in range(n) if i != 5], np.uint32)
    alll = np.zeros_like(lists)
    alll[:, :live.size] = live
    every = oracle.bz_backward_f32(bz, uni, tgt, td, bg, bgd, nrm, W, H, lists=alll,
                                   counts=np.full(len(cnt), live.size, np.int32))
    assert np.array_equal(binned["C"], every["C"])
    assert np.array_equal(binned["T"], every["T"])
    assert np.array_equal(binned["grads"], every["grads"])
    # cap: 1100 copies of one stroke in a single tile keep ids 0..1023
    many = np.repeat(straight([-0.5, 0, 0.5], [0.5, 0, 0.5], 0.05, 0.01, (1, 1, 1), 0.5)[None], 1100, 0)
    lists, cnt, full = oracle.bz_bin(many, oracle.bz_uniforms(eye_vp(), eye_vp()), 16, 16)
    assert cnt[0] == 1024 and full[0] == 1100 and np.array_equal(lists[0], np.arange(1024))
    # behind the camera (clip.w <= 0 for any control point): not listed
    bad = straight([-0.3, 0, 0.5], [0.3, 0, 0.5], 0.05, 0.01, (1, 1, 1), 0.5)
    bad[10] = -1.0
    lists, cnt, _ = oracle.bz_bin(bad[None], oracle.bz_uniforms(persp_w_is_z(), eye_vp()), 16, 16)
    assert cnt[0] == 0


# ------------------------------------------------------------ gradients (FD)
def _roof(z=0.5, width=0.05, soft=0.02, rgb=(0.8, 0.3, 0.2), a=0.7):
    """Symmetric arch below the screen origin, apex (t = 1/2) at y = -0.0575 z: the
    projections of (0, 0) onto both segments next to the apex fall beyond it, so the
    nearest polyline point is that vertex and stays at t = 1/2 under small changes."""
    return curve([-0.5, -0.08, z], [-0.2, -0.05, z], [0.2, -0.05, z], [0.5, -0.08, z], width, soft, rgb, a)


def _fd(oracle, b, uni, tex, W, H, mode, idx, h):
    tgt, td = tex[0], tex[1]
    bp, bm = b.copy(), b.copy()
    bp[idx] += h
    bm[idx] -= h
    Lp = surrogate_loss(oracle.bz_backward_f32(bp[None], uni, *tex, W, H), tgt, td, W, H, mode)
    Lm = surrogate_loss(oracle.bz_backward_f32(bm[None], uni, *tex, W, H), tgt, td, W, H, mode)
    return (Lp - Lm) / (2 * h)


PARAM_SLOT = [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 16, 17, 18, 19, 3, 7]
PARAM_FP = [1e4] * 12 + [1e5] * 4 + [1e4] * 2


@pytest.mark.parametrize("mode", [0.0, 1.0])
def test_vertex_gradients_match_finite_differences(oracle, mode):
    """1x1 view, edge and colour modes (colour mode on a foreground pixel with
    equal control-point depths): positions, colour and opacity gradients are the
    derivative of the seeded loss; width and softness gradients are that
    derivative divided by the soft-band width 2 * softness * pressure."""
    W = H = 1
    b = _roof(z=1.0)
    vp = persp_w_is_z() if mode > 0.5 else eye_vp()
    uni = oracle.bz_uniforms(vp, eye_vp(), mode=mode)
    tex = textures(W, H, tgt_rgb=(0.1, 0.6, 0.9), tdepth=0.5, bg=(0.2, 0.2, 0.7), bgd=0.95)
    if mode < 0.5:
        tex[0][..., 0] = 0.9  # edge target
        b = _roof(z=0.5)
    out = oracle.bz_backward_f32(b[None], uni, *tex, W, H)
    g = out["grads"][0].astype(np.float64) / np.array(PARAM_FP)
    assert 0.0 < 1 - out["T"][0, 0] < 0.999
    denom = 2 * 0.02 * 1.0
    for p in range(18):
        h = 1e-4 if p < 12 else 1e-4
        fd = _fd(oracle, b, uni, tex, W, H, mode, PARAM_SLOT[p], h)
        want = fd / denom if p >= 16 else fd
        assert abs(g[p] - want) < 3e-3 * max(1.0, abs(want)) + 2.5 / PARAM_FP[p], (p, g[p], want)


def test_direction_regulariser_gradient(oracle):
    """Fine layer (max_width > 0): the extra position gradient is the derivative of
    1.5 (tangent . flow)^2 of the pixel's nearest polyline segment, and softness
    gains 5 * 2 * softness per active pair."""
    W, H = 5, 1
    aspect = W / H
    A, B = np.array([-0.15 / aspect, -0.04, 0.5]), np.array([0.15 / aspect, 0.05, 0.5])
    # curve through the centre pixel (screen origin) at t = 0.3: segment k = 5
    t0 = 0.3
    D = B - A
    A = A - (A + t0 * D) * np.array([1, 1, 0])
    B = A + D
    b = straight(A, B, 0.05, 0.02, (0.5, 0.5, 0.5), 0.6)
    tex = textures(W, H)
    tex[0][0, 0, :3] = 0.0
    tex[0][0, 4, :3] = 0.8  # horizontal ramp: flow along +x at the centre pixel
    base = oracle.bz_uniforms(eye_vp(), eye_vp(), mode=0.0)
    fine = oracle.bz_uniforms(eye_vp(), eye_vp(), mode=0.0, max_width=0.1)
    g0 = oracle.bz_backward_f32(b[None], base, *tex, W, H)["grads"][0].astype(np.int64)
    g1 = oracle.bz_backward_f32(b[None], fine, *tex, W, H)["grads"][0].astype(np.int64)
    assert abs((g1[17] - g0[17]) - 10 * 0.02 * 1e4) <= 1
    flow = np.array([1.0, 0.0])

    def ldir(bb):
        P = [np.array([bb[4 * c] * aspect, bb[4 * c + 1]], np.float64) for c in range(4)]

        def at(t):
            o = 1 - t
            return o ** 3 * P[0] + 3 * o * o * t * P[1] + 3 * o * t * t * P[2] + t ** 3 * P[3]
        seg = at(5 / 16) - at(4 / 16)
        tau = seg / np.linalg.norm(seg)
        return 1.5 * np.dot(tau, flow) ** 2

    for p in range(12):
        slot = PARAM_SLOT[p]
        bp, bm = b.astype(np.float64), b.astype(np.float64)
        bp = bp.copy()
        bm = bm.copy()
        bp[slot] += 1e-6
        bm[slot] -= 1e-6
        fd = (ldir(bp) - ldir(bm)) / 2e-6
        got = (g1[p] - g0[p]) / 1e4
        assert abs(got - fd) < 2e-3 * max(1.0, abs(fd)) + 3e-4, (p, got, fd)


def test_loss_accumulators_single_pair(oracle):
    W = H = 1
    b = _roof()
    uni = oracle.bz_uniforms(eye_vp(), eye_vp(), mode=0.0)
    tex = textures(W, H, tgt_rgb=(0.1, 0.6, 0.9))
    out = oracle.bz_backward_f32(b[None], uni, *tex, W, H)
    a = 1 - out["T"][0, 0]
    dcol = out["grads"][0, 12:15].astype(np.float64) / 1e5          # dC * a
    la = out["loss_acc"][0] * 2.0 ** -32                           # |dC|^2 * a
    assert abs(la - (dcol ** 2).sum() / a) < 1e-3 * la + 1e-6
    e = out["C"][0, 0].astype(np.float64) - tex[0][0, 0, :3]
    want = int((e ** 2).sum() * out["T"][0, 0] * 1e4)
    assert abs(int(out["pixel_loss"][0]) - want) <= 1


# --------------------------------------------------------------------- step
def test_step_first_update_closed_form(oracle):
    rng = np.random.default_rng(7)
    n = 6
    bz = np.stack([straight(rng.uniform(-0.5, 0.5, 3), rng.uniform(-0.5, 0.5, 3), 0.02, 0.01,
                            rng.uniform(0.2, 0.8, 3), 0.5) for _ in range(n)])
    bz[3, 19] = 0.002  # dead: gradients drained, nothing updated
    g = rng.integers(-50000, 50000, (n, 18)).astype(np.int32)
    g[1, :] = 0
    m = np.zeros((n, 18), np.float32)
    v = np.zeros((n, 18), np.float32)
    ga = np.zeros(n, np.float32)
    uni = oracle.bz_uniforms(eye_vp(), eye_vp())
    W, H = 64, 32
    nb, ng, nm, nv, nga = oracle.bz_step_f32(bz, g, m, v, ga, uni, 1.0, W * H, no_kill=1.0)
    assert (ng == 0).all()
    lr = np.array([0.005] * 12 + [0.01] * 3 + [0.005, 0.002, 0.002])
    mu = np.array([0.005] * 12 + [0.01] * 3 + [0.005, 0.003, 0.003])
    fp = np.array(PARAM_FP)
    for i in range(n):
        if i == 3:
            assert np.array_equal(nb[i], bz[i]) and (nm[i] == 0).all()
            continue
        grad = g[i] / fp / (W * H)
        upd = np.clip(lr * grad / (np.abs(grad) + 1e-8), -mu, mu)  # m_hat = g, v_hat = g^2 at t = 1
        want = bz[i].astype(np.float64).copy()
        for p in range(18):
            s = PARAM_SLOT[p]
            lo = -1e9 if p < 12 else (0.0 if p < 16 else 0.001)
            hi = 1e9 if p < 12 else (1.0 if p < 15 else (0.99 if p == 15 else (0.1 if p == 16 else 0.03)))
            want[s] = np.clip(want[s] - upd[p], lo, hi)
        if want[19] < 0.001 or want[3] <= 0.001:
            want[19] = 0.0
        assert np.abs(nb[i] - want).max() < 2e-6, i
        assert abs(nga[i] - np.sqrt((grad[:12] ** 2).sum())) < 1e-6 * max(1.0, nga[i])
    assert np.array_equal(nb[1], bz[1])


def test_step_prune_and_offscreen(oracle):
    uni = oracle.bz_uniforms(eye_vp(), eye_vp(), prune_alpha=0.3)
    b_ok = straight([-0.2, 0, 0.5], [0.2, 0, 0.5], 0.02, 0.01, (0.5, 0.5, 0.5), 0.5)
    b_faint = b_ok.copy()
    b_faint[19] = 0.2            # below the layer's prune threshold 0.3
    b_thin = b_ok.copy()
    b_thin[3] = 0.001            # width at the floor: pruned
    b_off = straight([3.0, 0, 0.5], [3.5, 0.1, 0.5], 0.02, 0.01, (0.5, 0.5, 0.5), 0.5)
    bz = np.stack([b_ok, b_faint, b_thin, b_off])
    z = np.zeros((4, 18), np.int32)
    out = oracle.bz_step_f32(bz, z, np.zeros((4, 18), np.float32), np.zeros((4, 18), np.float32),
                             np.zeros(4, np.float32), uni, 1.0, 100.0, no_kill=0.0)[0]
    assert out[0, 19] == 0.5 and out[1, 19] == 0.0 and out[2, 19] == 0.0 and out[3, 19] == 0.0
    kept = oracle.bz_step_f32(bz, z, np.zeros((4, 18), np.float32), np.zeros((4, 18), np.float32),
                              np.zeros(4, np.float32), uni, 1.0, 100.0, no_kill=1.0)[0]
    assert kept[3, 19] == 0.5


# ---------------------------------------------------------------------- ADC
def _bez(P, t):
    o = 1 - t
    return o ** 3 * P[0] + 3 * o * o * t * P[1] + 3 * o * t * t * P[2] + t ** 3 * P[3]


def _pts(b):
    return [b[4 * k:4 * k + 3].astype(np.float64) for k in range(4)]


def test_adc_kill_split_clone_spawn(oracle):
    W, H = 32, 16
    uni = oracle.bz_uniforms(eye_vp(), eye_vp())
    long_ = straight([-0.6, -0.2, 0.5], [0.6, 0.3, 0.4], 0.02, 0.01, (0.1, 0.2, 0.3), 0.6)   # split
    short = straight([0.1, 0.1, 0.5], [0.15, 0.12, 0.5], 0.02, 0.01, (0.4, 0.5, 0.6), 0.6)   # clone
    stuck = straight([0.0, 0.0, 0.5], [0.1, 0.0, 0.5], 0.02, 0.01, (0.4, 0.5, 0.6), 0.6)     # killed
    off = straight([2.0, 2.0, 0.5], [2.2, 2.1, 0.5], 0.02, 0.01, (0.4, 0.5, 0.6), 0.6)       # off-screen
    quiet = straight([-0.3, 0.3, 0.5], [-0.2, 0.3, 0.5], 0.02, 0.01, (0.4, 0.5, 0.6), 0.6)   # untouched
    dead = straight([0, 0, 0.5], [0.1, 0, 0.5], 0.02, 0.01, (0, 0, 0), 0.0)
    bz = np.stack([long_, dead, short, stuck, dead, off, quiet, dead, dead])
    n = bz.shape[0]
    m = np.ones((n, 18), np.float32)
    v = np.ones((n, 18), np.float32)
    ga = np.zeros(n, np.float32)
    ga[0] = ga[2] = 1.0                    # grad_norm 0.02 > 2e-4
    la = np.zeros(n, np.int64)
    la[3] = int(1.0 * 2 ** 32)             # loss_norm 0.02 > 1e-3 with grad_norm 0: killed
    pl = np.zeros(W * H, np.int32)
    pl[5 * W + 20] = 777                   # the only loss peak (on a scanned pixel: stride 1)
    nb, nm, nv, nga, nla, npl, st = oracle.bz_adc_f32(bz, m, v, ga, la, pl, W, H, uni, adam_t=50.0)
    # dead list = [1, 4, 7, 8] (initial) + [3, 5] (killed), popped from the end:
    # pass 1.5 fills 5 and 3, the four spawns fill 8, 7, 4, 1, pass 3 finds none
    assert list(st) == [4, 2, 2, 4, 0, 0]
    # pass 1.5: curve 0 splits into slot 5, curve 2 clones into slot 3
    P = _pts(long_)
    for s in np.linspace(0, 1, 9):
        assert np.abs(_bez(_pts(nb[0]), s) - _bez(P, s / 2)).max() < 2e-6
        assert np.abs(_bez(_pts(nb[5]), s) - _bez(P, 0.5 + s / 2)).max() < 2e-6
    assert nb[0, 3] == long_[3] and nb[5, 3] == long_[3] and nb[5, 7] == long_[7]
    j = nb[3, :3] - short[:3]
    assert 0 < np.abs(j).max() <= 0.0005
    for k in range(4):
        assert np.allclose(nb[3, 4 * k:4 * k + 3] - short[4 * k:4 * k + 3], j, atol=1e-7)
    assert np.array_equal(nb[2], short)
    # pass 2: slot 8 spawned at the loss peak (claimed), 7, 4, 1 at fallback pixels
    for slot in (8, 7, 4, 1):
        c = (nb[slot, 0:3].astype(np.float64) + nb[slot, 12:15]) / 2
        px = ((c[0] + 1) / 2) * W - 0.5  # identity VP: ndc = world xy
        py = ((1 - c[1]) / 2) * H - 0.5
        assert abs(px - round(px)) < 1e-4 and abs(py - round(py)) < 1e-4
        assert 0 <= round(px) < W and 0 <= round(py) < H
        if slot == 8:
            assert (round(px), round(py)) == (20, 5)
        tangent = (nb[slot, 12:15].astype(np.float64) - nb[slot, 0:3]) / 2
        assert abs(np.linalg.norm(tangent) - 0.025) < 1e-6 and tangent[1] == 0.0
        assert np.allclose(nb[slot, 4:7], c - 0.33 * tangent, atol=1e-6)
        assert np.allclose(nb[slot, 8:11], c + 0.33 * tangent, atol=1e-6)
        assert nb[slot, 3] == np.float32(0.015) and nb[slot, 7] == np.float32(0.005)
        assert (nb[slot, 16:20] == 0.5).all()
    for s in (0, 1, 2, 3, 4, 5, 7, 8):
        assert (nm[s] == 0).all() and (nv[s] == 0).all()
    assert (nm[6] == 1).all() and np.array_equal(nb[6], quiet)
    assert (nga == 0).all() and (nla == 0).all() and (npl == 0).all()


def test_adc_second_split_pass_after_512_spawns(oracle):
    """With more than 512 dead slots left after pass 1.5, pass 3 splits / clones
    the high-gradient curves again (the first half of an earlier split included)."""
    W, H = 64, 64
    uni = oracle.bz_uniforms(eye_vp(), eye_vp())
    long_ = straight([-0.6, -0.2, 0.5], [0.6, 0.3, 0.4], 0.02, 0.01, (0.1, 0.2, 0.3), 0.6)
    short = straight([0.1, 0.1, 0.5], [0.15, 0.12, 0.5], 0.02, 0.01, (0.4, 0.5, 0.6), 0.6)
    n = 530
    bz = np.repeat(short[None], n, 0)
    bz[:, 19] = 0.0
    bz[0], bz[1] = long_, short
    ga = np.zeros(n, np.float32)
    ga[:2] = 1.0
    z = np.zeros((n, 18), np.float32)
    pl = np.arange(W * H, dtype=np.int32) % 97
    nb, nm, nv, nga, nla, npl, st = oracle.bz_adc_f32(bz, z, z, ga, np.zeros(n, np.int64), pl, W, H, uni,
                                                      adam_t=7.0)
    assert list(st) == [528, 0, 2, 512, 2, 12]
    # pass 1.5 put the second half of curve 0 in slot 529; pass 3 split the first half into slot 15
    P = _pts(long_)
    for s in np.linspace(0, 1, 5):
        assert np.abs(_bez(_pts(nb[529]), s) - _bez(P, 0.5 + s / 2)).max() < 2e-6
        assert np.abs(_bez(_pts(nb[0]), s) - _bez(P, s / 4)).max() < 2e-6
        assert np.abs(_bez(_pts(nb[15]), s) - _bez(P, 0.25 + s / 4)).max() < 2e-6
    assert (nb[2:14, 19] == 0).all() and (nb[14:, 19] > 0).all()


def test_adc_no_kill_and_fallback(oracle):
    W, H = 20, 10
    uni = oracle.bz_uniforms(eye_vp(), eye_vp())
    stuck = straight([0.0, 0.0, 0.5], [0.1, 0.0, 0.5], 0.02, 0.01, (0.4, 0.5, 0.6), 0.6)
    dead = stuck.copy()
    dead[19] = 0.0
    bz = np.stack([stuck, dead, dead])
    la = np.array([2 ** 33, 0, 0], np.int64)
    z = np.zeros((3, 18), np.float32)
    nb, *_, st = oracle.bz_adc_f32(bz, z, z, np.zeros(3, np.float32), la, np.zeros(W * H, np.int32), W, H, uni,
                                   adam_t=100.0, no_kill=1.0)
    assert nb[0, 19] == np.float32(0.6) and list(st) == [2, 0, 0, 2, 0, 0]
    for s in (1, 2):
        c = (nb[s, 0:3].astype(np.float64) + nb[s, 12:15]) / 2
        px = ((c[0] + 1) / 2) * W - 0.5
        py = ((1 - c[1]) / 2) * H - 0.5
        assert 0 <= px <= W - 1 and 0 <= py <= H - 1
        assert abs(px - round(px)) < 1e-4 and abs(py - round(py)) < 1e-4
