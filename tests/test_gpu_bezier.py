"""Bezier stroke optimizer on the GPU (GpuBezierOptimizerManager over nbz_*)
against the CPU oracle: tile lists, images, fixed-point gradients, loss
accumulators, pixel loss map, Adam step and densification, bit for bit."""
import math

import numpy as np
import pytest

from nprush_amd import GpuBezierOptimizerManager, camera, scenes

pytestmark = pytest.mark.gpu


def _view(W, H, eye=(0.4, 0.3, 3.0)):
    return camera.view_proj(eye, aspect=W / H)


def _textures(W, H, seed):
    rng = np.random.default_rng(seed)
    tgt = np.zeros((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3))
    tgt[..., 3] = 1.0
    td = np.where(rng.random((H, W)) < 0.7, rng.uniform(0.5, 0.98, (H, W)), 1.0).astype(np.float32)
    bg = np.zeros((H, W, 4), np.float32)
    bg[..., :3] = rng.random((H, W, 3))
    bgd = rng.uniform(0.9, 1.0, (H, W)).astype(np.float32)
    nrm = np.zeros((H, W, 4), np.float32)
    nrm[..., :3] = rng.random((H, W, 3))
    return tgt, td, bg, bgd, nrm


def _curves(n, seed, width=0.03):
    c = scenes.make_bezier_curves(n, seed, radius_median=width, walk_std=0.15)
    return c


def _manager(curves, W, H, tex, mode, max_width=0.0, bg_penalty=0.0, eye=(0.4, 0.3, 3.0), no_kill=False):
    vp, vpi = _view(W, H, eye)
    m = GpuBezierOptimizerManager(numBeziers=curves.shape[0])
    m.write_beziers(curves)
    m.writeVPMatrix(vp)
    m.writeVPInvMatrix(vpi)
    m.writeMode(mode)
    m.writeMaxWidth(max_width)
    m.writeBgPenalty(bg_penalty)
    m.writeNoKill(no_kill)
    m.setBackwardTarget(*tex, W, H)
    return m, vp, vpi


def _check_tiles(m, oracle, curves, uni, W, H):
    lists, cnt, full = oracle.bz_bin(curves, uni, W, H)
    ranges, ids = m.read_tiles()
    assert ranges.shape[0] == len(cnt)
    for t in range(len(cnt)):
        s, e = int(ranges[t, 0]), int(ranges[t, 1])
        k = min(e - s, 1024)
        assert k == cnt[t], t
        assert np.array_equal(ids[s:s + k], lists[t, :k]), t
    return cnt


@pytest.mark.parametrize("W,H,n,mode,max_width,bg_pen", [
    (96, 64, 400, 1.0, 0.0, 0.0),     # colour mode, base layer
    (96, 64, 400, 0.0, 0.0, 0.0),     # edge mode
    (83, 45, 300, 1.0, 0.05, 0.3),    # fine layer: regularisers, background penalty, ragged size
    (16, 16, 1, 1.0, 0.0, 0.0),
])
def test_backward_matches_oracle(oracle, W, H, n, mode, max_width, bg_pen):
    curves = _curves(n, 11 + n)
    curves[::17, 19] = 0.003  # dead curves are skipped
    tex = _textures(W, H, W + H)
    m, vp, vpi = _manager(curves, W, H, tex, mode, max_width, bg_pen)
    uni = oracle.bz_uniforms(vp, vpi, mode=mode, max_width=max_width, bg_penalty=bg_pen)
    m.backward()
    cnt = _check_tiles(m, oracle, curves, uni, W, H)
    exp = oracle.bz_backward_f32(curves, uni, *tex, W, H)
    C, D, T = m.read_image()
    assert np.array_equal(C, exp["C"]) and np.array_equal(D, exp["D"]) and np.array_equal(T, exp["T"])
    assert np.array_equal(m.read_grads(), exp["grads"])
    ga, la = m.read_adc()
    assert np.array_equal(la, exp["loss_acc"]) and (ga == 0).all()
    assert np.array_equal(m.read_pixel_loss(), exp["pixel_loss"])
    assert np.allclose(m.read_loss(), exp["loss"], rtol=1e-9, atol=1e-9)
    st = m.read_stats()
    tx = (W + 15) // 16
    pix = [min(16, W - 16 * (t % tx)) * min(16, H - 16 * (t // tx)) for t in range(len(cnt))]
    assert st["pairs"] == int(np.dot(cnt.astype(np.int64), pix))
    if n > 1:
        assert np.abs(exp["grads"]).max() > 0 and st["reverse_pairs"] > 0
    # a second backward adds into the same accumulators
    m.backward()
    exp2 = oracle.bz_backward_f32(curves, uni, *tex, W, H, grads=exp["grads"], loss_acc=exp["loss_acc"],
                                  pixel_loss=exp["pixel_loss"])
    assert np.array_equal(m.read_grads(), exp2["grads"])
    assert np.array_equal(m.read_adc()[1], exp2["loss_acc"])
    assert np.array_equal(m.read_pixel_loss(), exp2["pixel_loss"])
    # image only: same image, no gradients added
    m.forward()
    assert np.array_equal(m.read_image()[0], exp["C"])
    assert np.array_equal(m.read_grads(), exp2["grads"])
    m.close()


def test_tile_cap_and_dead_curves(oracle):
    """1100 overlapping strokes in one tile keep the first 1024 ascending ids;
    every curve dead renders the background."""
    W, H = 32, 32
    base = _curves(1, 5, width=0.05)[0]
    for k, p in enumerate(([-0.3, 0.0, 0.0], [-0.1, 0.05, 0.0], [0.1, 0.05, 0.0], [0.3, 0.1, 0.0])):
        base[4 * k:4 * k + 3] = p
    curves = np.repeat(base[None], 1100, 0)
    rng = np.random.default_rng(2)
    curves[:, 16:19] = rng.random((1100, 3))
    curves[:, 19] = rng.uniform(0.1, 0.3, 1100)
    tex = _textures(W, H, 9)
    m, vp, vpi = _manager(curves, W, H, tex, 1.0)
    uni = oracle.bz_uniforms(vp, vpi, mode=1.0)
    m.backward()
    cnt = _check_tiles(m, oracle, curves, uni, W, H)
    assert cnt.max() == 1024 and m.read_stats()["tiles_over_cap"] >= 1
    exp = oracle.bz_backward_f32(curves, uni, *tex, W, H)
    assert np.array_equal(m.read_image()[0], exp["C"])
    assert np.array_equal(m.read_grads(), exp["grads"])
    m.close()
    dead = curves[:50].copy()
    dead[:, 19] = 0.0
    m, vp, vpi = _manager(dead, W, H, tex, 1.0)
    m.backward()
    C, _, T = m.read_image()
    assert (T == 1.0).all() and np.array_equal(C, tex[2][..., :3])
    assert (m.read_grads() == 0).all()
    m.close()


@pytest.mark.parametrize("no_kill", [False, True])
def test_step_matches_oracle(oracle, no_kill):
    W, H, n = 64, 48, 500
    curves = _curves(n, 3)
    curves[::23, 19] = 0.002
    curves[5, 0:3] = [9.0, 9.0, 0.0]  # far off-screen
    curves[6, 0:3] = [9.2, 9.1, 0.1]
    curves[6, 4:7] = [9.3, 9.0, 0.1]
    tex = _textures(W, H, 4)
    m, vp, vpi = _manager(curves, W, H, tex, 1.0, no_kill=no_kill)
    m.writeKillThresholds(0.2, 0.0)
    uni = oracle.bz_uniforms(vp, vpi, mode=1.0, prune_alpha=0.2)
    rng = np.random.default_rng(8)
    g = rng.integers(-300000, 300000, (n, 18)).astype(np.int32)
    m0 = rng.normal(0, 1e-3, (n, 18)).astype(np.float32)
    v0 = rng.uniform(0, 1e-5, (n, 18)).astype(np.float32)
    ga0 = rng.uniform(0, 0.01, n).astype(np.float32)
    m.write_grads(g)
    m.write_adam(m0, v0, 6.0)
    m.write_adc(ga0, np.zeros(n, np.int64))
    m.step()
    eb, eg, em, ev, ega = oracle.bz_step_f32(curves, g, m0, v0, ga0, uni, 7.0, W * H, no_kill=1.0 if no_kill else 0.0)
    assert np.array_equal(m.read_beziers(), eb)
    assert (m.read_grads() == 0).all() and (eg == 0).all()
    mm, vv, t = m.read_adam()
    assert t == 7.0 and np.array_equal(mm, em) and np.array_equal(vv, ev)
    assert np.array_equal(m.read_adc()[0], ega)
    m.close()


@pytest.mark.parametrize("n,dead_frac,seed", [(300, 0.3, 1), (2000, 0.05, 2), (700, 0.9, 3)])
def test_adc_matches_oracle(oracle, n, dead_frac, seed):
    W, H = 80, 60
    rng = np.random.default_rng(seed)
    curves = _curves(n, 40 + seed)
    curves[rng.random(n) < dead_frac, 19] = 0.0
    off = rng.random(n) < 0.05                      # some entirely off-screen
    for k in range(4):
        curves[off, 4 * k:4 * k + 3] += 20.0
    long_ = rng.random(n) < 0.5
    curves[long_, 12:15] = curves[long_, 0:3] + rng.normal(0, 0.4, (long_.sum(), 3))  # long: split
    tex = _textures(W, H, 7)
    m, vp, vpi = _manager(curves, W, H, tex, 1.0)
    uni = oracle.bz_uniforms(vp, vpi, mode=1.0)
    m0 = rng.normal(0, 1, (n, 18)).astype(np.float32)
    v0 = rng.uniform(0, 1, (n, 18)).astype(np.float32)
    ga = np.where(rng.random(n) < 0.3, rng.uniform(0.011, 0.05, n), rng.uniform(0, 0.009, n)).astype(np.float32)
    la = np.where(rng.random(n) < 0.2, rng.integers(2 ** 32 // 10, 2 ** 33, n), 0).astype(np.int64)
    pl = np.where(rng.random(W * H) < 0.1, rng.integers(1, 5000, W * H), 0).astype(np.int32)
    m.write_adam(m0, v0, 50.0)
    m.write_adc(ga, la)
    m.write_pixel_loss(pl)
    m.adc()
    eb, em, ev, ega, ela, epl, est = oracle.bz_adc_f32(curves, m0, v0, ga, la, pl, W, H, uni, adam_t=50.0)
    assert np.array_equal(m.read_beziers(), eb)
    mm, vv, _ = m.read_adam()
    assert np.array_equal(mm, em) and np.array_equal(vv, ev)
    g2, l2 = m.read_adc()
    assert (g2 == 0).all() and (l2 == 0).all() and (m.read_pixel_loss() == 0).all()
    st = m.read_stats()
    got = [st[k] for k in ("adc_dead", "adc_killed", "adc_split1", "adc_spawned", "adc_split3", "adc_dead_left")]
    assert got == list(est)
    assert est[1] > 0 and est[2] > 0
    if dead_frac > 0.5:
        assert est[3] > 0
    m.close()


def test_training_loop_matches_oracle(oracle):
    """dispatch() = backward + step (+ densification every adc_period) for 6
    steps over two alternating views, bit for bit."""
    W, H, n = 64, 64, 300
    curves = _curves(n, 21)
    curves[::5, 19] = 0.0
    tex = _textures(W, H, 12)
    m, vp, vpi = _manager(curves, W, H, tex, 1.0)
    vp2, vpi2 = _view(W, H, eye=(-0.5, 0.2, 2.8))
    m.setAdcPeriod(3)
    views = [(vp, vpi), (vp2, vpi2)]
    bz = curves.copy()
    g = np.zeros((n, 18), np.int32)
    mm = np.zeros((n, 18), np.float32)
    vv = np.zeros((n, 18), np.float32)
    ga = np.zeros(n, np.float32)
    la = np.zeros(n, np.int64)
    pl = np.zeros(W * H, np.int32)
    for k in range(1, 7):
        a, ai = views[k % 2]
        m.writeVPMatrix(a)
        m.writeVPInvMatrix(ai)
        m.dispatch()
        uni = oracle.bz_uniforms(a, ai, mode=1.0)
        out = oracle.bz_backward_f32(bz, uni, *tex, W, H, grads=g, loss_acc=la, pixel_loss=pl)
        g, la, pl = out["grads"], out["loss_acc"], out["pixel_loss"]
        bz, g, mm, vv, ga = oracle.bz_step_f32(bz, g, mm, vv, ga, uni, float(k), W * H)
        if k % 3 == 0:
            bz, mm, vv, ga, la, pl, _ = oracle.bz_adc_f32(bz, mm, vv, ga, la, pl, W, H, uni, adam_t=float(k))
        assert np.array_equal(m.read_beziers(), bz), k
    m2, v2, t = m.read_adam()
    assert t == 6.0 and np.array_equal(m2, mm) and np.array_equal(v2, vv)
    g2, l2 = m.read_adc()
    assert np.array_equal(g2, ga) and np.array_equal(l2, la)
    assert np.array_equal(m.read_pixel_loss(), pl)
    m.close()
