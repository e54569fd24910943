"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact for tile lists, fixed-point gradients, images, Adam and ADC (the
kernels and the oracle evaluate the same deterministic FP32 expressions).
"""
import math

import numpy as np
import pytest

from nprush_amd import GpuSplatForwardPipelineManager, GpuSplatOptimizerManager, camera, debug_math, scenes

pytestmark = pytest.mark.gpu


def _target(H, W, seed=1, fg=False):
    rng = np.random.default_rng(seed)
    tgt = np.zeros((H, W, 4), np.float32)
    tgt[..., :3] = scenes.quantize8(rng.random((H, W, 3)))
    tgt[..., 3] = 1.0
    dep = np.ones((H, W), np.float32)
    if fg:
        yy, xx = np.mgrid[0:H, 0:W]
        mask = (xx - W / 2) ** 2 + (yy - H / 2) ** 2 < (min(W, H) / 3) ** 2
        dep[mask] = scenes.quantize8(0.97 - 0.001 * rng.random(mask.sum()))
    return tgt, dep


def _run_backward(splats, vp, vi, tgt, dep, W, H, cap=1024):
    mgr = GpuSplatOptimizerManager(numSplats=splats.shape[0], splats=splats, max_tile_splats=cap)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, dep, W, H)
    mgr.backward()
    return mgr


@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 5, 7])
def test_math_bit_identity(oracle, op):
    rng = np.random.default_rng(op)
    if op in (0, 3):
        x = rng.uniform(-130, 130, 20000).astype(np.float32)
    elif op in (1, 4):
        x = np.exp(rng.uniform(-80, 80, 20000)).astype(np.float32)
    elif op == 2:
        x = np.exp(rng.uniform(-6, 8, 20000)).astype(np.float32)
    else:
        x = rng.uniform(-5e7, 5e7, 20000).astype(np.float32)
    y = rng.uniform(-4, 10, 20000).astype(np.float32) if op == 2 else None
    special = np.array([0.0, -0.0, 1.0, 0.5, 2.0, np.inf, -np.inf, np.nan, 1e-40, 126.0, -126.0, 128.0],
                       np.float32)
    x = np.concatenate([x, special])
    if y is not None:
        y = np.concatenate([y, np.full(special.size, 1.5, np.float32)])
    g = debug_math(op, x, y)
    c = oracle.math_f32(op, x, y)
    assert np.array_equal(g.view(np.uint32), c.view(np.uint32)) or np.array_equal(g, c, equal_nan=True)


def test_float_to_int_instruction_matches_wgsl_conversion(oracle):
    """The gradient path converts with gfx950's v_cvt_i32_f32; it must equal the
    reference's saturating i32(f32) (NaN -> 0) on every class of input."""
    lib = oracle.load()
    rng = np.random.default_rng(6)
    x = np.concatenate([
        np.array([0.0, -0.0, 0.5, -0.5, 1.5, -1.5, 2147483520.0, -2147483648.0, 2147483648.0, -2147483904.0,
                  3e9, -3e9, np.inf, -np.inf, np.nan, 1e-40, -1e-40, 16777217.0], np.float32),
        rng.normal(0, 1e4, 5000).astype(np.float32),
        (rng.normal(0, 1, 2000) * 10.0 ** rng.uniform(0, 12, 2000)).astype(np.float32)])
    got = debug_math(6, x).view(np.int32)
    want = np.array([lib.ora_f2i(float(v)) for v in x], dtype=np.int32)
    assert np.array_equal(got, want)


def test_config0_tiles_and_backward(oracle):
    cfg = scenes.CONFIG0
    W, H = cfg["width"], cfg["height"]
    splats = scenes.make_splats(cfg["n"], cfg["seed"])
    vp, vi = scenes.config0_camera()
    tgt, dep = _target(H, W)
    mgr = _run_backward(splats, vp, vi, tgt, dep, W, H)
    ranges, ids, full = mgr.read_tiles()
    o_ids, o_cnt, o_full = oracle.bin_tiles(splats, vp, W, H)
    lens = (ranges[:, 1] - ranges[:, 0]).astype(np.int64)
    assert np.array_equal(full.astype(np.int64), o_full), "uncapped tile counts differ"
    assert (lens >= np.minimum(o_full, 1024)).all() and (lens <= o_full).all()
    for t in range(ranges.shape[0]):
        got = ids[ranges[t, 0]: ranges[t, 0] + min(lens[t], 1024)]
        assert np.array_equal(got, o_ids[t, : o_cnt[t]]), f"tile {t} list differs"
    g_ref, c_ref, d_ref, loss_ref, _ = oracle.backward_f32(splats, vp, vi, tgt, dep, W, H)
    c, d = mgr.read_image()
    assert np.array_equal(c, c_ref)
    assert np.array_equal(d, d_ref)
    g = mgr.read_grads()
    assert np.array_equal(g, g_ref), f"{(g != g_ref).sum()} gradient entries differ"
    l2, l1 = mgr.read_loss()
    assert abs(l2 - loss_ref[0]) <= 1e-9 * abs(loss_ref[0]) and abs(l1 - loss_ref[1]) <= 1e-9 * abs(loss_ref[1])
    st = mgr.read_stats()
    assert st["tiles_over_cap"] == int((o_full > 1024).sum())
    # pairs = sum over tiles of (capped list length x pixels of the tile), summed from per-tile counter slots
    tx = (W + 15) // 16
    pix = np.array([min(16, W - 16 * (t % tx)) * min(16, H - 16 * (t // tx)) for t in range(len(o_cnt))], np.int64)
    assert st["pairs"] == int(np.dot(o_cnt.astype(np.int64), pix))
    assert 0 < st["active_pairs"] <= st["pairs"]
    mgr.close()


@pytest.mark.parametrize("W,H,n,cap", [(40, 24, 500, 1024), (70, 33, 3000, 64), (16, 16, 1, 1024)])
def test_ragged_and_foreground(oracle, W, H, n, cap):
    splats = scenes.make_splats(n, 7)
    splats[:, 0:3] *= 0.6
    vp, vi = camera.view_proj((0.3, 0.2, 2.5), aspect=W / H)
    tgt, dep = _target(H, W, seed=3, fg=True)
    mgr = _run_backward(splats, vp, vi, tgt, dep, W, H, cap=cap)
    g_ref, c_ref, d_ref, _, _ = oracle.backward_f32(splats, vp, vi, tgt, dep, W, H, cap=cap)
    c, d = mgr.read_image()
    assert np.array_equal(c, c_ref) and np.array_equal(d, d_ref)
    assert np.array_equal(mgr.read_grads(), g_ref)
    mgr.close()


def test_all_dead_scene(oracle):
    W, H = 32, 32
    splats = scenes.with_dead_capacity(scenes.make_splats(0, 0), 100)
    vp, vi = scenes.config0_camera()
    tgt, dep = _target(H, W)
    mgr = _run_backward(splats, vp, vi, tgt, dep, W, H)
    c, _ = mgr.read_image()
    assert np.allclose(c, 0.05) and mgr.read_stats()["keys"] == 0
    assert not mgr.read_grads().any()
    mgr.close()


def test_adam_step(oracle):
    n = 5000
    rng = np.random.default_rng(5)
    splats = scenes.make_splats(n, 2)
    grads = rng.integers(-200000, 200000, size=(n, 16)).astype(np.int32)
    m = (rng.normal(0, 1e-3, (n, 16))).astype(np.float32)
    v = (np.abs(rng.normal(0, 1e-6, (n, 16)))).astype(np.float32)
    acc = np.abs(rng.normal(0, 1e-3, n)).astype(np.float32)
    mgr = GpuSplatOptimizerManager(numSplats=n, splats=splats)
    mgr.write_grads(grads)
    mgr.write_adam(m, v, 17.0)
    mgr.write_accum(acc)
    mgr.step()
    s_ref, g_ref, m_ref, v_ref, a_ref, t_ref = oracle.step_f32(splats, grads, m, v, acc, 17.0)
    s, (m2, v2, t2), a = mgr.read_splats(), mgr.read_adam(), mgr.read_accum()
    assert t2 == t_ref == 18.0
    assert np.array_equal(s, s_ref) and np.array_equal(m2, m_ref) and np.array_equal(v2, v_ref)
    assert np.array_equal(a, a_ref) and not mgr.read_grads().any()
    mgr.close()


def _adc_state(n, seed, dead_frac=0.4, hot_frac=0.3):
    rng = np.random.default_rng(seed)
    s = scenes.make_splats(n, seed)
    s[:, 3] *= np.where(rng.random(n) < 0.5, 1.0, 0.2).astype(np.float32)
    dead = rng.random(n) < dead_frac
    s[dead, 7] = 0.0
    acc = np.where(rng.random(n) < hot_frac, rng.uniform(0.002, 0.01, n), rng.uniform(0, 0.0012, n))
    m = rng.normal(0, 1e-3, (n, 16)).astype(np.float32)
    v = np.abs(rng.normal(0, 1e-6, (n, 16))).astype(np.float32)
    return s, m, v, acc.astype(np.float32)


@pytest.mark.parametrize("n,seed,dead_frac", [(64, 1, 0.4), (5000, 2, 0.3), (20000, 3, 0.05), (3000, 4, 0.9)])
def test_adc_exact(oracle, n, seed, dead_frac):
    s, m, v, acc = _adc_state(n, seed, dead_frac)
    mgr = GpuSplatOptimizerManager(numSplats=n, splats=s)
    mgr.write_adam(m, v, 50.0)
    mgr.write_accum(acc)
    mgr.adc()
    s_ref, m_ref, v_ref, a_ref, st = oracle.adc_f32(s, m, v, acc, 50.0)
    s2, (m2, v2, _), a2 = mgr.read_splats(), mgr.read_adam(), mgr.read_accum()
    stats = mgr.read_stats()
    assert stats["adc_qualifiers"] == st[0] and stats["adc_children"] == st[1]
    assert stats["adc_splits"] == st[2] and stats["adc_clones"] == st[3]
    assert np.array_equal(s2, s_ref)
    assert np.array_equal(m2, m_ref) and np.array_equal(v2, v_ref)
    assert np.array_equal(a2, a_ref)
    mgr.close()


@pytest.mark.parametrize("quantize", [0, 1])
def test_render(oracle, quantize):
    W, H, n = 64, 40, 1500
    splats = scenes.make_splats(n, 11)
    splats[:, 0:3] *= 0.7
    vp, _ = camera.view_proj((0.0, 0.5, 2.8), aspect=W / H)
    opt = GpuSplatOptimizerManager(numSplats=n, splats=splats)
    fwd = GpuSplatForwardPipelineManager(numSplats=n, splatBuffer=opt)
    fwd.writeVPMatrix(vp)
    fwd.setTarget(None, None, W, H)
    fwd.dispatch(None, clear=True, quantize=bool(quantize))
    c, d = fwd.read_target()
    c_ref, d_ref = oracle.render_f32(splats, vp, W, H, clear=True, quantize=bool(quantize))
    assert np.array_equal(c, c_ref) and np.array_equal(d, d_ref)
    fwd.close()
    opt.close()


def test_dispatch_cadence(oracle):
    W, H, n = 32, 32, 800
    splats = scenes.with_dead_capacity(scenes.make_splats(400, 9), n)
    splats[:400, 0:3] *= 0.5
    vp, vi = scenes.config0_camera()
    tgt, dep = _target(H, W)
    mgr = GpuSplatOptimizerManager(numSplats=n, splats=splats, adc_period=3)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, dep, W, H)
    s = splats.copy()
    m = np.zeros_like(s)
    v = np.zeros_like(s)
    acc = np.zeros(n, np.float32)
    t = 0.0
    for k in range(1, 7):
        mgr.dispatch()
        g, _, _, _, _ = oracle.backward_f32(s, vp, vi, tgt, dep, W, H)
        s, _, m, v, acc, t = oracle.step_f32(s, g, m, v, acc, t)
        if k % 3 == 0:
            s, m, v, acc, _ = oracle.adc_f32(s, m, v, acc, t)
    assert np.array_equal(mgr.read_splats(), s)
    mgr.close()
