"""EXT path on the GPU against the EXT oracle (FP32 mode).

Bit-exact: tile lists (depth order, ids), per-tile counts, composited images,
Adam.  Gradients: the GPU sums the FP32 per-pair terms with FP32 wave sums and
FP64 across waves, the oracle in FP64 sequentially, so they agree to FP32
rounding of the per-wave sums: |g_gpu - g_ora| <= 1e-4 * max |g_ora| over the
parameter group (means, opacity, scales, quaternion, SH) plus 1e-5 |g_ora|.
"""
import numpy as np
import pytest

from nprush_amd import camera, scenes
from nprush_amd.managers import GpuSplatOptimizerManager

pytestmark = pytest.mark.gpu

GROUPS = {"means": [0, 1, 2], "opacity": [3], "scales": [4, 5, 6], "quat": [8, 9, 10, 11]}


def _target(oracle, seed, n, deg, vp, vi, W, H):
    s, sh = scenes.make_ext_splats(n, seed, sh_degree=deg)
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    zero = np.zeros((H, W, 4), np.float32)
    _, _, c, _ = oracle.ext_backward_f32(s, sh, deg, vp, vi, zero, W, H, tiles, lists, counts, want_grads=False)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = scenes.quantize8(c)
    return tgt


def _run(s, sh, deg, vp, vi, tgt, W, H):
    mgr = GpuSplatOptimizerManager(numSplats=len(s), mode="ext", sh_degree=deg, splats=s, sh=sh,
                                   adc_enabled=False)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, np.ones((H, W), np.float32), W, H)
    mgr.backward()
    return mgr


def _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H):
    ranges, ids, full = mgr.read_tiles()
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    assert np.array_equal(full.astype(np.int64), counts.astype(np.int64))
    for t in range(len(tiles)):
        a, b = ranges[t]
        assert b - a == counts[t]
        assert np.array_equal(ids[a:b], lists[t, :counts[t]]), t
    return tiles, lists, counts


def _check_grads(g_gpu, g, gs, deg):
    errs = {}
    for name, cols in GROUPS.items():
        ref = g[:, cols]
        got = g_gpu[:, cols].astype(np.float64)
        scale = np.abs(ref).max()
        err = np.abs(got - ref)
        assert np.all(err <= 1e-4 * scale + 1e-5 * np.abs(ref)), (name, err.max(), scale)
        errs[name] = float(err.max() / max(scale, 1e-30))
    nk = (deg + 1) ** 2
    ref = gs[:, :3 * nk]
    got = g_gpu[:, 16:16 + 3 * nk].astype(np.float64)
    scale = np.abs(ref).max()
    err = np.abs(got - ref)
    assert np.all(err <= 1e-4 * scale + 1e-5 * np.abs(ref)), ("sh", err.max(), scale)
    errs["sh"] = float(err.max() / max(scale, 1e-30))
    # untouched entries stay zero
    assert np.all(g_gpu[:, [7, 12, 13, 14, 15]] == 0)
    assert np.all(g_gpu[:, 16 + 3 * nk:] == 0)
    return errs


@pytest.mark.parametrize("deg", [0, 3])
def test_ext_config0_tiles_image_grads(oracle, deg):
    """Config-0 scene (16,384 Gaussians, 256x256, camera at (0,0,3), target from
    seed 1) in EXT mode."""
    W = H = 256
    n = 16384
    vp, vi = scenes.config0_camera()
    s, sh = scenes.make_ext_splats(n, 0, sh_degree=deg)
    tgt = _target(oracle, 1, n, deg, vp, vi, W, H)
    mgr = _run(s, sh, deg, vp, vi, tgt, W, H)
    tiles, lists, counts = _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H)
    g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tiles, lists, counts)
    img, _ = mgr.read_image()
    assert np.array_equal(img, c)
    l2, l1 = mgr.read_loss()
    assert abs(l1 - loss[0]) <= 1e-9 * loss[0] and abs(l2 - loss[1]) <= 1e-9 * loss[1]
    errs = _check_grads(mgr.read_grads_f32(), g, gs, deg)
    print("max relative gradient error per group:", errs)
    st = mgr.read_stats()
    assert st["keys"] == int(counts.sum())
    mgr.close()


def test_ext_ragged_offscreen_and_empty_tiles(oracle):
    """200x120 target (ragged tiles), splats behind the camera, beyond the image
    and with very low opacity, and tiles that no splat covers."""
    W, H = 200, 120
    deg = 2
    n = 3000
    rng = np.random.default_rng(11)
    s, sh = scenes.make_ext_splats(n, 4, sh_degree=deg)
    s[:, 0] = s[:, 0] * 0.5 - 0.4           # leave the right side of the image empty
    s[:200, 2] = 5.0                          # behind the camera (w < 0.1)
    s[200:300, 0] += 6.0                      # far off screen
    s[300:400, 3] = -8.0                      # opacity below 1/255
    vp, vi = camera.view_proj((0.0, 0.0, 3.0), aspect=W / H)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3)).astype(np.float32)
    mgr = _run(s, sh, deg, vp, vi, tgt, W, H)
    tiles, lists, counts = _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H)
    assert (counts == 0).any()
    g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tiles, lists, counts)
    img, _ = mgr.read_image()
    assert np.array_equal(img, c)
    ntx = (W + 15) // 16
    for t in np.nonzero(counts == 0)[0]:
        tx, ty = t % ntx, t // ntx
        assert np.all(img[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16] == np.float32(0.05))
    _check_grads(mgr.read_grads_f32(), g, gs, deg)
    mgr.close()


def test_ext_dense_early_termination(oracle):
    """Opaque, dense splats: most pixels stop early (T < 1e-4)."""
    W = H = 64
    deg = 1
    n = 4000
    s, sh = scenes.make_ext_splats(n, 7, sh_degree=deg)
    s[:, 0:3] *= 0.35
    s[:, 3] = 4.0                             # opacity ~0.98
    s[:, 4:7] = np.log(0.2)
    vp, vi = camera.view_proj((0.2, 0.1, 2.5), aspect=1.0)
    rng = np.random.default_rng(2)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3)).astype(np.float32)
    mgr = _run(s, sh, deg, vp, vi, tgt, W, H)
    tiles, lists, counts = _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H)
    g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tiles, lists, counts)
    img, _ = mgr.read_image()
    assert np.array_equal(img, c)
    st = mgr.read_stats()
    assert st["pairs"] == int(loss[2])                  # same stopping entry per pixel
    assert loss[2] < 0.5 * int(counts.sum()) * 256       # termination cut most pairs
    _check_grads(mgr.read_grads_f32(), g, gs, deg)
    mgr.close()


def test_ext_speculative_key_buffers_overflow_redo(oracle):
    """A view of the same shape as the last one sizes its key buffers from that
    view's count (+25 % + 64K) and keeps its own count on the device.  A view
    with far more keys overflows them: the dropped keys are flagged, the chain
    rule adds nothing, and the view is redone with buffers for its count.  The
    results (tile lists, image, gradients, densification statistics) are then
    those of the oracle, counted once; the next view of the first camera fits."""
    W = H = 256
    deg = 3
    n = 16384
    s, sh = scenes.make_ext_splats(n, 0, sh_degree=deg)
    s[:, 4:7] += 1.0                                  # larger splats: many keys up close
    far = camera.view_proj((0.0, 0.0, 20.0), aspect=1.0)
    near = camera.view_proj((0.0, 0.0, 2.0), aspect=1.0)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = 0.4
    mgr = _run(s, sh, deg, *far, tgt, W, H)          # first view: waits for its key count
    m_far = mgr.read_stats()["keys"]
    for k, (vp, vi) in enumerate((near, far)):
        mgr.write_grads_f32(np.zeros((n, 64), np.float32))
        mgr.write_accum(np.zeros_like(mgr.read_accum()))
        mgr.writeSplatVPMatrix(vp, vi)
        mgr.backward()
        tiles, lists, counts = _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H)
        if k == 0:                                    # the near view overflowed the far view's capacity
            assert counts.sum() > 1.25 * m_far + 65536
        assert mgr.read_stats()["keys"] == int(counts.sum())
        g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tiles, lists, counts)
        img, _ = mgr.read_image()
        assert np.array_equal(img, c)
        _check_grads(mgr.read_grads_f32(), g, gs, deg)
        on_screen = np.zeros(n, bool)
        for t in range(len(tiles)):
            on_screen[lists[t, :counts[t]]] = True
        acc = mgr.read_accum()
        assert np.array_equal(acc[:, 1], on_screen.astype(np.float32))  # one view counted once
    mgr.close()


def _drawn_depth_span(oracle, s, sh, deg, vp, vi, W, H, eye):
    """float-bit span of the depths (distance along the view axis, here -z of
    a camera on the +z axis) of the splats in some tile list"""
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    drawn = np.zeros(len(s), bool)
    for t in range(len(tiles)):
        drawn[lists[t, :counts[t]]] = True
    w = (np.float32(eye[2]) - s[drawn, 2]).astype(np.float32).view(np.uint32)
    return int(w.max()) - int(w.min())


def test_ext_narrow_depth_order_and_redo(oracle):
    """A view of the shape seen last sorts its depths in two 12-bit passes over
    the depth bits less the least drawn depth's when they span fewer than
    2^24 - 1 float steps; a wider span is flagged, the chain rule adds nothing,
    and the view is redone with the 32-bit sort (as are the next views of that
    shape).  Lists (pairs of splats at one position: equal depths, id order),
    image, gradients and statistics are the oracle's, each view counted once."""
    W = H = 128
    deg = 1
    n = 8192
    s, sh = scenes.make_ext_splats(n, 3, sh_degree=deg)
    s[1::2, 0:3] = s[0::2, 0:3]
    eyes = {"tight": (0.0, 0.0, 6.0), "wide": (0.0, 0.0, 1.3)}
    cams = {k: camera.view_proj(e, aspect=1.0) for k, e in eyes.items()}
    assert _drawn_depth_span(oracle, s, sh, deg, *cams["tight"], W, H, eyes["tight"]) < (1 << 23)
    assert _drawn_depth_span(oracle, s, sh, deg, *cams["wide"], W, H, eyes["wide"]) > (1 << 24) + (1 << 21)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = 0.3
    mgr = _run(s, sh, deg, *cams["tight"], tgt, W, H)
    for name in ("tight", "wide", "wide", "tight"):
        vp, vi = cams[name]
        mgr.write_grads_f32(np.zeros((n, 64), np.float32))
        mgr.write_accum(np.zeros_like(mgr.read_accum()))
        mgr.writeSplatVPMatrix(vp, vi)
        mgr.backward()
        tiles, lists, counts = _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H)
        assert mgr.read_stats()["keys"] == int(counts.sum())
        g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tiles, lists, counts)
        img, _ = mgr.read_image()
        assert np.array_equal(img, c), name
        _check_grads(mgr.read_grads_f32(), g, gs, deg)
        on_screen = np.zeros(n, bool)
        for t in range(len(tiles)):
            on_screen[lists[t, :counts[t]]] = True
        assert np.array_equal(mgr.read_accum()[:, 1], on_screen.astype(np.float32)), name
    mgr.close()


@pytest.mark.parametrize("n", [1, 37, 300])
def test_ext_small_scenes_speculative_views(oracle, n):
    """Scenes of fewer Gaussians than one preprocess workgroup, a camera that sees
    none of them, then one that sees them: each view twice, so the second runs
    speculatively (narrow depth order over one or two histogram blocks, no fill
    of the 2D gradients).  Lists, image and gradients are the oracle's."""
    W = H = 64
    deg = 2
    s, sh = scenes.make_ext_splats(n, 11, sh_degree=deg)
    s[:, 4:7] += 1.5
    away = camera.view_proj((0.0, 0.0, 3.0), target=(0.0, 0.0, 6.0), aspect=1.0)
    front = camera.view_proj((0.3, 0.2, 3.0), aspect=1.0)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = 0.5
    mgr = _run(s, sh, deg, *away, tgt, W, H)
    for vp, vi in (away, front, front):
        mgr.write_grads_f32(np.zeros((n, 64), np.float32))
        mgr.writeSplatVPMatrix(vp, vi)
        mgr.backward()
        tiles, lists, counts = _check_tiles(mgr, oracle, s, sh, deg, vp, vi, W, H)
        g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tiles, lists, counts)
        img, _ = mgr.read_image()
        assert np.array_equal(img, c)
        if counts.sum():
            _check_grads(mgr.read_grads_f32(), g, gs, deg)
        else:
            assert np.all(mgr.read_grads_f32() == 0)
    assert counts.sum() > 0                           # the front camera sees the scene
    mgr.close()


def test_ext_forward_only_and_accumulation(oracle):
    W = H = 96
    deg = 3
    n = 2000
    s, sh = scenes.make_ext_splats(n, 5, sh_degree=deg)
    vp, vi = camera.view_proj((0.5, -0.2, 2.8), aspect=1.0)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = 0.3
    mgr = _run(s, sh, deg, vp, vi, tgt, W, H)
    img_b, _ = mgr.read_image()
    loss_b = mgr.read_loss()
    g1 = mgr.read_grads_f32()
    acc1 = mgr.read_accum()
    mgr.forward()
    img_f, _ = mgr.read_image()
    assert np.array_equal(img_b, img_f)
    assert mgr.read_loss() == loss_b
    assert np.array_equal(mgr.read_grads_f32(), g1)     # forward adds nothing
    mgr.backward()
    g2 = mgr.read_grads_f32()
    assert np.allclose(g2, 2 * g1, rtol=1e-5, atol=1e-6 * np.abs(g1).max())
    acc2 = mgr.read_accum()
    assert np.array_equal(acc2[:, 1], 2 * acc1[:, 1])
    _, _, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    on_screen = np.zeros(n, bool)
    for t in range(len(tiles)):
        on_screen[lists[t, :counts[t]]] = True
    assert np.array_equal(acc1[:, 1] > 0, on_screen)
    mgr.step()
    assert np.all(mgr.read_grads_f32() == 0)
    mgr.close()


def test_ext_adam_bit_exact(oracle):
    n = 1500
    deg = 2
    rng = np.random.default_rng(9)
    s, sh = scenes.make_ext_splats(n, 2, sh_degree=deg)
    g = (rng.standard_normal((n, 64)) * 10.0 ** rng.uniform(-6, 1, (n, 64))).astype(np.float32)
    m = (rng.standard_normal((n, 64)) * 1e-3).astype(np.float32)
    v = np.abs(rng.standard_normal((n, 64)) * 1e-4).astype(np.float32)
    mgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=deg, splats=s, sh=sh, adc_enabled=False)
    mgr.write_adam(m, v, 7.0)
    mgr.write_grads_f32(g)
    mgr.step()
    es, esh, em, ev, et = oracle.ext_adam_f32(s, sh, g, m, v, deg, 7.0)
    gm, gv, gt = mgr.read_adam()
    assert gt == et == 8.0
    assert np.array_equal(mgr.read_splats(), es)
    assert np.array_equal(mgr.read_sh(), esh)
    assert np.array_equal(gm, em) and np.array_equal(gv, ev)
    assert np.all(mgr.read_grads_f32() == 0)
    mgr.close()


@pytest.mark.parametrize("n,dead_frac,seed", [(5000, 0.3, 0), (3000, 0.02, 1), (64, 0.5, 2)])
def test_ext_adc_bit_exact(oracle, n, dead_frac, seed):
    """EXT densification on the GPU against the oracle, two calls in a row
    (the second uses the next seed), with more or fewer dead slots than
    qualifiers and a mix of splits and clones."""
    rng = np.random.default_rng(seed)
    s, sh = scenes.make_ext_splats(n, seed, sh_degree=3)
    s[rng.random(n) < dead_frac, 3] = -9.0
    s[:, 4:7] += rng.normal(0.0, 1.0, (n, 3)).astype(np.float32)   # scales spread over the split threshold
    m = rng.standard_normal((n, 64)).astype(np.float32)
    v = np.abs(rng.standard_normal((n, 64))).astype(np.float32)
    mgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=3, splats=s, sh=sh, adc_enabled=False)
    mgr.set_adc_ext(2e-4, 3.3, 0.005, 0.01)
    mgr.write_adam(m, v, 3.0)
    for call in range(2):
        acc = np.zeros((n, 2), np.float32)
        acc[:, 1] = rng.integers(0, 4, n)
        acc[:, 0] = (rng.random(n) * 6e-4 * acc[:, 1]).astype(np.float32)
        mgr.write_accum(acc)
        mgr.adc()
        s, sh, m, v, acc_o, st = oracle.ext_adc_f32(s, sh, m, v, acc, 2e-4, 3.3, 0.005, 0.01, seed=call)
        assert np.array_equal(mgr.read_splats(), s)
        assert np.array_equal(mgr.read_sh(), sh)
        gm, gv, _ = mgr.read_adam()
        assert np.array_equal(gm, m) and np.array_equal(gv, v)
        assert np.all(mgr.read_accum() == 0)
        g = mgr.read_stats()
        assert [g["adc_qualifiers"], g["adc_children"], g["adc_splits"], g["adc_clones"], g["adc_rounds"]] == \
            [int(st[0]), int(st[1]), int(st[2]), int(st[3]), int(st[4])]
    mgr.close()


def test_ext_training_reduces_l1(oracle):
    """Config-0 scene in EXT mode (SH degree 0, 256x256, target from seed 1):
    60 dispatches (backward + Adam, densification every 25 with a dead pool)
    lower the L1 loss."""
    W = H = 256
    n_live, cap = 16384, 20480
    vp, vi = scenes.config0_camera()
    tgt = _target(oracle, 1, n_live, 0, vp, vi, W, H)
    s, sh = scenes.make_ext_splats(n_live, 0, sh_degree=0)
    pool = np.zeros((cap, 16), np.float32)
    pool[:n_live] = s
    pool[n_live:, 3] = -10000.0
    pool[n_live:, 4:7] = -4.0
    pool[n_live:, 8] = 1.0
    shp = np.zeros((cap, 48), np.float32)
    shp[:n_live] = sh
    mgr = GpuSplatOptimizerManager(numSplats=cap, mode="ext", sh_degree=0, splats=pool, sh=shp, adc_period=25)
    mgr.set_adc_ext(2e-4, 3.3, 0.005, 0.01)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, np.ones((H, W), np.float32), W, H)
    losses = []
    for k in range(60):
        mgr.dispatch()
        losses.append(mgr.read_loss()[1])
        if k == 24:
            st = mgr.read_stats()
            assert st["adc_rounds"] >= cap - n_live        # the dead pool was seen
    assert losses[-1] < 0.9 * losses[0], (losses[0], losses[-1])
    mgr.close()


@pytest.mark.parametrize("tile,batch", [(8, 32), (8, 128), (16, 32), (16, 128), (32, 32), (32, 64), (32, 128)])
def test_ext_tile_and_batch_variants(oracle, tile, batch):
    """Every composite variant the autotuner can choose: tile lists bit-identical
    to the oracle's at that tile size, images bit-identical to the 16-pixel
    oracle image (a pixel's result does not depend on the tiling), gradients
    within tolerance."""
    W, H = 200, 136
    deg = 3
    n = 6000
    s, sh = scenes.make_ext_splats(n, 8, sh_degree=deg)
    vp, vi = camera.view_proj((0.4, -0.3, 2.7), aspect=W / H)
    rng = np.random.default_rng(5)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3)).astype(np.float32)
    mgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=deg, splats=s, sh=sh, adc_enabled=False,
                                   ext_tile=tile, ext_batch=batch)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, np.ones((H, W), np.float32), W, H)
    mgr.backward()
    ranges, ids, full = mgr.read_tiles()
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H, tile=tile)
    assert np.array_equal(full.astype(np.int64), counts.astype(np.int64))
    for t in range(len(tiles)):
        a, b = ranges[t]
        assert np.array_equal(ids[a:b], lists[t, :counts[t]]), t
    t16, l16, c16 = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, t16, l16, c16)
    img, _ = mgr.read_image()
    assert np.array_equal(img, c)
    _check_grads(mgr.read_grads_f32(), g, gs, deg)
    mgr.close()


def test_ext_autotune(oracle):
    W, H = 320, 200
    n = 8000
    s, sh = scenes.make_ext_splats(n, 9, sh_degree=1)
    vp, vi = camera.view_proj((0.0, 0.5, 3.0), aspect=W / H)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = 0.4
    mgr = _run(s, sh, 1, vp, vi, tgt, W, H)
    img0, _ = mgr.read_image()
    g0 = mgr.read_grads_f32()
    tile, batch, ms = mgr.autotune()
    assert tile in (8, 16, 32) and batch in (32, 64, 128)
    assert np.all(ms > 0) and ms[(8, 16, 32).index(tile), (32, 64, 128).index(batch)] == ms.min()
    assert np.array_equal(mgr.read_grads_f32(), g0)      # timing runs add no gradients
    mgr.write_grads_f32(np.zeros_like(g0))
    mgr.backward()
    img1, _ = mgr.read_image()
    assert np.array_equal(img1, img0)
    # the timing runs' 2D gradients (no chain rule after them) do not leak into
    # the next view: its gradients are the first view's (FP32 sums of FP64
    # accumulators in another order)
    g1 = mgr.read_grads_f32()
    assert np.all(np.abs(g1 - g0) <= 1e-4 * np.abs(g0).max(axis=0) + 1e-5 * np.abs(g0))
    mgr.backward()                                        # again, without a fill in between
    g2 = mgr.read_grads_f32()
    assert np.all(np.abs(g2 - 2 * g0) <= 2e-4 * np.abs(g0).max(axis=0) + 1e-5 * np.abs(g0))
    mgr.close()


@pytest.mark.parametrize("tile,batch", [(8, 32), (16, 64), (32, 128)])
def test_ext_block_cull_shapes(oracle, tile, batch):
    """Shapes that probe the forward's per-wave block test (3-sigma box, then a
    lower bound of r^2 over the 8x8 block against the alpha cut): needles (axis
    ratios up to e^4), sub-pixel and tens-of-pixel splats, opacities from just
    above the 1/255 cut to 0.99, centres spread over sub-pixel offsets.  A block
    left out wrongly would drop a composited pair: images, the evaluated-pair
    count and gradients against the oracle."""
    W, H = 176, 144
    deg = 1
    n = 5000
    rng = np.random.default_rng(23)
    s, sh = scenes.make_ext_splats(n, 9, sh_degree=deg)
    s[:, 0:3] *= 0.6
    ls = rng.uniform(-7.0, -2.5, (n, 3))
    needle = rng.random(n) < 0.5
    ls[needle, 1] = ls[needle, 0] - rng.uniform(2.0, 4.0, needle.sum())
    s[:, 4:7] = ls
    op = rng.uniform(1.05 / 255.0, 0.99, n)
    s[:, 3] = np.log(op / (1.0 - op))
    vp, vi = camera.view_proj((0.3, 0.2, 2.6), aspect=W / H)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3)).astype(np.float32)
    mgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=deg, splats=s, sh=sh, adc_enabled=False,
                                   ext_tile=tile, ext_batch=batch)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, np.ones((H, W), np.float32), W, H)
    mgr.backward()
    # the oracle at the same tile size: evaluated pairs depend on the lists
    tl, ll, cl = oracle.ext_bin(s, sh, deg, vp, vi, W, H, tile=tile)
    g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, tl, ll, cl, tile=tile)
    img, _ = mgr.read_image()
    assert np.array_equal(img, c)
    st = mgr.read_stats()
    assert st["pairs"] == int(loss[2])
    assert st["active_pairs"] == int(loss[3])       # composited pairs, as the reverse pass takes them
    _check_grads(mgr.read_grads_f32(), g, gs, deg)
    mgr.close()


@pytest.mark.parametrize("W,H", [(2064, 48), (48, 2064)])
def test_ext_key_path_beyond_256_tiles(oracle, W, H):
    """A tile grid with more than 256 columns (or rows) at 8-pixel tiles is
    binned by the key path (tile << 32 | id keys in depth order, a stable sort
    on the tile bits) instead of the row/column buckets: same lists, image and
    gradients as the oracle."""
    deg = 1
    n = 5000
    s, sh = scenes.make_ext_splats(n, 21, sh_degree=deg)
    s[:, 0] *= 1.0 if W > H else 0.05      # spread the scene over the long side
    s[:, 1] *= 1.0 if H > W else 0.05
    vp, vi = camera.view_proj((0.0, 0.0, 2.5), aspect=W / H)
    rng = np.random.default_rng(6)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3)).astype(np.float32)
    mgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=deg, splats=s, sh=sh, adc_enabled=False,
                                   ext_tile=8, ext_batch=32)
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.setBackwardTarget(tgt, np.ones((H, W), np.float32), W, H)
    mgr.backward()
    ranges, ids, full = mgr.read_tiles()
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H, tile=8)
    assert max((W + 7) // 8, (H + 7) // 8) > 256 and counts.sum() > 0
    assert np.array_equal(full.astype(np.int64), counts.astype(np.int64))
    for t in range(len(tiles)):
        a, b = ranges[t]
        assert np.array_equal(ids[a:b], lists[t, :counts[t]]), t
    t16, l16, c16 = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    g, gs, c, loss = oracle.ext_backward_f32(s, sh, deg, vp, vi, tgt, W, H, t16, l16, c16)
    img, _ = mgr.read_image()
    assert np.array_equal(img, c)
    _check_grads(mgr.read_grads_f32(), g, gs, deg)
    mgr.close()
