"""Frames of the reference's schedule (GpuRunner.svelte.ts:284-296, 644-709) on
the GPU against the oracle, bit for bit: optimisation resolution 455 x 256 for
a 16:9 panel, sRGB blur of the colour target and linear blur of the encoded
depth (radius 16, sigma 8), setBackwardTarget -> dispatch (backward, Adam,
densification on its cadence) -> forward render at the optimisation
resolution (clear, rgba8 rounding per blend); the camera alternates between
two ring views, so the VP is rewritten every frame.  The target is a second
splat set rendered by the same raster (the reference's mesh assets are not
shipped)."""
import numpy as np
import pytest

from nprush_amd import GpuSplatForwardPipelineManager, GpuSplatOptimizerManager, camera, scenes
from nprush_amd.runner import GpuRunner

pytestmark = pytest.mark.gpu


def test_runner_frames_match_oracle(oracle):
    aspect = 16 / 9
    n_live, n = 600, 1200
    target_set = scenes.make_splats(n_live, 21)
    target_set[:, 0:3] *= 0.6
    learner = scenes.with_dead_capacity(scenes.make_splats(n_live, 20), n)
    learner[:n_live, 0:3] *= 0.6
    runner = GpuRunner(n, aspect, splats=learner, adc_period=2)
    W, H = runner.optim_width, runner.optim_height
    assert (W, H) == (455, 256)
    cams = camera.ring_cameras(4, 3.0, W, H)[:2]
    # the "model render": the target set through the same raster at the optimisation resolution
    tm = GpuSplatOptimizerManager(numSplats=n_live, splats=target_set)
    tf = GpuSplatForwardPipelineManager(splatBuffer=tm)
    targets = []
    for vp, vi in cams:
        tf.setTarget(None, None, W, H)
        tf.writeVPMatrix(vp)
        tf.dispatch(None, clear=True)
        c, d = tf.read_target()
        ec, ed = oracle.render_f32(target_set, vp, W, H)
        assert np.array_equal(c, ec) and np.array_equal(d, ed)
        targets.append((c, d))
    tf.close()
    tm.close()

    s = learner.copy()
    m = np.zeros_like(s)
    v = np.zeros_like(s)
    acc = np.zeros(n, np.float32)
    t = 0.0
    for k in range(1, 6):
        vp, vi = cams[k % 2]
        tc, td = targets[k % 2]
        runner.set_camera(vp, vi)
        out_c, out_d = runner.frame(tc, td)
        bc = oracle.blur_f32(tc, W, H, radius=16, sigma=8.0, is_srgb=True)
        bd = oracle.blur_f32(td, W, H, radius=16, sigma=8.0, is_srgb=False)
        gc, gd = runner.prepare_target(tc, td)
        assert np.array_equal(gc, bc) and np.array_equal(gd, bd), k
        g, _, _, _, _ = oracle.backward_f32(s, vp, vi, bc, bd[..., 0], W, H)
        s, _, m, v, acc, t = oracle.step_f32(s, g, m, v, acc, t)
        if k % 2 == 0:
            s, m, v, acc, _ = oracle.adc_f32(s, m, v, acc, t)
        ec, ed = oracle.render_f32(s, vp, W, H)
        assert np.array_equal(runner.optimizer.read_splats(), s), k
        assert np.array_equal(out_c, ec) and np.array_equal(out_d, ed), k
    runner.close()
