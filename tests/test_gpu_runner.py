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


def test_runner_scene_target_matches_oracle(oracle):
    """The frame's target from an attached scene (GpuRunner.svelte.ts:578-642):
    the raster (colour, encoded depth) and the path tracer's output, which
    accumulates while the camera holds still and restarts when it moves."""
    from nprush_amd import meshes

    runner = GpuRunner(400, 1.0, splats=scenes.with_dead_capacity(scenes.make_splats(200, 3), 400))
    W, H = runner.optim_width, runner.optim_height
    ms = meshes.scene(subdiv=2, seed=5)
    env = meshes.environment(64, 32)
    runner.attach_scene([ms[0], ms[2]], env)
    v = np.concatenate([ms[0][0], ms[2][0]])
    idx = np.concatenate([ms[0][1], ms[2][1] + len(ms[0][0])])
    nodes, tris = oracle.pt_bvh(v, idx)
    eyes = [(3.0, -2.5, 1.8), (-2.0, 3.0, 2.2)]
    for eye in eyes:
        vp, vi = camera.view_proj(eye, target=(0.0, 0.0, 0.6), up=(0.0, 0.0, 1.0), aspect=W / H)
        view = camera.look_at(eye, (0.0, 0.0, 0.6), (0.0, 0.0, 1.0))
        runner.set_camera(vp, vi, view)
        acc = None
        for f in range(2):                      # still camera: samples accumulate from frame 0
            pt_col, col, dep = runner.scene_textures()
            oc, od = oracle.mesh_raster(vp, view, vi, 1.0, W, H, v, nodes, tris, env)
            assert np.array_equal(col, oc) and np.array_equal(dep, od)
            acc = oracle.pt_trace(vi, f, W, H, v, nodes, tris, env, accum=acc)
            assert np.array_equal(pt_col, oracle.pt_resolve(acc))
        out_c, out_d = runner.frame(pt_col, dep)
        assert out_c.shape == (H, W, 4) and np.isfinite(out_c).all()
    runner.close()
