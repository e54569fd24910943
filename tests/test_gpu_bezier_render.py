"""Display raster of the curves on the GPU (GpuBezierForwardPipelineManager over
nbf_*) against the CPU oracle's restatement of bezier_forward.wgsl and its blend
state, bit for bit: cleared and loaded targets, with and without the per-blend
8-bit rounding, curves behind the camera, near-transparent curves, ragged target
sizes, an empty curve set, and two targets of different size in a row."""
import numpy as np
import pytest

from nprush_amd import GpuBezierForwardPipelineManager, GpuBezierOptimizerManager, camera, scenes

pytestmark = pytest.mark.gpu


def _brush(bw, bh, seed):
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 256, (bh, bw, 4), dtype=np.uint8)
    b[..., 3] = 255
    return b


def _curves(n, seed):
    c = scenes.make_bezier_curves(n, seed, radius_median=0.03, walk_std=0.15)
    if n >= 13:
        c[::13, 19] = 0.0005   # below the draw threshold everywhere
        c[5::29, 10] = 4.0     # p2 behind the camera (eye at z = 3): not drawn
        c[7::31, 3] = 0.0      # width floored at 1e-4
    return c


def _setup(n, W, H, seed, brush):
    curves = _curves(n, seed)
    vp, _ = camera.view_proj((0.4, 0.3, 3.0), aspect=W / H)
    m = GpuBezierOptimizerManager(numBeziers=n)
    m.write_beziers(curves)
    f = GpuBezierForwardPipelineManager(numBeziers=n, bezierBuffer=m, brushTexture=brush)
    f.writeVPMatrix(vp)
    return m, f, curves, vp


@pytest.mark.parametrize("W,H,n,quantize", [
    (96, 64, 400, True),
    (83, 45, 300, False),    # ragged tiles
    (256, 192, 3000, True),
    (16, 16, 1, True),
])
def test_render_matches_oracle(oracle, W, H, n, quantize):
    print(f"bezier render W={W} H={H} n={n} quantize={quantize}", flush=True)
    brush = _brush(32, 16, n)
    m, f, curves, vp = _setup(n, W, H, 100 + n, brush)
    f.setTarget(None, W, H)
    f.dispatch(clear=True, quantize=quantize)
    got = f.read_target()
    exp = oracle.bz_render_f32(curves, vp, W, H, brush, clear=True, quantize=quantize)
    assert np.array_equal(got, exp), f"{(got != exp).any(-1).sum()} pixels differ"
    if n > 1:
        assert (got[..., 3] > 0).sum() > 0.02 * W * H
    # drawn again over a loaded image
    base = np.random.default_rng(n).random((H, W, 4)).astype(np.float32)
    f.setTarget(base, W, H)
    f.dispatch(clear=False, quantize=quantize)
    got = f.read_target()
    exp = oracle.bz_render_f32(curves, vp, W, H, brush, rgba=base, clear=False, quantize=quantize)
    assert np.array_equal(got, exp), f"{(got != exp).any(-1).sum()} pixels differ (load)"
    f.close()
    m.close()


def test_render_two_sizes_in_a_row(oracle):
    print("bezier render 64x48 then 40x40", flush=True)
    brush = _brush(16, 8, 3)
    m, f, curves, vp = _setup(500, 64, 48, 7, brush)
    for W, H in ((64, 48), (40, 40)):
        vp, _ = camera.view_proj((0.4, 0.3, 3.0), aspect=W / H)
        f.writeVPMatrix(vp)
        f.setTarget(None, W, H)
        f.dispatch(clear=True, quantize=True)
        exp = oracle.bz_render_f32(curves, vp, W, H, brush, clear=True, quantize=True)
        assert np.array_equal(f.read_target(), exp), (W, H)
    f.close()
    m.close()


def test_render_empty_curve_set():
    print("bezier render with 0 curves", flush=True)
    m = GpuBezierOptimizerManager(numBeziers=0)
    f = GpuBezierForwardPipelineManager(numBeziers=0, bezierBuffer=m, brushTexture=_brush(4, 4, 0))
    vp, _ = camera.view_proj((0.4, 0.3, 3.0), aspect=1.0)
    f.writeVPMatrix(vp)
    base = np.random.default_rng(1).random((32, 32, 4)).astype(np.float32)
    f.setTarget(base, 32, 32)
    f.dispatch(clear=False, quantize=False)
    assert np.array_equal(f.read_target(), base)
    f.dispatch(clear=True, quantize=True)
    assert np.all(f.read_target() == 0.0)
    f.close()
    m.close()
