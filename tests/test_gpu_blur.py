"""Target blur (GpuBlurPipelineManager.blur, blur.wgsl:13-48) on the GPU:
bit-identical to the oracle on random 8-bit images, both colour spaces,
ragged sizes, the reference's radius 16 / sigma 8 and small radii."""
import numpy as np
import pytest

from nprush_amd import GpuBlurPipelineManager

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,radius,sigma,srgb", [
    (256, 256, 16, 8.0, True),
    (256, 144, 16, 8.0, False),
    (37, 19, 5, 2.5, True),
    (64, 48, 0, 0.0, True),      # sigma 0: weight NaN, as the reference computes it
    (200, 120, 1, 0.5, False),
])
def test_blur_matches_oracle(oracle, W, H, radius, sigma, srgb):
    rng = np.random.default_rng(W * 7 + H)
    img = (rng.integers(0, 256, (H, W, 4)) / 255.0).astype(np.float32)
    b = GpuBlurPipelineManager()
    got = b.blur(None, img, None, None, W, H, radius, sigma, srgb)
    exp = oracle.blur_f32(img, W, H, radius=radius, sigma=sigma, is_srgb=srgb)
    assert np.array_equal(got, exp)
    b.close()
