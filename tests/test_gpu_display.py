"""The debug display compositor (splat_render.wgsl) on the GPU through
npd_compose against the oracle, bit for bit: random textures of different
sizes and aspect ratios (letter- and pillarboxed panels), path-trace samples
present in part of the image, every layer combination, several canvas sizes."""
import numpy as np
import pytest

from nprush_amd import compose_display

pytestmark = pytest.mark.gpu


def _compose_gpu(textures, enabled, W, H):
    return compose_display(textures, enabled, W, H)


@pytest.mark.parametrize("W,H", [(1000, 500), (1920, 1080), (737, 911)])
def test_compose_matches_oracle(oracle, W, H):
    rng = np.random.default_rng(W + H)
    tw, th = max(1, W // 2), max(1, int(H * 0.82))
    sizes = [(tw, th), (tw, th), (256, 256), (320, 180), (tw, th), (tw // 2 + 1, th // 2 + 1), (tw, th), (455, 256)]
    tex = [rng.random((h, w, 4)).astype(np.float32) * 1.2 - 0.1 for (w, h) in sizes]
    tex[7][..., 3] = np.where(rng.random(tex[7].shape[:2]) < 0.5, 0.0, tex[7][..., 3])
    for enabled in [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)]:
        got = _compose_gpu(tex, enabled, W, H)
        exp = oracle.compose_f32(tex, enabled, W, H)
        assert np.array_equal(got, exp), (W, H, enabled, int((got != exp).any(axis=2).sum()))
