"""The frame schedule's host logic (CPU): the optimisation resolution rule of
GpuRunner.svelte.ts:284-296 (short side 256, long side round(256 * aspect),
JavaScript rounding of halves upwards)."""
from nprush_amd.runner import OPTIM_SHORT, optim_size


def test_optim_size_short_side_256_aspect_matched():
    assert OPTIM_SHORT == 256
    assert optim_size(1.0) == (256, 256)
    assert optim_size(16 / 9) == (455, 256)      # 455.11
    assert optim_size(4 / 3) == (341, 256)       # 341.33
    assert optim_size(9 / 16) == (256, 455)
    assert optim_size(2.5) == (640, 256)
    assert optim_size(1.001953125) == (257, 256)  # 256.5 rounds up, as Math.round does
