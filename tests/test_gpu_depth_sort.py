"""EXT depth order on the GPU (nps_debug_depth_sort): the 32-bit pass sort and
the two-pass range-compressed sort of speculative views against a stable sort
on the depth bits (numpy), at ragged sizes, with ties, splats not drawn, and
depth spans on both sides of the compressed range's limit (2^24 - 1 float
steps above the least drawn depth)."""
import numpy as np
import pytest

from nprush_amd.managers import debug_depth_sort

pytestmark = pytest.mark.gpu

CAP = (1 << 24) - 1
CULLED = 0xFFFFFFFF


def _keys(hi):
    hi = np.asarray(hi, np.uint64)
    return (hi << np.uint64(32)) | np.arange(hi.size, dtype=np.uint64)


def _stable(keys):
    return keys[np.argsort(keys >> np.uint64(32), kind="stable")]


@pytest.mark.parametrize("n", [1, 255, 4096, 3 * 4096 + 17, 100_003])
def test_depth_sort_ragged_ties_culled(n):
    rng = np.random.default_rng(n)
    base = np.float32(1.3).view(np.uint32)
    hi = base + rng.integers(0, 4000, size=n).astype(np.uint32) * np.uint32(3000)   # many ties
    hi[rng.random(n) < 0.2] = CULLED
    keys = _keys(hi)
    ref = _stable(keys)
    for narrow in (False, True):
        got, fl = debug_depth_sort(keys, narrow)
        assert np.array_equal(got, ref), narrow
        assert fl.tolist() == [0, 0]


def test_depth_sort_range_limit():
    rng = np.random.default_rng(7)
    n = 50_000
    base = np.float32(0.75).view(np.uint32)
    hi = base + rng.integers(0, CAP, size=n).astype(np.uint32)
    hi[10] = base
    hi[20] = base + CAP - 1                         # the widest span the compressed code keeps
    hi[30:40] = CULLED
    keys = _keys(hi)
    got, fl = debug_depth_sort(keys, True)
    assert fl.tolist() == [0, 0]
    assert np.array_equal(got, _stable(keys))
    hi[25] = base + CAP                             # one step wider: flagged
    keys = _keys(hi)
    got, fl = debug_depth_sort(keys, True)
    assert fl.tolist() == [1, 1]
    assert np.array_equal(np.sort(got), np.sort(keys))   # still a permutation of the keys
    got, fl = debug_depth_sort(keys, False)        # the 32-bit sort has no limit
    assert fl.tolist() == [0, 0]
    assert np.array_equal(got, _stable(keys))


def test_depth_sort_all_culled_and_empty():
    keys = _keys(np.full(5000, CULLED, np.uint32))
    for narrow in (False, True):
        got, fl = debug_depth_sort(keys, narrow)
        assert np.array_equal(got, keys) and fl.tolist() == [0, 0]
        got, fl = debug_depth_sort(keys[:0], narrow)
        assert got.size == 0 and fl.tolist() == [0, 0]
