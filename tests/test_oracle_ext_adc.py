"""Pins of the EXT densification oracle: slot bookkeeping by brute force on
hand-built pools, clone = exact copy, the split's geometry against its closed
form (children at centre + R (z * s) with log-scales reduced by ln 1.6, z
standard normal), the normal generator's moments, Adam-state zeroing and the
accumulator reset."""
import math

import numpy as np

from nprush_amd import scenes


def _pool(n, seed=0):
    s, sh = scenes.make_ext_splats(n, seed, sh_degree=3)
    rng = np.random.default_rng(seed + 10)
    m = rng.standard_normal((n, 64)).astype(np.float32)
    v = np.abs(rng.standard_normal((n, 64))).astype(np.float32)
    return s, sh, m, v


def test_slot_bookkeeping(oracle):
    n = 12
    s, sh, m, v = _pool(n)
    s[:, 3] = 2.0                                   # live
    s[[1, 4, 9], 3] = -8.0                           # dead (op < 0.005)
    s[:, 4:7] = np.log(0.001)                        # small: clone
    acc = np.zeros((n, 2), np.float32)
    acc[[0, 2, 3, 5], 0] = [1e-3, 5e-4, 1e-5, 3e-4]
    acc[[0, 2, 3, 5], 1] = [2, 1, 1, 1]              # averages 5e-4, 5e-4, 1e-5, 3e-4
    acc[[6], 0] = 1.0                                 # count 0: never qualifies
    s2, sh2, m2, v2, a2, st = oracle.ext_adc_f32(s, sh, m, v, acc)
    # qualifiers 0, 2, 5 in index order -> dead slots 1, 4, 9
    assert list(st) == [3, 3, 0, 3, 3]
    for q, d in ((0, 1), (2, 4), (5, 9)):
        assert np.array_equal(s2[d], s[q]) and np.array_equal(sh2[d], sh[q])
        assert np.all(m2[d] == 0) and np.all(v2[d] == 0)
        assert np.array_equal(s2[q], s[q]) and np.array_equal(m2[q], m[q])
    assert np.all(a2 == 0)
    untouched = [3, 6, 7, 8, 10, 11]
    assert np.array_equal(s2[untouched], s[untouched])


def test_more_qualifiers_than_dead_slots_and_parking(oracle):
    n = 10
    s, sh, m, v = _pool(n, 1)
    s[:, 3] = 1.0
    s[[7], 3] = -9.0                                 # one dead slot
    s[[8], 3] = -7.0                                 # a second one
    s[:, 4:7] = np.log(0.001)
    acc = np.zeros((n, 2), np.float32)
    acc[:, 0] = 1.0
    acc[:, 1] = 1.0                                  # every live slot qualifies
    acc[[7, 8], 0] = 0.0
    s2, _, m2, v2, _, st = oracle.ext_adc_f32(s, sh, m, v, acc)
    assert list(st) == [8, 2, 0, 2, 2]
    assert np.array_equal(s2[7], s[0]) and np.array_equal(s2[8], s[1])
    # no dead slot left over here; with none qualifying every dead slot is parked
    acc[:] = 0
    s3, _, m3, v3, _, st = oracle.ext_adc_f32(s, sh, m, v, acc)
    assert list(st) == [0, 0, 0, 0, 2]
    assert s3[7, 3] == -10000.0 and s3[8, 3] == -10000.0
    assert np.all(m3[[7, 8]] == 0) and np.all(v3[[7, 8]] == 0)


def _rot(q):
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_split_geometry(oracle):
    n = 2
    s, sh, m, v = _pool(n, 2)
    s[0, 3] = 1.0
    s[1, 3] = -9.0
    s[0, 4:7] = np.log([0.2, 0.05, 0.1])            # max 0.2 > 0.01 * 3.3: split
    s[0, 8:12] = [0.9, 0.3, -0.2, 0.25]
    acc = np.array([[1.0, 1.0], [0.0, 0.0]], np.float32)
    s2, sh2, m2, v2, _, st = oracle.ext_adc_f32(s, sh, m, v, acc, seed=7)
    assert list(st) == [1, 1, 1, 0, 1]
    Rm = _rot(s[0, 8:12])
    sc = np.exp(s[0, 4:7].astype(np.float64))
    for slot, child in ((0, 0), (1, 1)):
        z = np.array([oracle.ext_adc_normal(7, slot, child, k) for k in range(3)])
        exp_pos = s[0, :3].astype(np.float64) + Rm @ (z * sc)
        assert np.allclose(s2[slot, :3], exp_pos, atol=1e-6)
        assert np.allclose(s2[slot, 4:7], s[0, 4:7].astype(np.float64) - math.log(1.6), atol=1e-6)
        assert np.array_equal(s2[slot, [3, 8, 9, 10, 11]], s[0, [3, 8, 9, 10, 11]])
        assert np.array_equal(sh2[slot], sh[0])
        assert np.all(m2[slot] == 0) and np.all(v2[slot] == 0)
    assert not np.array_equal(s2[0, :3], s2[1, :3])


def test_normals_moments(oracle):
    z = np.array([oracle.ext_adc_normal(3, i, c, k) for i in range(3000) for c in range(2) for k in range(3)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
    # fourth moment of a standard normal is 3
    assert abs((z ** 4).mean() - 3.0) < 0.25
