"""Pins of the EXT oracle (CPU): its FP64 gradients against central finite
differences of the L1 loss it defines, the SH basis against the analytic real
spherical harmonics, and closed forms of the depth order and early stop."""
import math

import numpy as np
import pytest

from nprush_amd import camera, scenes


def _scene(n=10, seed=3, deg=3):
    s, sh = scenes.make_ext_splats(n, seed, sh_degree=deg, higher_std=0.3)
    s[:, 0:3] *= 0.3
    s[:, 4:7] = np.log(np.random.default_rng(seed).uniform(0.04, 0.1, (n, 3)))
    s[:, 3] = np.random.default_rng(seed + 1).uniform(-1.0, 0.5, n)
    return s, sh


@pytest.mark.parametrize("deg", [0, 3])
def test_ext_gradients_match_finite_differences(oracle, deg):
    W = H = 24
    s, sh = _scene(deg=deg)
    vp, vi = camera.view_proj((0.4, 0.3, 2.7), aspect=1.0)
    rng = np.random.default_rng(5)
    tgt = np.zeros((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3))
    tiles, lists, counts = oracle.ext_bin(s, sh, deg, vp, vi, W, H)
    cam = oracle.cam_from_inv(vi)
    args = (deg, vp.astype(np.float64), cam, tgt, W, H, tiles, lists, counts)
    s64, sh64 = s.astype(np.float64), sh.astype(np.float64)
    g, gs, _, _ = oracle.ext_backward_f64(s64, sh64, *args)
    h = 1e-6

    def loss(a, b):
        return oracle.ext_backward_f64(a, b, *args, want_grads=False)[3][0]

    nk = (deg + 1) ** 2
    checked = 0
    for i in range(len(s)):
        for k in [0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 11]:
            sp, sm = s64.copy(), s64.copy()
            sp[i, k] += h
            sm[i, k] -= h
            fd = (loss(sp, sh64) - loss(sm, sh64)) / (2 * h)
            assert abs(fd - g[i, k]) <= 2e-4 * max(1.0, abs(fd)), (i, k, fd, g[i, k])
            checked += 1
        for k in range(3 * nk):
            hp, hm = sh64.copy(), sh64.copy()
            hp[i, k] += h
            hm[i, k] -= h
            fd = (loss(s64, hp) - loss(s64, hm)) / (2 * h)
            assert abs(fd - gs[i, k]) <= 2e-4 * max(1.0, abs(fd)), (i, "sh", k, fd, gs[i, k])
    assert checked > 100


def _real_sh(x, y, z):
    pi = math.pi
    return [0.5 / math.sqrt(pi),
            -math.sqrt(3 / (4 * pi)) * y, math.sqrt(3 / (4 * pi)) * z, -math.sqrt(3 / (4 * pi)) * x,
            0.5 * math.sqrt(15 / pi) * x * y, -0.5 * math.sqrt(15 / pi) * y * z,
            0.25 * math.sqrt(5 / pi) * (3 * z * z - 1), -0.5 * math.sqrt(15 / pi) * x * z,
            0.25 * math.sqrt(15 / pi) * (x * x - y * y),
            -0.25 * math.sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y), 0.5 * math.sqrt(105 / pi) * x * y * z,
            -0.25 * math.sqrt(21 / (2 * pi)) * y * (5 * z * z - 1), 0.25 * math.sqrt(7 / pi) * z * (5 * z * z - 3),
            -0.25 * math.sqrt(21 / (2 * pi)) * x * (5 * z * z - 1), 0.25 * math.sqrt(105 / pi) * z * (x * x - y * y),
            -0.25 * math.sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y)]


def test_sh_basis_matches_analytic_harmonics(oracle):
    """A splat with a single SH coefficient c_k = 1 has colour 0.5 + Y_k(d) for the
    real orthonormal harmonics in Cartesian form (written independently here):
    the rendered pixel is a*(0.5 + Y_k) + (1-a)*0.05, a read off a Y = 0 render."""
    rng = np.random.default_rng(2)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    cam = np.array([0.0, 0.0, 3.0])
    pos = cam + 2.5 * d
    vp, vi = camera.view_proj(tuple(cam), target=tuple(pos), aspect=1.0)
    camc = oracle.cam_from_inv(vi)
    W = H = 16
    s = np.zeros((1, 16), np.float64)
    s[0, 0:3] = np.float32(pos)
    s[0, 3] = 20.0
    s[0, 4:7] = np.log(0.3)
    s[0, 8] = 1.0
    dd = s[0, 0:3] - camc
    dd /= np.linalg.norm(dd)
    Y = _real_sh(*dd)
    zero = np.zeros((1, 48), np.float64)
    tiles, lists, counts = oracle.ext_bin(s.astype(np.float32), zero.astype(np.float32), 3, vp, vi, W, H)
    render = lambda sh: oracle.ext_backward_f64(s, sh, 3, vp.astype(np.float64), camc,
                                                np.zeros((H, W, 4), np.float32), W, H, tiles, lists, counts,
                                                want_grads=False)[2][8, 8, 0]
    a = (render(zero) - 0.05) / 0.45
    assert 0.5 < a <= 0.99
    for k in range(16):
        sh = zero.copy()
        sh[0, 3 * k] = 1.0
        assert abs(render(sh) - (a * max(0.5 + Y[k], 0.0) + (1 - a) * 0.05)) < 1e-12, k


def test_ext_depth_order_and_early_stop(oracle):
    """Compositing runs in increasing view depth whatever the index order, and a
    pixel stops before the splat that would take T below 1e-4: three large
    opaque splats (alpha clamped to 0.99) -- blue in front, then green, then red.
    After blue T = 1 - 0.99 (0.00999999 in FP32); green would leave
    9.99998e-5 < 1e-4, so only blue and the background are composited."""
    W = H = 16
    vp, vi = camera.view_proj((0.0, 0.0, 3.0), aspect=1.0)
    s = np.zeros((3, 16), np.float32)
    sh = np.zeros((3, 48), np.float32)
    for i, (z, rgb) in enumerate([(0.5, (0, 0, 1)), (-0.5, (1, 0, 0)), (0.0, (0, 1, 0))]):
        s[i, 2] = z
        s[i, 3] = 20.0
        s[i, 4:7] = np.log(3.0)
        s[i, 8] = 1.0
        sh[i, 0:3] = (np.array(rgb) - 0.5) / scenes.SH_C0
    tiles, lists, counts = oracle.ext_bin(s, sh, 0, vp, vi, W, H)
    assert list(lists[0, :counts[0]]) == [0, 2, 1]  # depth 2.5 < 3.0 < 3.5
    _, _, c, _ = oracle.ext_backward_f32(s, sh, 0, vp, vi, np.zeros((H, W, 4), np.float32), W, H, tiles, lists,
                                         counts, want_grads=False)
    px = c[8, 8].astype(np.float64)
    assert abs(px[2] - (0.99 + 0.05 * 0.01)) < 1e-6
    assert abs(px[1] - 0.05 * 0.01) < 1e-7 and abs(px[0] - 0.05 * 0.01) < 1e-7


def test_images_do_not_depend_on_tile_size(oracle):
    """Each pixel applies the splat's 3-sigma box itself and the binning box is
    wider, so tile lists of any size are supersets of what a pixel can see:
    images are identical for 8, 16 and 32 pixel tiles (gradients up to the FP64
    summation order)."""
    W, H = 100, 72
    s, sh = scenes.make_ext_splats(1500, 6, sh_degree=1)
    s[:, 0:3] *= 0.6
    vp, vi = camera.view_proj((0.3, 0.2, 2.6), aspect=W / H)
    rng = np.random.default_rng(3)
    tgt = np.ones((H, W, 4), np.float32)
    tgt[..., :3] = rng.random((H, W, 3))
    out = {}
    for tile in (8, 16, 32):
        tiles, lists, counts = oracle.ext_bin(s, sh, 1, vp, vi, W, H, tile=tile)
        out[tile] = oracle.ext_backward_f32(s, sh, 1, vp, vi, tgt, W, H, tiles, lists, counts, tile=tile)
    g16, gs16, c16, l16 = out[16]
    for tile in (8, 32):
        g, gs, c, l = out[tile]
        assert np.array_equal(c, c16)
        assert abs(l[0] - l16[0]) <= 1e-12 * l16[0]
        assert np.allclose(g, g16, rtol=1e-12, atol=1e-12 * np.abs(g16).max())
        assert np.allclose(gs, gs16, rtol=1e-12, atol=1e-12 * np.abs(gs16).max())
    # a box that ignored the pixel test would see more splats with bigger tiles
    assert out[32][3][2] > out[8][3][2]


def test_whole_view_binning_matches_per_tile_binning(oracle):
    """The O(keys) whole-view binning (used to time the CPU path) gives the same
    lists as the per-tile brute force, and the CSR backward the same image."""
    W, H = 120, 90
    s, sh = scenes.make_ext_splats(2000, 12, sh_degree=2)
    vp, vi = camera.view_proj((0.5, 0.1, 2.9), aspect=W / H)
    for tile in (8, 16):
        tiles, lists, counts = oracle.ext_bin(s, sh, 2, vp, vi, W, H, tile=tile)
        offs, ids = oracle.ext_bin_all(s, sh, 2, vp, vi, W, H, tile=tile)
        assert np.array_equal(np.diff(offs), counts.astype(np.int64))
        for t in range(len(tiles)):
            assert np.array_equal(ids[offs[t]:offs[t + 1]], lists[t, :counts[t]])
    tgt = np.full((H, W, 4), 0.5, np.float32)
    a = oracle.ext_backward_f32(s, sh, 2, vp, vi, tgt, W, H, tiles, lists, counts, tile=16)
    offs8, ids8 = oracle.ext_bin_all(s, sh, 2, vp, vi, W, H, tile=8)
    b = oracle.ext_backward_csr_f32(s, sh, 2, vp, vi, tgt, W, H, offs8, ids8, tile=8)
    assert np.array_equal(a[2], b[2])
