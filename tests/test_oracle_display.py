"""Pins of the display-compositor oracle (splat_render.wgsl) by closed forms on
constant textures: the separator and strip borders, the path-trace/raster
choice by alpha, the layer blends, the grey depth panel and the letterbox
background.  Canvas 1000 x 500 with the target sized as the viewer sizes it
(half the width, the height above the 18 % strip: 500 x 410)."""
import numpy as np

W, H = 1000, 500
TW, TH = 500, 410


def _tex(val, w=TW, h=TH):
    return np.broadcast_to(np.asarray(val, np.float32), (h, w, 4)).copy()


def _base(**over):
    t = {"target": _tex([0.1, 0.2, 0.3, 1.0]), "splat": _tex([0.4, 0.5, 0.6, 1.0]),
         "depth": _tex([0.7, 0.0, 0.0, 1.0]), "edge": _tex([0.25, 0.0, 0.0, 1.0]),
         "bez": _tex([0.0, 0.0, 0.0, 0.0]), "bcol": _tex([0.0, 0.0, 0.0, 0.0]),
         "col": _tex([0.0, 0.0, 0.0, 0.0]), "pt": _tex([0.9, 0.8, 0.7, 0.0])}
    t.update(over)
    return [t[k] for k in ("target", "splat", "depth", "edge", "bez", "bcol", "col", "pt")]


def _uv():
    ux = (np.arange(W, dtype=np.float32) + np.float32(0.5)) / np.float32(W)
    uy = (np.arange(H, dtype=np.float32) + np.float32(0.5)) / np.float32(H)
    return ux, uy


def test_main_area_layout(oracle):
    out = oracle.compose_f32(_base(), (0, 0, 0), W, H)
    ux, uy = _uv()
    main = uy <= np.float32(1.0) - np.float32(0.18)
    left = ux < np.float32(0.498)
    sep = (~left) & (ux < np.float32(0.502))
    right = ux >= np.float32(0.502)
    # left: no path-trace sample yet (alpha 0) -> the raster target
    assert np.all(out[np.ix_(main, left)] == np.float32([0.1, 0.2, 0.3, 1.0]))
    assert np.all(out[np.ix_(main, sep)] == 1.0)
    # right: all layers off -> the splat view (alpha 1)
    assert np.all(out[np.ix_(main, right)] == np.float32([0.4, 0.5, 0.6, 1.0]))
    # a path-trace sample (alpha > 0) replaces the raster
    out = oracle.compose_f32(_base(pt=_tex([0.9, 0.8, 0.7, 0.5])), (0, 0, 0), W, H)
    assert np.all(out[np.ix_(main, left)] == np.float32([0.9, 0.8, 0.7, 0.5]))


def test_layer_blends(oracle):
    ux, uy = _uv()
    main = uy <= np.float32(1.0) - np.float32(0.18)
    right = ux >= np.float32(0.502)
    t = _base(bcol=_tex([0.2, 0.2, 0.2, 0.5]), col=_tex([0.0, 0.1, 0.0, 0.25]), bez=_tex([2.0, 0.0, 0.0, 1.0]))
    # base colour only: b (1 - 0.5) + 0.2
    out = oracle.compose_f32(t, (0, 1, 0), W, H)
    exp = np.float32([0.4, 0.5, 0.6]) * np.float32(0.5) + np.float32(0.2)
    assert np.allclose(out[np.ix_(main, right)][..., :3], exp, atol=1e-7)
    # then colour over it
    out = oracle.compose_f32(t, (0, 1, 1), W, H)
    exp2 = exp * np.float32(0.75) + np.float32([0.0, 0.1, 0.0])
    assert np.allclose(out[np.ix_(main, right)][..., :3], exp2, atol=1e-7)
    # edge layer red 2 is clamped to 1: white
    out = oracle.compose_f32(t, (1, 1, 1), W, H)
    assert np.all(out[np.ix_(main, right)][..., :3] == 1.0)
    # edge layer 0.25: mix towards white by a quarter
    t[4] = _tex([0.25, 0.0, 0.0, 1.0])
    out = oracle.compose_f32(t, (1, 0, 0), W, H)
    assert np.allclose(out[np.ix_(main, right)][..., :3], np.float32([0.4, 0.5, 0.6]) * 0.75 + 0.25, atol=1e-7)


def test_strip_panels(oracle):
    ux, uy = _uv()
    strip = uy > np.float32(1.0) - np.float32(0.18)
    sv = (uy - (np.float32(1.0) - np.float32(0.18))) / np.float32(0.18)
    pu = ux * np.float32(6.0) - np.floor(ux * np.float32(6.0))
    pidx = np.floor(ux * np.float32(6.0))
    border_rows = strip & (sv < np.float32(0.01))
    sep_cols = (pu < np.float32(0.005)) | (pu > np.float32(0.995))
    out = oracle.compose_f32(_base(), (0, 0, 0), W, H)
    assert np.all(out[border_rows] == np.float32([0.3, 0.3, 0.3, 1.0]))
    assert np.all(out[np.ix_(strip, sep_cols)] == np.float32([0.3, 0.3, 0.3, 1.0]))
    inner = strip & (sv >= np.float32(0.01))
    cols = ~sep_cols
    # panel aspect = (500/410 * 2 * 0.82) / 1.08 = 1.852 > the target's 1.220: pillarboxed with background
    panel1 = cols & (pidx == 1)   # target depth, grey from .r
    vals = out[np.ix_(inner, panel1)].reshape(-1, 4)
    grey = np.float32([0.7, 0.7, 0.7, 1.0])
    bg = np.float32([0.05, 0.05, 0.05, 1.0])
    assert np.all(np.all(vals == grey, axis=1) | np.all(vals == bg, axis=1))
    assert np.any(np.all(vals == bg, axis=1)) and np.any(np.all(vals == grey, axis=1))
    # the content width is panel width * 1.220 / 1.852 = 65.9 %: the background share matches
    frac_bg = np.mean(np.all(vals == bg, axis=1))
    assert abs(frac_bg - (1 - (500 / 410) / ((500 / 410) * 2 * 0.82 / 1.08))) < 0.02
    panel0 = cols & (pidx == 0)  # splat view, colour
    vals = out[np.ix_(inner, panel0)].reshape(-1, 4)
    assert np.all(np.all(vals == np.float32([0.4, 0.5, 0.6, 1.0]), axis=1) | np.all(vals == bg, axis=1))
