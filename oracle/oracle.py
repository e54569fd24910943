"""ctypes front end of the CPU oracle (oracle/libnps_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libnps_oracle.so")
_lib = None

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_L = C.POINTER(C.c_int64)
i32 = C.c_int32

NO_REG = 1
NO_PULL = 2
NO_SKIP = 4


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        sig = {
            "ora_bin": [_F, i32, _F, i32, i32, _I, i32, i32, _U, _I, _I],
            "ora_backward_f32": [_F, i32, _F, _F, _F, _F, i32, i32, i32, _I, i32, _I, _F, _F, _D, _I],
            "ora_backward_f64": [_D, i32, _D, _D, _F, _F, i32, i32, _I, i32, _U, _I, i32, _D, _D, _D, _D, i32],
            "ora_step_f32": [_F, _I, _F, _F, _F, i32, _F],
            "ora_adc_f32": [_F, _F, _F, _F, i32, C.c_float, _I],
            "ora_render_f32": [_F, i32, _F, i32, i32, _F, _F, i32, i32],
        }
        for k, args in sig.items():
            getattr(lib, k).argtypes = args
            getattr(lib, k).restype = i32
        for k in ["ora_exp2f", "ora_log2f", "ora_expf", "ora_logf", "ora_sinf", "ora_cosf"]:
            getattr(lib, k).argtypes = [C.c_float]
            getattr(lib, k).restype = C.c_float
        lib.ora_powf.argtypes = [C.c_float, C.c_float]
        lib.ora_powf.restype = C.c_float
        lib.ora_f2i.argtypes = [C.c_float]
        lib.ora_f2i.restype = C.c_int32
        lib.ora_num_threads.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def tiles_of(W, H):
    return ((W + 15) // 16) * ((H + 15) // 16)


def bin_tiles(splats, vp, W, H, tiles=None, cap=1024):
    lib = load()
    s = f32(splats)
    n = s.shape[0]
    tiles = np.arange(tiles_of(W, H), dtype=np.int32) if tiles is None else np.ascontiguousarray(tiles, np.int32)
    ids = np.zeros((len(tiles), cap), dtype=np.uint32)
    cnt = np.zeros(len(tiles), dtype=np.int32)
    full = np.zeros(len(tiles), dtype=np.int32)
    lib.ora_bin(_p(s, _F), n, _p(f32(vp), _F), W, H, _p(tiles, _I), len(tiles), cap, _p(ids, _U), _p(cnt, _I),
                _p(full, _I))
    return ids, cnt, full


def backward_f32(splats, vp, vp_inv, tgt_rgba, tgt_depth, W, H, cap=1024, tiles=None, grads=None):
    """REF backward.  Returns (grads int32 [n,16], C [H,W,3], D [H,W], loss[2], full_counts)."""
    lib = load()
    s = f32(splats)
    n = s.shape[0]
    g = np.zeros((n, 16), dtype=np.int32) if grads is None else np.ascontiguousarray(grads, dtype=np.int32)
    c = np.zeros((H, W, 3), dtype=np.float32)
    d = np.zeros((H, W), dtype=np.float32)
    loss = np.zeros(2, dtype=np.float64)
    t = None if tiles is None else np.ascontiguousarray(tiles, np.int32)
    nt = tiles_of(W, H) if t is None else len(t)
    full = np.zeros(nt, dtype=np.int32)
    lib.ora_backward_f32(_p(s, _F), n, _p(f32(vp), _F), _p(f32(vp_inv), _F), _p(f32(tgt_rgba), _F),
                         _p(f32(tgt_depth), _F), W, H, cap, _p(t, _I), 0 if t is None else len(t), _p(g, _I),
                         _p(c, _F), _p(d, _F), _p(loss, _D), _p(full, _I))
    return g, c, d, loss, full


def backward_f64(splats64, vp64, vpinv64, tgt_rgba, tgt_depth, W, H, tiles, lists, counts, want_grads=True,
                 flags=0):
    lib = load()
    s = np.ascontiguousarray(splats64, dtype=np.float64)
    n = s.shape[0]
    g = np.zeros((n, 16), dtype=np.float64) if want_grads else None
    c = np.zeros((H, W, 3), dtype=np.float64)
    d = np.zeros((H, W), dtype=np.float64)
    loss = np.zeros(2, dtype=np.float64)
    tiles = np.ascontiguousarray(tiles, np.int32)
    lists = np.ascontiguousarray(lists, np.uint32)
    counts = np.ascontiguousarray(counts, np.int32)
    lib.ora_backward_f64(_p(s, _D), n, _p(np.ascontiguousarray(vp64, np.float64), _D),
                         _p(np.ascontiguousarray(vpinv64, np.float64), _D), _p(f32(tgt_rgba), _F),
                         _p(f32(tgt_depth), _F), W, H, _p(tiles, _I), len(tiles), _p(lists, _U), _p(counts, _I),
                         lists.shape[1], _p(g, _D), _p(c, _D), _p(d, _D), _p(loss, _D), flags)
    return g, c, d, loss


def step_f32(splats, grads, m, v, accum, t):
    lib = load()
    s, g = f32(splats).copy(), np.ascontiguousarray(grads, np.int32).copy()
    m, v, a = f32(m).copy(), f32(v).copy(), f32(accum).copy()
    tt = np.array([t], dtype=np.float32)
    lib.ora_step_f32(_p(s, _F), _p(g, _I), _p(m, _F), _p(v, _F), _p(a, _F), s.shape[0], _p(tt, _F))
    return s, g, m, v, a, float(tt[0])


def adc_f32(splats, m, v, accum, t):
    lib = load()
    s, m, v, a = f32(splats).copy(), f32(m).copy(), f32(v).copy(), f32(accum).copy()
    st = np.zeros(4, dtype=np.int32)
    lib.ora_adc_f32(_p(s, _F), _p(m, _F), _p(v, _F), _p(a, _F), s.shape[0], float(t), _p(st, _I))
    return s, m, v, a, st


def render_f32(splats, vp, W, H, color=None, depth=None, clear=True, quantize=True):
    lib = load()
    s = f32(splats)
    c = np.zeros((H, W, 4), np.float32) if color is None else f32(color).copy()
    d = np.zeros((H, W, 4), np.float32) if depth is None else f32(depth).copy()
    lib.ora_render_f32(_p(s, _F), s.shape[0], _p(f32(vp), _F), W, H, _p(c, _F), _p(d, _F), int(clear),
                       int(quantize))
    return c, d


def math_f32(op, x, y=None):
    lib = load()
    fn = {0: lib.ora_exp2f, 1: lib.ora_log2f, 3: lib.ora_expf, 4: lib.ora_logf, 5: lib.ora_sinf, 7: lib.ora_cosf}
    x = f32(x)
    if op == 2:
        return np.array([lib.ora_powf(float(a), float(b)) for a, b in zip(x, f32(y))], dtype=np.float32)
    return np.array([fn[op](float(a)) for a in x], dtype=np.float32)


# ---------------------------------------------------------------- EXT path
def _ext_sig(lib):
    lib.ora_ext_bin.argtypes = [_F, _F, i32, i32, _F, _F, i32, i32, i32, _I, i32, i32, _U, _I]
    lib.ora_ext_backward_f32.argtypes = [_F, _F, i32, i32, _F, _F, _F, i32, i32, i32, _I, i32, _U, _I, i32, _D, _D, _F,
                                         _D]
    lib.ora_ext_backward_f64.argtypes = [_D, _D, i32, i32, _D, _D, _F, i32, i32, i32, _I, i32, _U, _I, i32, _D, _D, _D,
                                         _D]
    for k in ("ora_ext_bin", "ora_ext_backward_f32", "ora_ext_backward_f64"):
        getattr(lib, k).restype = i32


def cam_from_inv(vp_inv):
    vi = np.asarray(vp_inv, np.float32).astype(np.float64)
    return np.array([vi[8] / vi[11], vi[9] / vi[11], vi[10] / vi[11]], np.float64)


def ext_bin(splats, sh, deg, vp, vp_inv, W, H, stride=None, tiles=None, tile=16):
    lib = load()
    _ext_sig(lib)
    s, h = f32(splats), f32(sh)
    n = s.shape[0]
    nt = ((W + tile - 1) // tile) * ((H + tile - 1) // tile)
    tiles = np.arange(nt, dtype=np.int32) if tiles is None else np.ascontiguousarray(tiles, np.int32)
    stride = n if stride is None else stride
    lists = np.zeros((len(tiles), max(stride, 1)), np.uint32)
    counts = np.zeros(len(tiles), np.int32)
    lib.ora_ext_bin(_p(s, _F), _p(h, _F), n, deg, _p(f32(vp), _F), _p(f32(vp_inv), _F), W, H, tile, _p(tiles, _I),
                    len(tiles), max(stride, 1), _p(lists, _U), _p(counts, _I))
    return tiles, lists, counts


def ext_backward_f32(splats, sh, deg, vp, vp_inv, tgt, W, H, tiles, lists, counts, want_grads=True, tile=16):
    lib = load()
    _ext_sig(lib)
    s, h = f32(splats), f32(sh)
    n = s.shape[0]
    g = np.zeros((n, 16), np.float64) if want_grads else None
    gs = np.zeros((n, 48), np.float64) if want_grads else None
    c = np.zeros((H, W, 3), np.float32)
    loss = np.zeros(5, np.float64)
    lib.ora_ext_backward_f32(_p(s, _F), _p(h, _F), n, deg, _p(f32(vp), _F), _p(f32(vp_inv), _F), _p(f32(tgt), _F), W,
                             H, tile, _p(tiles, _I), len(tiles), _p(lists, _U), _p(counts, _I), lists.shape[1], _p(g, _D),
                             _p(gs, _D), _p(c, _F), _p(loss, _D))
    return g, gs, c, loss


def ext_backward_f64(splats64, sh64, deg, vp64, cam64, tgt, W, H, tiles, lists, counts, want_grads=True, tile=16):
    lib = load()
    _ext_sig(lib)
    s = np.ascontiguousarray(splats64, np.float64)
    h = np.ascontiguousarray(sh64, np.float64)
    n = s.shape[0]
    g = np.zeros((n, 16), np.float64) if want_grads else None
    gs = np.zeros((n, 48), np.float64) if want_grads else None
    c = np.zeros((H, W, 3), np.float64)
    loss = np.zeros(5, np.float64)
    lib.ora_ext_backward_f64(_p(s, _D), _p(h, _D), n, deg, _p(np.ascontiguousarray(vp64, np.float64), _D),
                             _p(np.ascontiguousarray(cam64, np.float64), _D), _p(f32(tgt), _F), W, H, tile, _p(tiles, _I),
                             len(tiles), _p(lists, _U), _p(counts, _I), lists.shape[1], _p(g, _D), _p(gs, _D),
                             _p(c, _D), _p(loss, _D))
    return g, gs, c, loss


def ext_adam_f32(splats, sh, grads64, m, v, deg, t):
    """EXT Adam (FP32) on copies; returns (splats, sh, m, v, t)."""
    lib = load()
    lib.ora_ext_adam_f32.argtypes = [_F, _F, _F, _F, _F, i32, i32, _F]
    lib.ora_ext_adam_f32.restype = i32
    s, h = f32(splats).copy(), f32(sh).copy()
    mm, vv = f32(m).copy(), f32(v).copy()
    tt = np.array([t], np.float32)
    lib.ora_ext_adam_f32(_p(s, _F), _p(h, _F), _p(f32(grads64), _F), _p(mm, _F), _p(vv, _F), s.shape[0], deg,
                         _p(tt, _F))
    return s, h, mm, vv, float(tt[0])


def blur_f32(src_rgba, W, H, radius=16, sigma=None, is_srgb=True):
    """GpuBlurPipelineManager.blur on the CPU: returns the H x W x 4 output."""
    lib = load()
    lib.ora_blur_f32.argtypes = [_F, _F, _F, i32, i32, i32, C.c_float, i32]
    lib.ora_blur_f32.restype = i32
    sigma = radius / 2 if sigma is None else sigma
    src = f32(src_rgba)
    dst = np.zeros((H, W, 4), np.float32)
    tmp = np.zeros((H, W, 4), np.float32)
    rc = lib.ora_blur_f32(_p(src, _F), _p(dst, _F), _p(tmp, _F), W, H, int(radius), float(sigma), 1 if is_srgb else 0)
    assert rc == 0
    return dst


def ext_adc_f32(splats, sh, m, v, accum, threshold=2e-4, extent=3.3, prune_opacity=0.005, percent_dense=0.01,
                seed=0):
    """EXT densification on copies; returns (splats, sh, m, v, accum, stats[5])."""
    lib = load()
    lib.ora_ext_adc_f32.argtypes = [_F, _F, _F, _F, _F, i32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32,
                                    _L]
    lib.ora_ext_adc_f32.restype = i32
    s, h = f32(splats).copy(), f32(sh).copy()
    mm, vv, acc = f32(m).copy(), f32(v).copy(), f32(accum).copy()
    st = np.zeros(5, np.int64)
    lib.ora_ext_adc_f32(_p(s, _F), _p(h, _F), _p(mm, _F), _p(vv, _F), _p(acc, _F), s.shape[0], threshold, extent,
                        prune_opacity, percent_dense, seed, _p(st, _L))
    return s, h, mm, vv, acc, st


def ext_adc_normal(seed, slot, child, k):
    lib = load()
    lib.ora_ext_adc_normal.argtypes = [C.c_uint32] * 4
    lib.ora_ext_adc_normal.restype = C.c_float
    return float(lib.ora_ext_adc_normal(seed, slot, child, k))


def ext_bin_all(splats, sh, deg, vp, vp_inv, W, H, tile=16):
    """Whole-view tile lists in O(keys), CSR: (offs [T+1] int64, ids [M] uint32)."""
    lib = load()
    lib.ora_ext_bin_all.argtypes = [_F, _F, i32, i32, _F, _F, i32, i32, i32, _L, _U]
    lib.ora_ext_bin_all.restype = C.c_int64
    s, h = f32(splats), f32(sh)
    T = ((W + tile - 1) // tile) * ((H + tile - 1) // tile)
    offs = np.zeros(T + 1, np.int64)
    m = lib.ora_ext_bin_all(_p(s, _F), _p(h, _F), s.shape[0], deg, _p(f32(vp), _F), _p(f32(vp_inv), _F), W, H, tile,
                            _p(offs, _L), None)
    ids = np.zeros(max(int(m), 1), np.uint32)
    lib.ora_ext_bin_all(_p(s, _F), _p(h, _F), s.shape[0], deg, _p(f32(vp), _F), _p(f32(vp_inv), _F), W, H, tile,
                        _p(offs, _L), _p(ids, _U))
    return offs, ids[:m]


def ext_backward_csr_f32(splats, sh, deg, vp, vp_inv, tgt, W, H, offs, ids, want_grads=True, tile=16):
    T = ((W + tile - 1) // tile) * ((H + tile - 1) // tile)
    if len(offs) != T + 1:
        raise ValueError(f"offs has {len(offs)} entries, {T + 1} expected for {tile}-pixel tiles")
    lib = load()
    lib.ora_ext_backward_csr_f32.argtypes = [_F, _F, i32, i32, _F, _F, _F, i32, i32, i32, _U, _L, _D, _D, _F, _D]
    lib.ora_ext_backward_csr_f32.restype = i32
    s, h = f32(splats), f32(sh)
    n = s.shape[0]
    g = np.zeros((n, 16), np.float64) if want_grads else None
    gs = np.zeros((n, 48), np.float64) if want_grads else None
    c = np.zeros((H, W, 3), np.float32)
    loss = np.zeros(5, np.float64)
    idsc = np.ascontiguousarray(ids, np.uint32) if len(ids) else np.zeros(1, np.uint32)
    lib.ora_ext_backward_csr_f32(_p(s, _F), _p(h, _F), n, deg, _p(f32(vp), _F), _p(f32(vp_inv), _F), _p(f32(tgt), _F),
                                 W, H, tile, _p(idsc, _U), _p(np.ascontiguousarray(offs, np.int64), _L), _p(g, _D),
                                 _p(gs, _D), _p(c, _F), _p(loss, _D))
    return g, gs, c, loss


def ext_backward_csr_f64(splats, sh, deg, vp, vp_inv, tgt, W, H, offs, ids, want_grads=True, tile=16):
    """FP64 twin of ext_backward_csr_f32 on the same CSR lists (splats, SH and VP
    widened from FP32, camera centre from cam_from_inv)."""
    T = ((W + tile - 1) // tile) * ((H + tile - 1) // tile)
    if len(offs) != T + 1:
        raise ValueError(f"offs has {len(offs)} entries, {T + 1} expected for {tile}-pixel tiles")
    lib = load()
    lib.ora_ext_backward_csr_f64.argtypes = [_D, _D, i32, i32, _D, _D, _F, i32, i32, i32, _U, _L, _D, _D, _D, _D]
    lib.ora_ext_backward_csr_f64.restype = i32
    s = np.ascontiguousarray(np.asarray(splats, np.float32), np.float64)
    h = np.ascontiguousarray(np.asarray(sh, np.float32), np.float64)
    n = s.shape[0]
    g = np.zeros((n, 16), np.float64) if want_grads else None
    gs = np.zeros((n, 48), np.float64) if want_grads else None
    c = np.zeros((H, W, 3), np.float64)
    loss = np.zeros(5, np.float64)
    idsc = np.ascontiguousarray(ids, np.uint32) if len(ids) else np.zeros(1, np.uint32)
    vp64 = np.ascontiguousarray(np.asarray(vp, np.float32).reshape(16), np.float64)
    lib.ora_ext_backward_csr_f64(_p(s, _D), _p(h, _D), n, deg, _p(vp64, _D), _p(cam_from_inv(vp_inv), _D),
                                 _p(f32(tgt), _F), W, H, tile, _p(idsc, _U), _p(np.ascontiguousarray(offs, np.int64), _L),
                                 _p(g, _D), _p(gs, _D), _p(c, _D), _p(loss, _D))
    return g, gs, c, loss


def compose_f32(textures, enabled, W, H):
    """The debug display compositor (splat_render.wgsl) on the CPU: textures = 8
    H_k x W_k x 4 arrays in binding order, enabled = (edge, base colour, colour)
    layer flags; returns the H x W x 4 canvas."""
    lib = load()
    lib.ora_compose_f32.argtypes = [C.POINTER(_F), _I, _F, _F, i32, i32]
    lib.ora_compose_f32.restype = i32
    arrs = [f32(t) for t in textures]
    ptrs = (_F * 8)(*[_p(a, _F) for a in arrs])
    dims = np.array([[a.shape[1], a.shape[0]] for a in arrs], np.int32).reshape(16)
    en = np.asarray(enabled, np.float32)
    out = np.zeros((H, W, 4), np.float32)
    lib.ora_compose_f32(ptrs, _p(dims, _I), _p(en, _F), _p(out, _F), W, H)
    return out


# ---------------------------------------------------------------- Bezier path
BZ_CAP = 1024


def bz_uniforms(vp, vp_inv, mode=1.0, max_width=0.0, prune_alpha=0.0, prune_width=0.0, bg_penalty=0.0):
    """BezierUniforms (bezier_backward.wgsl:20-32) as the oracle's flat struct."""
    u = np.zeros(37, np.float32)
    u[:16] = f32(vp).reshape(16)
    u[16:21] = [mode, max_width, prune_alpha, prune_width, bg_penalty]
    u[21:37] = f32(vp_inv).reshape(16)
    return u


def _tex4(a, H, W):
    a = f32(a)
    if a.ndim == 2:
        out = np.zeros((H, W, 4), np.float32)
        out[..., 0] = a
        return out
    return np.ascontiguousarray(a.reshape(H, W, 4))


def bz_bin(bz, uni, W, H, tiles=None):
    """Tile lists (first 1024 ascending ids).  Returns (lists [T,1024], counts, full counts)."""
    lib = load()
    lib.ora_bz_bin.argtypes = [_F, i32, _F, i32, i32, _I, i32, _U, _I, _I]
    lib.ora_bz_bin.restype = i32
    b = f32(bz)
    tiles = np.arange(tiles_of(W, H), dtype=np.int32) if tiles is None else np.ascontiguousarray(tiles, np.int32)
    lists = np.zeros((len(tiles), BZ_CAP), np.uint32)
    cnt = np.zeros(len(tiles), np.int32)
    full = np.zeros(len(tiles), np.int32)
    rc = lib.ora_bz_bin(_p(b, _F), b.shape[0], _p(f32(uni), _F), W, H, _p(tiles, _I), len(tiles), _p(lists, _U),
                        _p(cnt, _I), _p(full, _I))
    assert rc == 0
    return lists, cnt, full


def bz_backward_f32(bz, uni, tgt, tgt_depth, bg, bg_depth, normal, W, H, tiles=None, lists=None, counts=None,
                    grads=None, loss_acc=None, pixel_loss=None):
    """bezier_backward over the given tiles (all by default).  Accumulates into copies
    of grads [n,18] i32, loss_acc [n] i64 (2^-32 fixed point) and pixel_loss [H*W] i32.
    Returns a dict with grads, loss_acc, pixel_loss, C [H,W,3], D [H,W], T [H,W], loss[2]."""
    lib = load()
    lib.ora_bz_backward_f32.argtypes = [_F, i32, _F, _F, _F, _F, _F, _F, i32, i32, _I, i32, _U, _I, _I, _L, _I, _F,
                                        _F, _F, _D]
    lib.ora_bz_backward_f32.restype = i32
    b = f32(bz)
    n = b.shape[0]
    tiles = np.arange(tiles_of(W, H), dtype=np.int32) if tiles is None else np.ascontiguousarray(tiles, np.int32)
    if lists is None:
        lists, counts, _ = bz_bin(b, uni, W, H, tiles)
    g = np.zeros((n, 18), np.int32) if grads is None else np.ascontiguousarray(grads, np.int32).copy()
    la = np.zeros(n, np.int64) if loss_acc is None else np.ascontiguousarray(loss_acc, np.int64).copy()
    pl = np.zeros(H * W, np.int32) if pixel_loss is None else np.ascontiguousarray(pixel_loss, np.int32).copy()
    c = np.zeros((H, W, 3), np.float32)
    d = np.zeros((H, W), np.float32)
    t = np.zeros((H, W), np.float32)
    loss = np.zeros(2, np.float64)
    texs = [_tex4(x, H, W) for x in (tgt, tgt_depth, bg, bg_depth, normal)]
    rc = lib.ora_bz_backward_f32(_p(b, _F), n, _p(f32(uni), _F), *[_p(x, _F) for x in texs], W, H, _p(tiles, _I),
                                 len(tiles), _p(np.ascontiguousarray(lists, np.uint32), _U),
                                 _p(np.ascontiguousarray(counts, np.int32), _I), _p(g, _I), _p(la, _L), _p(pl, _I),
                                 _p(c, _F), _p(d, _F), _p(t, _F), _p(loss, _D))
    assert rc == 0
    return dict(grads=g, loss_acc=la, pixel_loss=pl, C=c, D=d, T=t, loss=loss)


def bz_step_f32(bz, grads, m, v, grad_accum, uni, t, pixel_count, no_kill=0.0):
    """bezier_step on copies; returns (bz, grads (drained), m, v, grad_accum)."""
    lib = load()
    lib.ora_bz_step_f32.argtypes = [_F, i32, _I, _F, _F, _F, _F, C.c_float, C.c_float, C.c_float]
    lib.ora_bz_step_f32.restype = i32
    b = f32(bz).copy()
    g = np.ascontiguousarray(grads, np.int32).copy()
    mm, vv, ga = f32(m).copy(), f32(v).copy(), f32(grad_accum).copy()
    rc = lib.ora_bz_step_f32(_p(b, _F), b.shape[0], _p(g, _I), _p(mm, _F), _p(vv, _F), _p(ga, _F), _p(f32(uni), _F),
                             float(t), float(pixel_count), float(no_kill))
    assert rc == 0
    return b, g, mm, vv, ga


def bz_adc_f32(bz, m, v, grad_accum, loss_acc, pixel_loss, W, H, uni, adam_t, no_kill=0.0):
    """bezier_adc on copies; returns (bz, m, v, grad_accum, loss_acc, pixel_loss, stats[6])."""
    lib = load()
    lib.ora_bz_adc_f32.argtypes = [_F, i32, _F, _F, _F, _L, _I, i32, i32, _F, C.c_float, C.c_float, _L]
    lib.ora_bz_adc_f32.restype = i32
    b = f32(bz).copy()
    mm, vv, ga = f32(m).copy(), f32(v).copy(), f32(grad_accum).copy()
    la = np.ascontiguousarray(loss_acc, np.int64).copy()
    pl = np.ascontiguousarray(pixel_loss, np.int32).copy()
    st = np.zeros(6, np.int64)
    rc = lib.ora_bz_adc_f32(_p(b, _F), b.shape[0], _p(mm, _F), _p(vv, _F), _p(ga, _F), _p(la, _L), _p(pl, _I), W, H,
                            _p(f32(uni), _F), float(adam_t), float(no_kill), _p(st, _L))
    assert rc == 0
    return b, mm, vv, ga, la, pl, st


def bz_brush_values(brush_rgba8):
    """Texel values the sampler returns for an H x W x 4 uint8 brush: red / 255 (float32)."""
    b = np.ascontiguousarray(brush_rgba8, np.uint8)
    return np.ascontiguousarray(b[..., 0].astype(np.float32) / np.float32(255.0))


def bz_brush_sample(brush_rgba8, u, v):
    """The restated brush sampler at (u, v) (scalars)."""
    lib = load()
    lib.ora_bz_brush_sample.argtypes = [_F, i32, i32, C.c_float, C.c_float]
    lib.ora_bz_brush_sample.restype = C.c_float
    t = bz_brush_values(brush_rgba8)
    return float(lib.ora_bz_brush_sample(_p(t, _F), t.shape[1], t.shape[0], float(u), float(v)))


def bz_render_f32(bz, vp, W, H, brush_rgba8, rgba=None, clear=True, quantize=True):
    """Display raster of the curves (bezier_forward.wgsl); returns H x W x 4 float32."""
    lib = load()
    lib.ora_bz_render_f32.argtypes = [_F, i32, _F, i32, i32, _F, i32, i32, _F, i32, i32]
    lib.ora_bz_render_f32.restype = i32
    b = f32(bz).reshape(-1, 20)
    t = bz_brush_values(brush_rgba8)
    out = np.zeros((H, W, 4), np.float32) if rgba is None else f32(rgba).reshape(H, W, 4).copy()
    rc = lib.ora_bz_render_f32(_p(b, _F), b.shape[0], _p(f32(vp), _F), W, H, _p(t, _F), t.shape[1], t.shape[0],
                               _p(out, _F), int(clear), int(quantize))
    if rc != 0:
        raise RuntimeError(f"ora_bz_render_f32 failed ({rc})")
    return out


# ---------------------------------------------------------------- path trace
def _pt_sig(lib):
    if getattr(lib, "_pt_ready", False):
        return
    lib.ora_pt_bvh.argtypes = [_F, i32, _U, i32, _U, _U]
    lib.ora_pt_bvh.restype = i32
    lib.ora_pt_trace.argtypes = [_F, C.c_uint32, i32, i32, _F, _U, i32, _U, _F, i32, i32, _F]
    lib.ora_pt_trace.restype = None
    lib.ora_pt_resolve.argtypes = [_F, i32, _F]
    lib.ora_pt_resolve.restype = None
    lib.ora_env_background.argtypes = [_F, _F, i32, i32, i32, i32, _F]
    lib.ora_env_background.restype = None
    lib.ora_pt_closest_brute.argtypes = [_F, _F, _F, _U, i32, _F, _I]
    lib.ora_pt_closest_brute.restype = None
    lib.ora_pt_closest_bvh.argtypes = [_F, _F, _F, _U, i32, _U, _F]
    lib.ora_pt_closest_bvh.restype = None
    lib.ora_pt_hemisphere.argtypes = [_F, C.c_uint32, i32, _F]
    lib.ora_pt_hemisphere.restype = None
    lib.ora_pt_pcg.argtypes = [C.c_uint32]
    lib.ora_pt_pcg.restype = C.c_uint32
    lib.ora_atan2_asin.argtypes = [_F, _F, i32, _F, _F]
    lib.ora_atan2_asin.restype = None
    lib._pt_ready = True


def pt_bvh(verts, indices):
    """SAH BVH of bvh.ts: (nodes u32[n, 8], tris u32[m, 3])."""
    lib = load()
    _pt_sig(lib)
    v = f32(verts).reshape(-1, 12)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    nt = idx.size // 3
    nodes = np.zeros((2 * nt + 1, 8), np.uint32)
    tris = np.zeros((max(nt, 1), 3), np.uint32)
    nn = lib.ora_pt_bvh(_p(v, _F), v.shape[0], _p(idx, _U), nt, _p(nodes, _U), _p(tris, _U))
    return nodes[:nn].copy(), tris[:nt].copy()


def pt_trace(inv_vp, frame, W, H, verts, nodes, tris, env, accum=None):
    """One frame of path_trace.wgsl: returns accum (H, W, 4) += radiance, 1."""
    lib = load()
    _pt_sig(lib)
    acc = np.zeros((H, W, 4), np.float32) if accum is None else f32(accum).copy()
    v = f32(verts).reshape(-1, 12)
    nd = np.ascontiguousarray(nodes, dtype=np.uint32)
    tr = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
    if tr.size == 0:
        tr = np.zeros(3, np.uint32)
    e = f32(env)
    eh, ew = e.shape[0], e.shape[1]
    m = f32(inv_vp).reshape(16)
    lib.ora_pt_trace(_p(m, _F), C.c_uint32(frame), W, H, _p(v, _F), _p(nd, _U), nd.shape[0], _p(tr, _U), _p(e, _F),
                     ew, eh, _p(acc, _F))
    return acc


def pt_resolve(accum):
    lib = load()
    _pt_sig(lib)
    a = f32(accum)
    out = np.empty_like(a)
    lib.ora_pt_resolve(_p(a, _F), a.shape[0] * a.shape[1], _p(out, _F))
    return out


def env_background(inv_vp, env, W, H):
    lib = load()
    _pt_sig(lib)
    e = f32(env)
    out = np.empty((H, W, 4), np.float32)
    lib.ora_env_background(_p(f32(inv_vp).reshape(16), _F), _p(e, _F), e.shape[1], e.shape[0], W, H, _p(out, _F))
    return out


def pt_closest_brute(ro, rd, verts, indices):
    lib = load()
    _pt_sig(lib)
    t = np.zeros(1, np.float32)
    k = np.zeros(1, np.int32)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    lib.ora_pt_closest_brute(_p(f32(ro), _F), _p(f32(rd), _F), _p(f32(verts), _F), _p(idx, _U), idx.size // 3,
                             _p(t, _F), _p(k, _I))
    return float(t[0]), int(k[0])


def pt_closest_bvh(ro, rd, verts, nodes, tris):
    lib = load()
    _pt_sig(lib)
    out = np.zeros(8, np.float32)
    nd = np.ascontiguousarray(nodes, dtype=np.uint32)
    tr = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
    lib.ora_pt_closest_bvh(_p(f32(ro), _F), _p(f32(rd), _F), _p(f32(verts), _F), _p(nd, _U), nd.shape[0],
                           _p(tr, _U), _p(out, _F))
    return out


def pt_hemisphere(n, seed, count):
    lib = load()
    _pt_sig(lib)
    d = np.zeros((count, 3), np.float32)
    lib.ora_pt_hemisphere(_p(f32(n), _F), C.c_uint32(seed), count, _p(d, _F))
    return d


def pt_pcg(v):
    lib = load()
    _pt_sig(lib)
    return int(lib.ora_pt_pcg(C.c_uint32(v)))


def atan2_asin_f32(y, x):
    lib = load()
    _pt_sig(lib)
    y = f32(y)
    x = f32(x)
    a = np.empty_like(y)
    s = np.empty_like(y)
    lib.ora_atan2_asin(_p(y, _F), _p(x, _F), y.size, _p(a, _F), _p(s, _F))
    return a, s


def mesh_raster(vp, view, inv_vp, shading, W, H, verts, nodes, tris, env):
    """mesh.wgsl over the envmap background per pixel centre: (colour, depth),
    each (H, W, 4) rgba8 values."""
    lib = load()
    _pt_sig(lib)
    if not getattr(lib, "_raster_ready", False):
        lib.ora_mesh_raster.argtypes = [_F, _F, _F, C.c_float, i32, i32, _F, _U, i32, _U, _F, i32, i32, _F, _F]
        lib.ora_mesh_raster.restype = None
        lib._raster_ready = True
    v = f32(verts).reshape(-1, 12)
    nd = np.ascontiguousarray(nodes, dtype=np.uint32)
    tr = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
    if tr.size == 0:
        tr = np.zeros(3, np.uint32)
    e = f32(env)
    col = np.empty((H, W, 4), np.float32)
    dep = np.empty((H, W, 4), np.float32)
    lib.ora_mesh_raster(_p(f32(vp).reshape(16), _F), _p(f32(view).reshape(16), _F), _p(f32(inv_vp).reshape(16), _F),
                        float(shading), W, H, _p(v, _F), _p(nd, _U), nd.shape[0], _p(tr, _U), _p(e, _F), e.shape[1],
                        e.shape[0], _p(col, _F), _p(dep, _F))
    return col, dep
