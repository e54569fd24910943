"""Host classes mirroring the reference's WebGPU pipeline managers.

``GpuSplatOptimizerManager`` and ``GpuSplatForwardPipelineManager`` keep the
reference's class and method names (src/lib/gpu/GpuSplatOptimizerManager.ts,
src/lib/gpu/GpuSplatForwardPipelineManager.ts) and drive libnprush_hip.so
through its C ABI.  WebGPU objects become host arrays: a texture view passed to
``setBackwardTarget`` is an H x W x 4 float32 array (colour) plus an H x W
float32 array (depth .r); ``device``/``format`` arguments are accepted and
ignored.  Differences from the reference:

* errors raise ``NpsError`` instead of WebGPU's asynchronous error scopes;
* ``dispatch`` before ``setBackwardTarget`` raises (the reference returns
  silently, GpuSplatOptimizerManager.ts:366);
* splats are written explicitly (``write_splats``) instead of the reference's
  unseeded Math.random() initialisation (GpuSplatOptimizerManager.ts:59-86);
  ``reference_init`` reproduces that distribution with a seed.

``mode="ext"`` selects the EXT model of the same loop (3D Gaussians with SH
colour, depth-ordered uncapped tiles, early termination, L1 loss, FP32
gradients; no reference counterpart): splats use the EXT record, SH
coefficients live in their own buffer (``write_sh``), gradients are FP32
(``read_grads_f32``, 64 per splat), Adam state has 64 entries per splat and
``forward()`` renders without gradients.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr, fptr, iptr, lptr, uptr


def reference_init(num_splats: int, seed: int = 0, live: int = 512) -> np.ndarray:
    """GpuSplatOptimizerManager.ts:59-86 with a seeded generator: positions
    U(-0.3, 0.3), scales 0.1 + U(0, 0.15), RGB U(0,1), opacity 0.3 + U(0, 0.4)
    for the first `live` splats and 0 (dead) after, identity quaternion,
    shape (2, 0.5)."""
    rng = np.random.default_rng(seed)
    s = np.zeros((num_splats, 16), dtype=np.float32)
    for i in range(num_splats):
        r = rng.random(11)
        s[i, 0:3] = (r[0:3] * 2 - 1) * 0.3
        s[i, 3] = 0.1 + r[3] * 0.15
        s[i, 4:7] = r[4:7]
        s[i, 7] = 0.3 + r[7] * 0.4 if i < live else 0.0
        s[i, 8:12] = (1.0, 0.0, 0.0, 0.0)
        s[i, 12] = 0.1 + r[8] * 0.15
        s[i, 13] = 2.0
        s[i, 14] = 0.5
        s[i, 15] = 0.1 + r[9] * 0.15
    return s


class GpuSplatOptimizerManager:
    """new GpuSplatOptimizerManager({device, format, numSplats}) (OPT:43-125)."""

    def __init__(self, device=None, format=None, numSplats: int = 512, numBeziers=None, *,
                 max_tile_splats: int = 1024, adc_period: int = 25, adc_enabled: bool = True, hip_device: int = 0,
                 splats: np.ndarray | None = None, mode: str = "ref", sh_degree: int = 0,
                 sh: np.ndarray | None = None, ext_tile: int = 16, ext_batch: int = 64):
        lib = _lib.load()
        cfg = _lib.NpsConfig()
        check(lib.nps_config_default(C.byref(cfg)), None, "nps_config_default")
        cfg.max_tile_splats = max_tile_splats
        cfg.adc_period = adc_period
        cfg.adc_enabled = 1 if adc_enabled else 0
        cfg.device = hip_device
        if mode not in ("ref", "ext"):
            raise ValueError("mode must be 'ref' or 'ext'")
        cfg.mode = _lib.MODE_EXT if mode == "ext" else _lib.MODE_REF
        cfg.sh_degree = int(sh_degree)
        cfg.ext_tile = int(ext_tile)
        cfg.ext_batch = int(ext_batch)
        h = C.c_void_p()
        check(lib.nps_create(C.byref(cfg), int(numSplats), C.byref(h)), None, "nps_create")
        self._lib = lib
        self._h = h
        self.numSplats = int(numSplats)
        self.mode = mode
        self.sh_degree = int(sh_degree)
        self.opt_width = 64 if mode == "ext" else 16  # Adam entries per splat
        self.numParams = self.numSplats * self.opt_width
        self.dims = (0, 0)
        if splats is not None:
            self.write_splats(splats)
        if sh is not None:
            self.write_sh(sh)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nps_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- buffers (splatBuffer, gradBuffer, adamBuffer, adcBuffer; OPT:15-18) ----
    def write_splats(self, splats: np.ndarray):
        a = np.ascontiguousarray(splats, dtype=np.float32).reshape(self.numSplats, 16)
        check(self._lib.nps_write_splats(self._h, fptr(a), self.numSplats), self._h, "write_splats")

    def read_splats(self) -> np.ndarray:
        a = np.empty((self.numSplats, 16), dtype=np.float32)
        check(self._lib.nps_read_splats(self._h, fptr(a), self.numSplats), self._h, "read_splats")
        return a

    def write_grads(self, grads: np.ndarray):
        a = np.ascontiguousarray(grads, dtype=np.int32).reshape(self.numSplats, 16)
        check(self._lib.nps_write_grads(self._h, iptr(a), self.numSplats), self._h, "write_grads")

    def read_grads(self) -> np.ndarray:
        a = np.empty((self.numSplats, 16), dtype=np.int32)
        check(self._lib.nps_read_grads(self._h, iptr(a), self.numSplats), self._h, "read_grads")
        return a

    def write_adam(self, m: np.ndarray, v: np.ndarray, t: float):
        m = np.ascontiguousarray(m, dtype=np.float32).reshape(self.numSplats, self.opt_width)
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(self.numSplats, self.opt_width)
        check(self._lib.nps_write_adam(self._h, fptr(m), fptr(v), float(t), self.numSplats), self._h, "write_adam")

    def read_adam(self):
        m = np.empty((self.numSplats, self.opt_width), dtype=np.float32)
        v = np.empty((self.numSplats, self.opt_width), dtype=np.float32)
        t = C.c_float()
        check(self._lib.nps_read_adam(self._h, fptr(m), fptr(v), C.byref(t), self.numSplats), self._h, "read_adam")
        return m, v, float(t.value)

    def _accum_shape(self):
        return (self.numSplats, 2) if self.mode == "ext" else (self.numSplats,)

    def write_accum(self, accum: np.ndarray):
        a = np.ascontiguousarray(accum, dtype=np.float32).reshape(self._accum_shape())
        check(self._lib.nps_write_accum(self._h, fptr(a), self.numSplats), self._h, "write_accum")

    def read_accum(self) -> np.ndarray:
        a = np.empty(self._accum_shape(), dtype=np.float32)
        check(self._lib.nps_read_accum(self._h, fptr(a), self.numSplats), self._h, "read_accum")
        return a

    # ---- EXT buffers ----
    def write_sh(self, sh: np.ndarray):
        a = np.ascontiguousarray(sh, dtype=np.float32).reshape(self.numSplats, 48)
        check(self._lib.nps_write_sh(self._h, fptr(a), self.numSplats), self._h, "write_sh")

    def read_sh(self) -> np.ndarray:
        a = np.empty((self.numSplats, 48), dtype=np.float32)
        check(self._lib.nps_read_sh(self._h, fptr(a), self.numSplats), self._h, "read_sh")
        return a

    def write_grads_f32(self, grads: np.ndarray):
        a = np.ascontiguousarray(grads, dtype=np.float32).reshape(self.numSplats, 64)
        check(self._lib.nps_write_grads_f32(self._h, fptr(a), self.numSplats), self._h, "write_grads_f32")

    def read_grads_f32(self) -> np.ndarray:
        a = np.empty((self.numSplats, 64), dtype=np.float32)
        check(self._lib.nps_read_grads_f32(self._h, fptr(a), self.numSplats), self._h, "read_grads_f32")
        return a

    # ---- reference methods ----
    def writeSplatVPMatrix(self, mat, invMat, blurEnabled: bool = False):
        vp = np.ascontiguousarray(mat, dtype=np.float32).reshape(16)
        vi = np.ascontiguousarray(invMat, dtype=np.float32).reshape(16)
        check(self._lib.nps_set_camera(self._h, fptr(vp), fptr(vi), 1 if blurEnabled else 0), self._h,
              "writeSplatVPMatrix")

    def setBackwardTarget(self, targetTextureView: np.ndarray, targetDepthTextureView: np.ndarray, width: int,
                          height: int):
        rgba = np.ascontiguousarray(targetTextureView, dtype=np.float32).reshape(height, width, 4)
        depth = np.ascontiguousarray(targetDepthTextureView, dtype=np.float32).reshape(height, width)
        check(self._lib.nps_set_target(self._h, fptr(rgba), fptr(depth), int(width), int(height)), self._h,
              "setBackwardTarget")
        self.dims = (int(width), int(height))

    def dispatch(self, commandEncoder=None):
        check(self._lib.nps_dispatch(self._h), self._h, "dispatch")

    # ---- targets resident in HBM (no reference counterpart) ----
    def upload_target(self, rgba: np.ndarray, depth: np.ndarray):
        """Copies a target to device memory once; returns a slot for set_target_slot."""
        rgba = np.ascontiguousarray(rgba, dtype=np.float32)
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        slot = []
        for a in (rgba, depth):
            p = C.c_void_p()
            check(self._lib.nps_dev_alloc(self._h, a.nbytes, C.byref(p)), self._h, "nps_dev_alloc")
            check(self._lib.nps_dev_upload(self._h, p, a.ctypes.data_as(C.c_void_p), a.nbytes), self._h,
                  "nps_dev_upload")
            slot.append(p)
        return (slot[0], slot[1], rgba.shape[1], rgba.shape[0])

    def set_target_slot(self, slot):
        rgba_p, depth_p, w, h = slot
        check(self._lib.nps_set_target_dev(self._h, rgba_p, depth_p, int(w), int(h)), self._h, "set_target_dev")
        self.dims = (int(w), int(h))

    def free_target(self, slot):
        for p in slot[:2]:
            self._lib.nps_dev_free(self._h, p)

    # ---- the three passes dispatch() records, callable one at a time ----
    def backward(self):
        check(self._lib.nps_backward(self._h), self._h, "backward")

    def autotune(self):
        """EXT: picks the fastest composite tile size / LDS batch for the current
        camera and target; returns (tile, batch, ms[3, 3])."""
        t, b = C.c_int32(), C.c_int32()
        ms = np.zeros(9, dtype=np.float32)
        check(self._lib.nps_autotune(self._h, C.byref(t), C.byref(b), fptr(ms)), self._h, "autotune")
        return int(t.value), int(b.value), ms.reshape(3, 3)

    def forward(self):
        """EXT: render the current camera at the target's size (no gradients)."""
        check(self._lib.nps_forward(self._h), self._h, "forward")

    def step(self):
        check(self._lib.nps_step(self._h), self._h, "step")

    def adc(self):
        check(self._lib.nps_adc(self._h), self._h, "adc")

    # ---- learned densification support (EXT; no reference counterpart) ----
    def adc_features(self) -> np.ndarray:
        """32 features per parent the next adc() densifies, in its placement order."""
        nq = C.c_int32()
        check(self._lib.nps_adc_features(self._h, None, 0, C.byref(nq)), self._h, "adc_features")
        f = np.empty((max(nq.value, 1), 32), np.float32)
        check(self._lib.nps_adc_features(self._h, fptr(f), nq.value, C.byref(nq)), self._h, "adc_features")
        return f[:nq.value]

    def adc_children(self):
        """(parent slots, dead slots) of the last adc()'s placements."""
        cnt = C.c_int32()
        check(self._lib.nps_adc_children(self._h, None, None, 0, C.byref(cnt)), self._h, "adc_children")
        p = np.empty(max(cnt.value, 1), np.int32)
        c = np.empty(max(cnt.value, 1), np.int32)
        check(self._lib.nps_adc_children(self._h, iptr(p), iptr(c), cnt.value, C.byref(cnt)), self._h,
              "adc_children")
        return p[:cnt.value], c[:cnt.value]

    def adc_apply_deltas(self, deltas: np.ndarray):
        d = np.ascontiguousarray(deltas, dtype=np.float32).reshape(-1, 14)
        check(self._lib.nps_adc_apply_deltas(self._h, fptr(d), d.shape[0]), self._h, "adc_apply_deltas")

    def adc_features_dev(self):
        """(device pointer, rows) of the features adc_features() would return."""
        p = C.c_void_p()
        nq = C.c_int32()
        check(self._lib.nps_adc_features_dev(self._h, C.byref(p), C.byref(nq)), self._h, "adc_features_dev")
        return p.value, nq.value

    def adc_apply_deltas_dev(self, ptr: int, rows: int):
        check(self._lib.nps_adc_apply_deltas_dev(self._h, C.c_void_p(ptr), int(rows)), self._h,
              "adc_apply_deltas_dev")

    def dev_alloc(self, nbytes: int) -> int:
        """A device buffer on the handle's GPU (nps_dev_alloc); free with dev_free."""
        p = C.c_void_p()
        check(self._lib.nps_dev_alloc(self._h, int(nbytes), C.byref(p)), self._h, "nps_dev_alloc")
        return p.value

    def dev_free(self, ptr: int):
        self._lib.nps_dev_free(self._h, C.c_void_p(ptr))

    def dev_upload(self, ptr: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        check(self._lib.nps_dev_upload(self._h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes), self._h,
              "nps_dev_upload")

    def snapshot(self):
        check(self._lib.nps_snapshot(self._h, 0), self._h, "snapshot")

    def restore(self):
        check(self._lib.nps_snapshot(self._h, 1), self._h, "restore")

    def set_adc_ext(self, grad_threshold: float = 2e-4, scene_extent: float = 1.0, prune_opacity: float = 0.005,
                    percent_dense: float = 0.01):
        """EXT densification parameters (see nps_set_adc_ext)."""
        check(self._lib.nps_set_adc_ext(self._h, grad_threshold, scene_extent, prune_opacity, percent_dense),
              self._h, "set_adc_ext")

    def sync(self):
        check(self._lib.nps_sync(self._h), self._h, "sync")

    # ---- outputs the reference never materialises ----
    def read_image(self):
        w, h = self.dims
        c = np.empty((h, w, 3), dtype=np.float32)
        d = np.empty((h, w), dtype=np.float32)
        check(self._lib.nps_read_image(self._h, fptr(c), fptr(d)), self._h, "read_image")
        return c, d

    def read_loss(self):
        out = np.zeros(2, dtype=np.float64)
        check(self._lib.nps_read_loss(self._h, dptr(out)), self._h, "read_loss")
        return float(out[0]), float(out[1])

    def read_tiles(self):
        """(ranges [T,2], ids [M], full_counts [T]) of the last backward."""
        nt = C.c_int32()
        nid = C.c_int64()
        check(self._lib.nps_read_tiles(self._h, None, None, None, C.byref(nt), C.byref(nid)), self._h, "read_tiles")
        ranges = np.empty((nt.value, 2), dtype=np.uint32)
        ids = np.empty(max(nid.value, 1), dtype=np.uint32)
        full = np.empty(nt.value, dtype=np.uint32)
        check(self._lib.nps_read_tiles(self._h, uptr(ranges), uptr(ids), uptr(full), None, None), self._h,
              "read_tiles")
        return ranges, ids[: nid.value], full

    def read_tile_work(self) -> np.ndarray:
        """EXT: [T, 2] entries staged by the forward, reverse extent (last backward/forward)."""
        nt = C.c_int32()
        check(self._lib.nps_read_tiles(self._h, None, None, None, C.byref(nt), None), self._h, "read_tiles")
        w = np.empty((nt.value, 2), dtype=np.uint32)
        check(self._lib.nps_read_tile_work(self._h, uptr(w)), self._h, "read_tile_work")
        return w

    def read_stats(self) -> dict:
        s = np.zeros(12, dtype=np.int64)
        check(self._lib.nps_read_stats(self._h, lptr(s)), self._h, "read_stats")
        keys = ["keys", "tiles_over_cap", "pairs", "active_pairs", "adc_qualifiers", "adc_children", "adc_splits",
                "adc_clones", "adc_rounds", "reverse_wave_entries", "forward_wave_evals"]
        return dict(zip(keys, (int(x) for x in s)))

    def enable_timing(self, on: bool = True):
        check(self._lib.nps_enable_timing(self._h, 1 if on else 0), self._h, "enable_timing")

    def read_timing(self) -> dict:
        t = np.zeros(8, dtype=np.float32)
        check(self._lib.nps_read_timing(self._h, fptr(t)), self._h, "read_timing")
        keys = ["preprocess_keys", "sort", "ranges", "composite", "step", "adc", "backward", "chain"]
        return dict(zip(keys, (float(x) for x in t)))


class GpuSplatForwardPipelineManager:
    """new GpuSplatForwardPipelineManager({device, numSplats, splatBuffer}) (FWDM:17-92).

    ``splatBuffer`` is the optimizer manager whose splats are borrowed."""

    def __init__(self, device=None, numSplats: int | None = None, splatBuffer: GpuSplatOptimizerManager = None):
        if splatBuffer is None:
            raise ValueError("splatBuffer (a GpuSplatOptimizerManager) is required")
        lib = _lib.load()
        f = C.c_void_p()
        check(lib.npf_create(splatBuffer.handle, C.byref(f)), splatBuffer.handle, "npf_create")
        self._lib = lib
        self._f = f
        self._owner = splatBuffer
        self.numSplats = splatBuffer.numSplats
        self.dims = (0, 0)

    def close(self):
        if getattr(self, "_f", None):
            self._lib.npf_destroy(self._f)
            self._f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def writeVPMatrix(self, mat):
        vp = np.ascontiguousarray(mat, dtype=np.float32).reshape(16)
        check(self._lib.npf_write_vp(self._f, fptr(vp)), self._owner.handle, "writeVPMatrix")

    def setTarget(self, targetColorView=None, targetDepthView=None, width: int = 0, height: int = 0):
        check(self._lib.npf_set_target(self._f, int(width), int(height)), self._owner.handle, "setTarget")
        self.dims = (int(width), int(height))
        if targetColorView is not None or targetDepthView is not None:
            c = None if targetColorView is None else np.ascontiguousarray(targetColorView, dtype=np.float32)
            d = None if targetDepthView is None else np.ascontiguousarray(targetDepthView, dtype=np.float32)
            check(self._lib.npf_write_target(self._f, fptr(c) if c is not None else None,
                                             fptr(d) if d is not None else None), self._owner.handle, "setTarget")

    def dispatch(self, commandEncoder=None, clear: bool = False, quantize: bool = True):
        check(self._lib.npf_dispatch(self._f, 1 if clear else 0, 1 if quantize else 0), self._owner.handle,
              "dispatch")

    def read_target(self):
        w, h = self.dims
        c = np.empty((h, w, 4), dtype=np.float32)
        d = np.empty((h, w, 4), dtype=np.float32)
        check(self._lib.npf_read_target(self._f, fptr(c), fptr(d)), self._owner.handle, "read_target")
        return c, d


class GpuMultiViewStep:
    """Multi-view training step over several EXT scenes (npt_* of the C ABI; no
    reference counterpart).  ``step(graph=True)`` replays one captured hipGraph
    for all views + Adam; ``step(graph=False)`` runs view after view."""

    def __init__(self, scenes, lanes: int = 4):
        lib = _lib.load()
        self._scenes = list(scenes)
        arr = (C.c_void_p * len(self._scenes))(*[m.handle for m in self._scenes])
        t = C.c_void_p()
        rc = lib.npt_create(arr, len(self._scenes), int(lanes), C.byref(t))
        if rc != 0:
            raise _lib.NpsError(f"npt_create failed ({rc})")
        self._lib = lib
        self._t = t
        self._targets = []
        self.num_views = 0

    def _check(self, rc, what):
        if rc != 0:
            m = self._lib.npt_last_error(self._t)
            raise _lib.NpsError(f"{what} failed ({rc}): {m.decode() if m else ''}")

    def add_view(self, scene: int, vp, vp_inv, target_rgba: np.ndarray):
        """Uploads the target (H x W x 4) to the scene's GPU and registers the view."""
        owner = self._scenes[scene]
        rgba = np.ascontiguousarray(target_rgba, dtype=np.float32)
        h, w = rgba.shape[:2]
        p = C.c_void_p()
        check(self._lib.nps_dev_alloc(owner.handle, rgba.nbytes, C.byref(p)), owner.handle, "nps_dev_alloc")
        check(self._lib.nps_dev_upload(owner.handle, p, rgba.ctypes.data_as(C.c_void_p), rgba.nbytes), owner.handle,
              "nps_dev_upload")
        self._targets.append((owner, p))
        a = np.ascontiguousarray(vp, dtype=np.float32).reshape(16)
        b = np.ascontiguousarray(vp_inv, dtype=np.float32).reshape(16)
        self._check(self._lib.npt_add_view(self._t, int(scene), fptr(a), fptr(b), p, int(w), int(h)), "add_view")
        self.num_views += 1

    def step(self, graph: bool = True, serial: bool = False):
        """graph: the whole step as one captured graph; otherwise view by view,
        overlapping across scenes, or (serial) each view completed before the next."""
        self._check(self._lib.npt_step(self._t, 1 if graph else (2 if serial else 0)), "step")

    def autotune(self) -> np.ndarray:
        """Chooses tile size and LDS batch per (scene, resolution); returns [V, 2]."""
        self._check(self._lib.npt_autotune(self._t), "autotune")
        out = np.zeros((self.num_views, 2), dtype=np.int32)
        self._check(self._lib.npt_read_tune(self._t, iptr(out)), "read_tune")
        return out

    def read_losses(self) -> np.ndarray:
        out = np.zeros((self.num_views, 2), dtype=np.float64)
        self._check(self._lib.npt_read_losses(self._t, dptr(out)), "read_losses")
        return out

    def graph_info(self) -> dict:
        """Nodes of the last captured graph, lanes in use, and the last autotune's
        in-context replay times (ms) of {own settings, per-class choices} x
        {all lanes, half of them} (-1: not run)."""
        info = np.zeros(2, dtype=np.int64)
        ms = np.zeros(4, dtype=np.float32)
        self._check(self._lib.npt_graph_info(self._t, info.ctypes.data_as(C.POINTER(C.c_int64)), fptr(ms)),
                    "graph_info")
        return {"nodes": int(info[0]), "lanes": int(info[1]), "autotune_ms": [round(float(x), 3) for x in ms]}

    def close(self):
        if getattr(self, "_t", None):
            self._lib.npt_destroy(self._t)
            self._t = None
            for owner, p in self._targets:
                self._lib.nps_dev_free(owner.handle, p)
            self._targets = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBlurPipelineManager:
    """new GpuBlurPipelineManager(device) (src/lib/gpu/GpuBlurPipelineManager.ts:9-49)."""

    def __init__(self, device=None, *, hip_device: int = 0):
        lib = _lib.load()
        b = C.c_void_p()
        check(lib.npb_create(int(hip_device), C.byref(b)), None, "npb_create")
        self._lib = lib
        self._b = b

    def close(self):
        if getattr(self, "_b", None):
            self._lib.npb_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def blur(self, commandEncoder=None, srcView: np.ndarray = None, dstView=None, tempView=None, width: int = 0,
             height: int = 0, radius: int = 16, sigma: float | None = None, isSrgb: bool = False) -> np.ndarray:
        """GpuBlurPipelineManager.blur (GpuBlurPipelineManager.ts:51-99).  srcView is an
        H x W x 4 float32 array of texel values; returns the blurred texels (and
        writes them into dstView when given).  tempView is accepted and unused."""
        src = np.ascontiguousarray(srcView, dtype=np.float32).reshape(height, width, 4)
        out = np.empty((height, width, 4), dtype=np.float32)
        sigma = radius / 2 if sigma is None else sigma
        rc = self._lib.npb_blur(self._b, fptr(src), fptr(out), int(width), int(height), int(radius), float(sigma),
                                1 if isSrgb else 0)
        if rc != 0:
            m = self._lib.npb_last_error(self._b)
            raise _lib.NpsError(f"blur failed ({rc}): {m.decode() if m else ''}")
        if dstView is not None:
            dstView[...] = out
        return out


class GpuPathTracePipelineManager:
    """new GpuPathTracePipelineManager({device, envTexture})
    (src/lib/gpu/GpuPathTracePipelineManager.ts:8-248): the progressive path
    tracer that produces the optimisation target of the reference's frame.
    envTexture is an H x W x 4 float array (rgb used, HDR allowed)."""

    def __init__(self, device=None, envTexture: np.ndarray | None = None, *, hip_device: int = 0):
        lib = _lib.load()
        p = C.c_void_p()
        check(lib.npp_create(int(hip_device), C.byref(p)), None, "npp_create")
        self._lib = lib
        self._p = p
        self.accumWidth = 0
        self.accumHeight = 0
        if envTexture is not None:
            self.setEnvTexture(envTexture)

    def _ok(self, rc, what):
        if rc != 0:
            m = self._lib.npp_last_error(self._p)
            raise _lib.NpsError(f"{what} failed ({rc}): {m.decode() if m else ''}")

    def close(self):
        if getattr(self, "_p", None):
            self._lib.npp_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setEnvTexture(self, env: np.ndarray):
        e = np.ascontiguousarray(env, dtype=np.float32)
        self._ok(self._lib.npp_set_env(self._p, fptr(e), int(e.shape[1]), int(e.shape[0])), "npp_set_env")

    def setMeshes(self, meshes):
        """setMeshes (:101-150): meshes is a list of (vertices [n, 12], indices [m, 3])
        concatenated with each mesh's indices offset by the vertices before it."""
        vs, ids, off = [], [], 0
        for v, i in meshes:
            v = np.asarray(v, np.float32).reshape(-1, 12)
            vs.append(v)
            ids.append(np.asarray(i, np.uint32).reshape(-1) + np.uint32(off))
            off += v.shape[0]
        v = np.ascontiguousarray(np.concatenate(vs) if vs else np.zeros((0, 12), np.float32))
        i = np.ascontiguousarray(np.concatenate(ids) if ids else np.zeros(0, np.uint32))
        self._ok(self._lib.npp_set_meshes(self._p, fptr(v), int(v.shape[0]),
                                          i.ctypes.data_as(C.POINTER(C.c_uint32)), int(i.size // 3)),
                 "npp_set_meshes")

    def setOutputSize(self, width: int, height: int):
        self._ok(self._lib.npp_set_output_size(self._p, int(width), int(height)), "npp_set_output_size")
        self.accumWidth, self.accumHeight = int(width), int(height)

    def reset(self):
        self._ok(self._lib.npp_reset(self._p), "npp_reset")

    def writeInvViewProjMat(self, mat):
        m = np.ascontiguousarray(np.asarray(mat, np.float32).reshape(16))
        self._ok(self._lib.npp_write_inv_vp(self._p, fptr(m)), "npp_write_inv_vp")

    def dispatch(self, commandEncoder=None):
        self._ok(self._lib.npp_dispatch(self._p), "npp_dispatch")

    def read_output(self) -> np.ndarray:
        """outputTextureView: H x W x 4 rgba8 texel values (k/255)."""
        out = np.empty((self.accumHeight, self.accumWidth, 4), np.float32)
        self._ok(self._lib.npp_read_output(self._p, fptr(out)), "npp_read_output")
        return out

    def read_accum(self) -> np.ndarray:
        out = np.empty((self.accumHeight, self.accumWidth, 4), np.float32)
        self._ok(self._lib.npp_read_accum(self._p, fptr(out)), "npp_read_accum")
        return out

    def output_dev(self) -> int:
        """Device address of the output texture (for set_target_dev / blur_dev)."""
        q = C.c_void_p()
        self._ok(self._lib.npp_output_dev(self._p, C.byref(q)), "npp_output_dev")
        return int(q.value)

    def read_bvh(self):
        counts = np.zeros(2, np.int32)
        I = C.POINTER(C.c_int32)
        U = C.POINTER(C.c_uint32)
        self._ok(self._lib.npp_read_bvh(self._p, None, None, counts.ctypes.data_as(I)), "npp_read_bvh")
        nodes = np.zeros((int(counts[0]), 8), np.uint32)
        tris = np.zeros((max(int(counts[1]), 1), 3), np.uint32)
        self._ok(self._lib.npp_read_bvh(self._p, nodes.ctypes.data_as(U), tris.ctypes.data_as(U),
                                        counts.ctypes.data_as(I)), "npp_read_bvh")
        return nodes, tris[:int(counts[1])]

    def env_background(self, inv_vp, width: int, height: int) -> np.ndarray:
        """The environment background pass of the raster target (envmap.wgsl)."""
        m = np.ascontiguousarray(np.asarray(inv_vp, np.float32).reshape(16))
        out = np.empty((height, width, 4), np.float32)
        self._ok(self._lib.npp_env_background(self._p, fptr(m), fptr(out), int(width), int(height)),
                 "npp_env_background")
        return out


class GpuMeshRenderPipelineManager:
    """new GpuMeshRenderPipelineManager({device, format, uniformsManager, mesh})
    (src/lib/gpu/GpuMeshRenderPipelineManager.ts:17-280): the raster target of
    the reference's frame (colour and encoded depth of the mesh over the envmap
    background).  mesh and setGroundMesh take (vertices [n, 12], indices [m, 3]);
    the matcap texture is an H x W x 4 float array (GpuRunner passes its
    environment panorama).  The uniforms of GpuUniformsBufferManager (view-
    projection, view matrix, shading mode, inverse view-projection) are given
    with writeUniforms.  setPbrMesh's textured ground is not built (its texture
    assets are not shipped with the reference)."""

    def __init__(self, device=None, format=None, uniformsManager=None, mesh=None, *,
                 matcapTexture: np.ndarray | None = None, hip_device: int = 0):
        lib = _lib.load()
        p = C.c_void_p()
        check(lib.npp_create(int(hip_device), C.byref(p)), None, "npp_create")
        self._lib = lib
        self._p = p
        self._meshes = [] if mesh is None else [mesh]
        self._ground = None
        self._u = None
        if mesh is not None:
            self._upload()
        if matcapTexture is not None:
            self.setMatcapTexture(matcapTexture)

    def _ok(self, rc, what):
        if rc != 0:
            m = self._lib.npp_last_error(self._p)
            raise _lib.NpsError(f"{what} failed ({rc}): {m.decode() if m else ''}")

    def close(self):
        if getattr(self, "_p", None):
            self._lib.npp_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _upload(self):
        ms = list(self._meshes) + ([self._ground] if self._ground is not None else [])
        vs, ids, off = [], [], 0
        for v, i in ms:
            v = np.asarray(v, np.float32).reshape(-1, 12)
            vs.append(v)
            ids.append(np.asarray(i, np.uint32).reshape(-1) + np.uint32(off))
            off += v.shape[0]
        v = np.ascontiguousarray(np.concatenate(vs) if vs else np.zeros((0, 12), np.float32))
        i = np.ascontiguousarray(np.concatenate(ids) if ids else np.zeros(0, np.uint32))
        self._ok(self._lib.npp_set_meshes(self._p, fptr(v), int(v.shape[0]), i.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          int(i.size // 3)), "npp_set_meshes")

    def setGroundMesh(self, mesh):
        self._ground = mesh
        self._upload()

    def setPbrMesh(self, *args, **kwargs):
        raise NotImplementedError("the textured (PBR) ground needs texture assets the reference does not ship")

    def setMatcapTexture(self, tex: np.ndarray):
        e = np.ascontiguousarray(tex, dtype=np.float32)
        self._ok(self._lib.npp_set_env(self._p, fptr(e), int(e.shape[1]), int(e.shape[0])), "npp_set_env")

    def writeUniforms(self, viewProjMat, viewMat, invViewProjMat, shadingMode: float = 1.0):
        self._u = tuple(np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
                        for m in (viewProjMat, viewMat, invViewProjMat)) + (float(shadingMode),)

    def addDraw(self, renderPassEncoder=None, matcapTextureView=None, width: int = 0, height: int = 0):
        """The envmap pass and this manager's draws into a cleared W x H pass:
        returns (colour, depth), each H x W x 4 rgba8 texel values."""
        if self._u is None:
            raise _lib.NpsError("no uniforms (writeUniforms)")
        if matcapTextureView is not None:
            self.setMatcapTexture(matcapTextureView)
        vp, view, ivp, mode = self._u
        col = np.empty((height, width, 4), np.float32)
        dep = np.empty((height, width, 4), np.float32)
        self._ok(self._lib.npp_raster(self._p, fptr(vp), fptr(view), fptr(ivp), mode, fptr(col), fptr(dep),
                                      int(width), int(height)), "npp_raster")
        return col, dep


def reference_bezier_init(num_beziers: int, seed: int = 0) -> np.ndarray:
    """The reference's initial curves (GpuBezierOptimizerManager.ts:56-95: short
    randomly oriented squiggles near the origin, width 0.02, softness 0.005,
    alpha 0.5) drawn from a seeded generator instead of Math.random()."""
    rng = np.random.default_rng(seed)
    n = int(num_beziers)
    d = np.zeros((n, 20), np.float32)
    cx = rng.random(n) * 2 - 1
    cy = rng.random(n) * 2 - 1
    ln = 0.2 + rng.random(n) * 0.2
    ang = rng.random(n) * np.pi * 2
    dx, dy = np.cos(ang) * ln, np.sin(ang) * ln

    def jit():
        return (rng.random(n) - 0.5) * 0.08
    d[:, 0], d[:, 1], d[:, 2], d[:, 3] = cx - dx * 0.5, cy - dy * 0.5, (rng.random(n) * 2 - 1) * 0.3, 0.02
    d[:, 4], d[:, 5], d[:, 6], d[:, 7] = cx - dx * 0.15 + jit(), cy - dy * 0.15 + jit(), (rng.random(n) * 2 - 1) * 0.3, 0.005
    d[:, 8], d[:, 9], d[:, 10] = cx + dx * 0.15 + jit(), cy + dy * 0.15 + jit(), (rng.random(n) * 2 - 1) * 0.3
    d[:, 12], d[:, 13], d[:, 14] = cx + dx * 0.5, cy + dy * 0.5, (rng.random(n) * 2 - 1) * 0.3
    d[:, 16], d[:, 17], d[:, 18] = rng.random(n), rng.random(n), rng.random(n)
    d[:, 19] = 0.5
    return d


class GpuBezierOptimizerManager:
    """new GpuBezierOptimizerManager({device, numBeziers}) (src/lib/gpu/GpuBezierOptimizerManager.ts:17-406).

    Texture views are host arrays: H x W x 4 float32 for the target, background
    and normal images, H x W float32 (or the .r of H x W x 4) for the two depths.
    The constructor writes the reference's initial curve distribution from a
    seeded generator (``reference_bezier_init``)."""

    def __init__(self, device=None, numBeziers: int = 16, *, hip_device: int = 0, seed: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        rc = lib.nbz_create(int(hip_device), int(numBeziers), C.byref(h))
        if rc != 0:
            raise _lib.NpsError(f"nbz_create failed ({rc})")
        self._lib = lib
        self._h = h
        self.numBeziers = int(numBeziers)
        self.numParams = self.numBeziers * 18
        self.dims = (0, 0)
        self.write_beziers(reference_bezier_init(self.numBeziers, seed))

    def _check(self, rc, what):
        if rc != 0:
            m = self._lib.nbz_last_error(self._h)
            raise _lib.NpsError(f"{what} failed ({rc}): {m.decode() if m else ''}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nbz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- buffers
    def write_beziers(self, curves: np.ndarray):
        c = np.ascontiguousarray(curves, dtype=np.float32).reshape(self.numBeziers, 20)
        self._check(self._lib.nbz_write_beziers(self._h, fptr(c), self.numBeziers), "nbz_write_beziers")

    def read_beziers(self) -> np.ndarray:
        c = np.empty((self.numBeziers, 20), np.float32)
        self._check(self._lib.nbz_read_beziers(self._h, fptr(c), self.numBeziers), "nbz_read_beziers")
        return c

    def write_grads(self, grads: np.ndarray):
        g = np.ascontiguousarray(grads, dtype=np.int32).reshape(self.numBeziers, 18)
        self._check(self._lib.nbz_write_grads(self._h, iptr(g), self.numBeziers), "nbz_write_grads")

    def read_grads(self) -> np.ndarray:
        g = np.empty((self.numBeziers, 18), np.int32)
        self._check(self._lib.nbz_read_grads(self._h, iptr(g), self.numBeziers), "nbz_read_grads")
        return g

    def write_adam(self, m: np.ndarray, v: np.ndarray, t: float):
        mm = np.ascontiguousarray(m, dtype=np.float32).reshape(self.numBeziers, 18)
        vv = np.ascontiguousarray(v, dtype=np.float32).reshape(self.numBeziers, 18)
        self._check(self._lib.nbz_write_adam(self._h, fptr(mm), fptr(vv), float(t), self.numBeziers),
                    "nbz_write_adam")

    def read_adam(self):
        m = np.empty((self.numBeziers, 18), np.float32)
        v = np.empty((self.numBeziers, 18), np.float32)
        t = C.c_float()
        self._check(self._lib.nbz_read_adam(self._h, fptr(m), fptr(v), C.byref(t), self.numBeziers),
                    "nbz_read_adam")
        return m, v, float(t.value)

    def write_adc(self, grad_accum: np.ndarray, loss_accum: np.ndarray):
        ga = np.ascontiguousarray(grad_accum, dtype=np.float32).reshape(self.numBeziers)
        la = np.ascontiguousarray(loss_accum, dtype=np.int64).reshape(self.numBeziers)
        self._check(self._lib.nbz_write_adc(self._h, fptr(ga), lptr(la), self.numBeziers), "nbz_write_adc")

    def read_adc(self):
        """(grad_accum float32 [n], loss_accum int64 [n] in 2^-32 fixed point)."""
        ga = np.empty(self.numBeziers, np.float32)
        la = np.empty(self.numBeziers, np.int64)
        self._check(self._lib.nbz_read_adc(self._h, fptr(ga), lptr(la), self.numBeziers), "nbz_read_adc")
        return ga, la

    def write_pixel_loss(self, pl: np.ndarray):
        p = np.ascontiguousarray(pl, dtype=np.int32).reshape(-1)
        self._check(self._lib.nbz_write_pixel_loss(self._h, iptr(p)), "nbz_write_pixel_loss")

    def read_pixel_loss(self) -> np.ndarray:
        w, h = self.dims
        p = np.empty(w * h, np.int32)
        self._check(self._lib.nbz_read_pixel_loss(self._h, iptr(p)), "nbz_read_pixel_loss")
        return p

    # ---- the reference's methods
    def writeVPMatrix(self, mat):
        self._check(self._lib.nbz_write_vp_matrix(self._h, fptr(np.ascontiguousarray(mat, np.float32).reshape(16))),
                    "writeVPMatrix")

    def writeVPInvMatrix(self, mat):
        self._check(self._lib.nbz_write_vp_inv_matrix(self._h,
                                                      fptr(np.ascontiguousarray(mat, np.float32).reshape(16))),
                    "writeVPInvMatrix")

    def writeMode(self, mode: float = 0):
        self._check(self._lib.nbz_write_mode(self._h, float(mode)), "writeMode")

    def writeBgPenalty(self, weight: float = 0):
        self._check(self._lib.nbz_write_bg_penalty(self._h, float(weight)), "writeBgPenalty")

    def writeMaxWidth(self, maxWidth: float = 0):
        self._check(self._lib.nbz_write_max_width(self._h, float(maxWidth)), "writeMaxWidth")

    def writeKillThresholds(self, alphaThresh: float = 0, widthThresh: float = 0):
        self._check(self._lib.nbz_write_kill_thresholds(self._h, float(alphaThresh), float(widthThresh)),
                    "writeKillThresholds")

    def setAdcPeriod(self, period: int):
        self._check(self._lib.nbz_set_adc_period(self._h, int(period)), "setAdcPeriod")

    def writeNoKill(self, noKill: bool):
        self._check(self._lib.nbz_write_no_kill(self._h, 1 if noKill else 0), "writeNoKill")

    def setBackwardTarget(self, targetTextureView, targetDepthTextureView, bgColorTextureView, bgDepthTextureView,
                          normalTextureView, width: int, height: int):
        W, H = int(width), int(height)

        def rgba(a):
            return np.ascontiguousarray(np.asarray(a, np.float32).reshape(H, W, 4))

        def red(a):
            a = np.asarray(a, np.float32)
            return np.ascontiguousarray(a[..., 0] if a.ndim == 3 else a.reshape(H, W))
        t, td, b, bd, nr = (rgba(targetTextureView), red(targetDepthTextureView), rgba(bgColorTextureView),
                            red(bgDepthTextureView), rgba(normalTextureView))
        self._check(self._lib.nbz_set_backward_target(self._h, fptr(t), fptr(td), fptr(b), fptr(bd), fptr(nr), W, H),
                    "setBackwardTarget")
        self.dims = (W, H)

    def set_backward_target_dev(self, tgt_ptr, tdepth_ptr, bg_ptr, bgdepth_ptr, normal_ptr, width: int,
                                height: int):
        """setBackwardTarget with device pointers (e.g. torch tensors' data_ptr()) kept by reference."""
        self._check(self._lib.nbz_set_backward_target_dev(self._h, C.c_void_p(tgt_ptr), C.c_void_p(tdepth_ptr),
                                                          C.c_void_p(bg_ptr), C.c_void_p(bgdepth_ptr),
                                                          C.c_void_p(normal_ptr), int(width), int(height)),
                    "setBackwardTarget")
        self.dims = (int(width), int(height))

    def dispatch(self, commandEncoder=None):
        self._check(self._lib.nbz_dispatch(self._h), "dispatch")

    # ---- passes and readback
    def backward(self):
        self._check(self._lib.nbz_backward(self._h), "nbz_backward")

    def forward(self):
        self._check(self._lib.nbz_forward(self._h), "nbz_forward")

    def step(self):
        self._check(self._lib.nbz_step(self._h), "nbz_step")

    def adc(self):
        self._check(self._lib.nbz_adc(self._h), "nbz_adc")

    def sync(self):
        self._check(self._lib.nbz_sync(self._h), "nbz_sync")

    def read_image(self):
        """(colour [H,W,3], encoded depth [H,W], final transmittance [H,W])."""
        w, h = self.dims
        c = np.empty((h, w, 3), np.float32)
        d = np.empty((h, w), np.float32)
        t = np.empty((h, w), np.float32)
        self._check(self._lib.nbz_read_image(self._h, fptr(c), fptr(d), fptr(t)), "nbz_read_image")
        return c, d, t

    def read_loss(self) -> np.ndarray:
        out = np.zeros(2, np.float64)
        self._check(self._lib.nbz_read_loss(self._h, dptr(out)), "nbz_read_loss")
        return out

    def read_tiles(self):
        """(ranges [T,2], ids [M]) of the last backward / forward."""
        nt = C.c_int32()
        nk = C.c_int64()
        self._check(self._lib.nbz_read_tiles(self._h, None, None, C.byref(nt), C.byref(nk)), "nbz_read_tiles")
        ranges = np.empty((nt.value, 2), np.uint32)
        ids = np.empty(max(nk.value, 1), np.uint32)
        self._check(self._lib.nbz_read_tiles(self._h, uptr(ranges), uptr(ids), C.byref(nt), C.byref(nk)),
                    "nbz_read_tiles")
        return ranges, ids[:nk.value]

    def read_stats(self) -> dict:
        s = np.zeros(16, np.int64)
        self._check(self._lib.nbz_read_stats(self._h, lptr(s)), "nbz_read_stats")
        keys = ["keys", "tiles_over_cap", "pairs", "reverse_pairs", "reverse_wave_entries", "evaluated_wave_entries",
                "adc_dead", "adc_killed", "adc_split1", "adc_spawned", "adc_split3", "adc_dead_left",
                "reverse_evaluated_wave_entries"]
        return {k: int(v) for k, v in zip(keys, s)}

    def enable_timing(self, on: bool = True):
        self._check(self._lib.nbz_enable_timing(self._h, 1 if on else 0), "nbz_enable_timing")

    def read_timing(self) -> dict:
        ms = np.zeros(8, np.float32)
        self._check(self._lib.nbz_read_timing(self._h, fptr(ms)), "nbz_read_timing")
        return {"preprocess": float(ms[0]), "keys": float(ms[1]), "sort": float(ms[2]), "composite": float(ms[3]),
                "step": float(ms[4]), "adc": float(ms[5]), "backward": float(ms[6])}

    def dev_alloc(self, nbytes: int) -> int:
        """A device buffer on the handle's GPU (nbz_dev_alloc); free with dev_free."""
        p = C.c_void_p()
        self._check(self._lib.nbz_dev_alloc(self._h, int(nbytes), C.byref(p)), "nbz_dev_alloc")
        return p.value

    def dev_free(self, ptr: int):
        self._lib.nbz_dev_free(self._h, C.c_void_p(ptr))

    def dev_upload(self, ptr: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self._check(self._lib.nbz_dev_upload(self._h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes),
                    "nbz_dev_upload")


class GpuBezierForwardPipelineManager:
    """new GpuBezierForwardPipelineManager({device, numBeziers, bezierBuffer, brushTexture})
    (src/lib/gpu/GpuBezierForwardPipelineManager.ts:17-98): the display raster of the curves.

    ``bezierBuffer`` is the GpuBezierOptimizerManager whose curves are borrowed;
    ``brushTexture`` is an H x W x 4 uint8 array (rgba8 texels; the red channel
    is sampled)."""

    def __init__(self, device=None, numBeziers: int | None = None,
                 bezierBuffer: GpuBezierOptimizerManager = None, brushTexture=None):
        if bezierBuffer is None:
            raise ValueError("bezierBuffer (a GpuBezierOptimizerManager) is required")
        if brushTexture is None:
            raise ValueError("brushTexture (H x W x 4 uint8) is required")
        tex = np.ascontiguousarray(brushTexture, dtype=np.uint8)
        if tex.ndim != 3 or tex.shape[2] != 4:
            raise ValueError("brushTexture must be H x W x 4 uint8")
        lib = _lib.load()
        f = C.c_void_p()
        bezierBuffer._check(lib.nbf_create(bezierBuffer._h, tex.ctypes.data_as(C.c_void_p), tex.shape[1],
                                           tex.shape[0], C.byref(f)), "nbf_create")
        self._lib = lib
        self._f = f
        self._owner = bezierBuffer
        self.numBeziers = bezierBuffer.numBeziers
        self.dims = (0, 0)

    def close(self):
        if getattr(self, "_f", None):
            self._lib.nbf_destroy(self._f)
            self._f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def writeVPMatrix(self, mat):
        vp = np.ascontiguousarray(mat, dtype=np.float32).reshape(16)
        self._owner._check(self._lib.nbf_write_vp(self._f, fptr(vp)), "writeVPMatrix")

    def setTarget(self, targetView=None, width: int = 0, height: int = 0):
        self._owner._check(self._lib.nbf_set_target(self._f, int(width), int(height)), "setTarget")
        self.dims = (int(width), int(height))
        if targetView is not None:
            c = np.ascontiguousarray(targetView, dtype=np.float32)
            self._owner._check(self._lib.nbf_write_target(self._f, fptr(c)), "setTarget")

    def dispatch(self, commandEncoder=None, clear: bool = True, quantize: bool = True):
        self._owner._check(self._lib.nbf_dispatch(self._f, 1 if clear else 0, 1 if quantize else 0), "dispatch")

    def read_target(self) -> np.ndarray:
        w, h = self.dims
        c = np.empty((h, w, 4), dtype=np.float32)
        self._owner._check(self._lib.nbf_read_target(self._f, fptr(c)), "read_target")
        return c


class DensifierMLP:
    """The densification MLP (nml_*): 32 parent features -> 14 child deltas,
    BF16 matrix-core products with FP32 accumulation, trained with Adam."""

    NUM_PARAMS = 143630

    def __init__(self, max_batch: int = 8192, *, hip_device: int = 0, params: np.ndarray | None = None):
        lib = _lib.load()
        h = C.c_void_p()
        rc = lib.nml_create(int(hip_device), int(max_batch), C.byref(h))
        if rc != 0:
            raise _lib.NpsError(f"nml_create failed ({rc})")
        self._lib = lib
        self._h = h
        self.max_batch = int(max_batch)
        if params is not None:
            self.write_params(params)

    def _check(self, rc, what):
        if rc != 0:
            m = self._lib.nml_last_error(self._h)
            raise _lib.NpsError(f"{what} failed ({rc}): {m.decode() if m else ''}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.nml_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write_params(self, p: np.ndarray):
        q = np.ascontiguousarray(p, dtype=np.float32).reshape(self.NUM_PARAMS)
        self._check(self._lib.nml_write_params(self._h, fptr(q)), "nml_write_params")

    def read_params(self) -> np.ndarray:
        q = np.empty(self.NUM_PARAMS, np.float32)
        self._check(self._lib.nml_read_params(self._h, fptr(q)), "nml_read_params")
        return q

    def set_norm(self, in_mean, in_std, out_mean, out_std):
        a = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in (in_mean, in_std, out_mean, out_std)]
        self._check(self._lib.nml_set_norm(self._h, *[fptr(v) for v in a]), "nml_set_norm")

    def forward(self, x: np.ndarray) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 32)
        y = np.empty((xx.shape[0], 14), np.float32)
        self._check(self._lib.nml_forward(self._h, fptr(xx), xx.shape[0], fptr(y)), "nml_forward")
        return y

    def forward_dev(self, x_ptr: int, n: int, y_ptr: int):
        self._check(self._lib.nml_forward_dev(self._h, C.c_void_p(x_ptr), int(n), C.c_void_p(y_ptr)),
                    "nml_forward_dev")

    def grads(self, x: np.ndarray, t: np.ndarray):
        """(loss, gradient [143630]) of the standardised MSE over one batch."""
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 32)
        tt = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 14)
        g = np.empty(self.NUM_PARAMS, np.float32)
        loss = C.c_double()
        self._check(self._lib.nml_grads(self._h, fptr(xx), fptr(tt), xx.shape[0], fptr(g), C.byref(loss)),
                    "nml_grads")
        return float(loss.value), g

    def train(self, x: np.ndarray, t: np.ndarray, lr: float) -> float:
        """One Adam step on one batch of host arrays; returns the batch loss."""
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 32)
        tt = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 14)
        loss = C.c_double()
        self._check(self._lib.nml_train(self._h, fptr(xx), fptr(tt), xx.shape[0], float(lr), C.byref(loss)),
                    "nml_train")
        return float(loss.value)

    def train_dev(self, x_ptr: int, t_ptr: int, n: int, lr: float) -> float:
        loss = C.c_double()
        self._check(self._lib.nml_train_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(t_ptr), int(n), float(lr),
                                            C.byref(loss)), "nml_train_dev")
        return float(loss.value)

    def enable_timing(self, on: bool = True):
        self._check(self._lib.nml_enable_timing(self._h, 1 if on else 0), "nml_enable_timing")

    def read_timing(self) -> dict:
        """HIP-event ms of the last inference launch and the last training step."""
        ms = np.zeros(2, np.float32)
        self._check(self._lib.nml_read_timing(self._h, fptr(ms)), "nml_read_timing")
        return {"forward": float(ms[0]), "train_step": float(ms[1])}

    def dev_alloc(self, nbytes: int) -> int:
        """A device buffer on the handle's GPU (nml_dev_alloc); free with dev_free."""
        p = C.c_void_p()
        self._check(self._lib.nml_dev_alloc(self._h, int(nbytes), C.byref(p)), "nml_dev_alloc")
        return p.value

    def dev_free(self, ptr: int):
        self._lib.nml_dev_free(self._h, C.c_void_p(ptr))

    def dev_upload(self, ptr: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self._check(self._lib.nml_dev_upload(self._h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes),
                    "nml_dev_upload")


def debug_math(op: int, x: np.ndarray, y: np.ndarray | None = None, device: int = 0) -> np.ndarray:
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float32)
    check(lib.nps_debug_math(device, op, fptr(x), fptr(yy) if yy is not None else None, fptr(out), x.size), None,
          "nps_debug_math")
    return out


def debug_depth_sort(keys: np.ndarray, narrow: bool, device: int = 0):
    """EXT depth order of keys depth_bits << 32 | id on the GPU (nps_debug_depth_sort):
    (sorted keys, flags [any overflow, depth range too wide])."""
    lib = _lib.load()
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty_like(k)
    fl = np.zeros(2, np.uint32)
    check(lib.nps_debug_depth_sort(device, k.ctypes.data_as(C.POINTER(C.c_uint64)), k.size, 1 if narrow else 0,
                                   out.ctypes.data_as(C.POINTER(C.c_uint64)), uptr(fl)), None,
          "nps_debug_depth_sort")
    return out, fl


def compose_display(textures, enabled, width: int, height: int, device: int = 0) -> np.ndarray:
    """The viewer's debug display compositor (src/lib/gpu/splat_render.wgsl:24-181)
    through npd_compose.  textures: 8 (H_k, W_k, 4) float arrays in the
    shader's binding order (target, splat view, target depth, target edges,
    edge-layer, base-colour and colour Bezier views, path trace); enabled: the
    edge, base-colour and colour Bezier layers.  Returns the (height, width, 4)
    canvas."""
    if len(textures) != 8:
        raise ValueError("npd_compose takes 8 textures")
    lib = _lib.load()
    arrs = [np.ascontiguousarray(t, np.float32) for t in textures]
    for a in arrs:
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("textures must be (H, W, 4)")
    F = C.POINTER(C.c_float)
    ptrs = (F * 8)(*[a.ctypes.data_as(F) for a in arrs])
    dims = np.array([[a.shape[1], a.shape[0]] for a in arrs], np.int32).reshape(16)
    en = np.ascontiguousarray(enabled, np.float32).reshape(3)
    out = np.zeros((int(height), int(width), 4), np.float32)
    check(lib.npd_compose(int(device), ptrs, iptr(dims), fptr(en), fptr(out), int(width), int(height)), None,
          "npd_compose")
    return out
