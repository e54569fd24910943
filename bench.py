"""Benchmark: fwd+bwd ms per view at 1M splats, 1920x1080.

--mode ext (default): the scale-up configuration with SH degree 3.  One "step"
= one view of the optimizer loop: backward (preprocess with SH colour, depth
keys, binning into per-tile lists, composite fwd+bwd, per-splat chain rule)
+ Adam.
  1,000,000 Gaussians drawn with seed 0 (means U[-1,1]^3, log-scales
  N(-4,0.5), uniform quaternions, opacity logits N(0,1), SH degree 3: DC from
  RGB U[0,1], higher bands N(0,0.05)), 8 cameras on a radius-3 ring at
  1920x1080 (16x16 tiles), targets = the EXT forward image of a seed-1 scene
  drawn the same way, quantised to k/255.  Densification runs every 100
  steps, so the timed steps (fewer than 100) contain none.

--mode ref: the reference's own splat model (the default run also measures
it, as the "ref" object of the EXT line; --no-ref skips that leg).
Workload (one "step" = one reference dispatch for one view, i.e.
GpuSplatOptimizerManager.dispatch: backward (fwd+loss+bwd) + Adam step, ADC
every 25th step):
  1,000,000 live splats drawn with seed 0 plus 1,000,000 dead slots (capacity
  2M, the fixed-capacity pool the reference's ADC fills), 8 cameras on a
  radius-3 ring at 1920x1080 (16x16 tiles: 120x68 = 8,160 tiles), targets =
  the REF forward image of a seed-1 scene of 1M splats per camera, quantised to
  k/255, target depth 1 (background).  Targets are uploaded to HBM before the
  timed region.  Colour is direct RGB as in the reference (SH is not part of
  the reference path).

Prints one JSON line.  value = mean GPU time of the backward (preprocess, key
generation, sort, ranges, fused composite) per view, from HIP events on the
library's stream; ms_per_step = wall clock per dispatch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic work of the fused composite kernel, counted from the reference
# algorithm (splat_backward.wgsl:230-455) after the per-splat terms are
# hoisted; transcendentals (sqrt, log2, exp2) weighted x4 (quarter-rate VALU).
#   forward, every composited (pixel, splat) pair: 34 FLOP + 4 transcendentals
#   reverse, every pair with alpha >= 0.001: 240 FLOP + 3 transcendentals
#     (colour/opacity/depth terms, shape a/b, conic and axis gradients, screen
#      and depth position terms, quaternion partials, fixed-point conversion)
FWD_FLOP_PER_PAIR = 34 + 4 * 4
BWD_FLOP_PER_ACTIVE_PAIR = 240 + 3 * 4
# EXT composite (exp weighted x4 as above):
#   forward, every (pixel, entry) pair evaluated before the pixel stops:
#     r^2 (10) + exp (1 + 4) + alpha and transmittance (4) + colour (6) = 25
#   reverse, every pair the pixel composited: recomputed alpha and
#     transmittance (19) + pair geometry (10) + colour/opacity/transmittance
#     seeds (22) + screen-centre and conic gradients (12) + colour (4) = 67
EXT_FWD_FLOP_PER_PAIR = 25
EXT_BWD_FLOP_PER_ACTIVE_PAIR = 67
# Bezier composite (bezier_backward.wgsl:196-563), div and sqrt counted once:
#   every (pixel, curve) evaluation of the distance loop (forward, and again in
#     the reverse recompute): 16 segments x 26 (segment, projection parameter,
#     clamp, projection, squared distance, sqrt, strict minimum) + 40 (t,
#     pressure, smoothstep, Bernstein depth) + 10 (compositing) = 466;
#   every reverse pair with alpha >= 0.001: 300 (alpha and transmittance
#     seeds, segment recompute, smoothstep derivatives, depth term, Bernstein
#     weights, 12 backprojections, fixed-point conversion, loss term).
BZ_FLOP_PER_EVAL = 466
BZ_FLOP_PER_REV_PAIR = 300
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def build_workload(n_live, capacity, W, H, views, radius):
    from nprush_amd import GpuSplatOptimizerManager, camera, scenes

    cams = camera.ring_cameras(views, radius, W, H)
    tscene = scenes.make_splats(n_live, 1)
    tmgr = GpuSplatOptimizerManager(numSplats=n_live, splats=tscene)
    dummy = np.zeros((H, W, 4), np.float32)
    ones = np.ones((H, W), np.float32)
    tmgr.setBackwardTarget(dummy, ones, W, H)
    targets = []
    for vp, vi in cams:
        tmgr.writeSplatVPMatrix(vp, vi)
        tmgr.backward()
        c, _ = tmgr.read_image()
        rgba = np.ones((H, W, 4), np.float32)
        rgba[..., :3] = scenes.quantize8(c)
        targets.append(rgba)
    tmgr.close()
    splats = scenes.with_dead_capacity(scenes.make_splats(n_live, 0), capacity)
    return cams, targets, ones, splats


def build_workload_ext(n, W, H, views, radius, deg=3):
    from nprush_amd import GpuSplatOptimizerManager, camera, scenes

    cams = camera.ring_cameras(views, radius, W, H)
    ts, tsh = scenes.make_ext_splats(n, 1, sh_degree=deg)
    tmgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=deg, splats=ts, sh=tsh, adc_enabled=False)
    tmgr.setBackwardTarget(np.zeros((H, W, 4), np.float32), np.ones((H, W), np.float32), W, H)
    targets = []
    for vp, vi in cams:
        tmgr.writeSplatVPMatrix(vp, vi)
        tmgr.forward()
        c, _ = tmgr.read_image()
        rgba = np.ones((H, W, 4), np.float32)
        rgba[..., :3] = scenes.quantize8(c)
        targets.append(rgba)
    tmgr.close()
    s, sh = scenes.make_ext_splats(n, 0, sh_degree=deg)
    return cams, targets, s, sh


def batched_config2(steps=6, warmup=1, lanes=8):
    """The uneven multi-view workload: 4 scenes of 50k/150k/300k/500k Gaussians
    (seeds 0-3, SH degree 3), 64 views with resolutions drawn uniformly (seed 4)
    from {256^2, 512^2, 1024x768, 1920x1080} on radius-3 rings at random angles
    and heights, targets = the forward image of a second scene per scene (seed
    100 + s), quantised to k/255.  One step = all 64 views fwd+bwd + Adam per
    scene.  Returns per-step ms of the captured graph and of view-after-view
    launches (strictly serial, and overlapping across scenes)."""
    from nprush_amd import GpuMultiViewStep, GpuSplatOptimizerManager, camera, scenes

    sizes = scenes.CONFIG2_SIZES
    counts = scenes.CONFIG2_COUNTS
    views = scenes.config2_views()
    targets = []
    for sidx, n in enumerate(counts):
        ts, tsh = scenes.make_ext_splats(n, 100 + sidx, sh_degree=3)
        tm = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=3, splats=ts, sh=tsh, adc_enabled=False)
        for (sc, w, h, vp, vi) in views:
            if sc != sidx:
                continue
            tm.setBackwardTarget(np.zeros((h, w, 4), np.float32), np.ones((h, w), np.float32), w, h)
            tm.writeSplatVPMatrix(vp, vi)
            tm.forward()
            c, _ = tm.read_image()
            rgba = np.ones((h, w, 4), np.float32)
            rgba[..., :3] = scenes.quantize8(c)
            targets.append(rgba)
        tm.close()
    tgt_of = {}
    it = iter(targets)
    for sidx in range(4):
        for k, v in enumerate(views):
            if v[0] == sidx:
                tgt_of[k] = next(it)
    out = {}
    for graph, tuned, serial in ((False, False, True), (False, False, False), (True, False, False),
                                 (True, True, False)):
        key = ("graph" if graph else "per_view") + ("_tuned" if tuned else "") + ("_serial" if serial else "")
        mgrs = []
        for sidx, n in enumerate(counts):
            s, sh = scenes.make_ext_splats(n, sidx, sh_degree=3)
            mgrs.append(GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=3, splats=s, sh=sh,
                                                 adc_enabled=False))
        step = GpuMultiViewStep(mgrs, lanes=lanes)
        for k, (sc, w, h, vp, vi) in enumerate(views):
            step.add_view(sc, vp, vi, tgt_of[k])
        if tuned:
            tune = step.autotune()
            out["tune"] = {f"{sz[0]}x{sz[1]}": [int(x) for x in tune[[i for i, v in enumerate(views)
                                                                       if (v[1], v[2]) == sz][0]]]
                           for sz in sizes}
            out["tune_in_context"] = step.graph_info()
        for _ in range(warmup):
            step.step(graph=graph, serial=serial)
        t0 = time.perf_counter()
        for _ in range(steps):
            step.step(graph=graph, serial=serial)
        out[key] = (time.perf_counter() - t0) * 1000.0 / steps
        if graph:
            out["nodes_" + key] = step.graph_info()["nodes"]
        losses = step.read_losses()
        out["l1_" + key] = float(losses[:, 1].sum())
        step.close()
        for m in mgrs:
            m.close()
    return {
        "workload": "4 scenes 50k/150k/300k/500k SH3, 64 views of 256^2/512^2/1024x768/1920x1080 (seed 4)",
        "lanes": lanes,
        # per-view launches, strictly sequential (each view completes before the next)
        "per_view_serial_ms_per_step": round(out["per_view_serial"], 3),
        # per-view launches on the scenes' streams (views of different scenes overlap)
        "per_view_ms_per_step": round(out["per_view"], 3),
        "graph_ms_per_step": round(out["graph"], 3),
        "graph_autotuned_ms_per_step": round(out["graph_tuned"], 3),
        "autotuned_tile_batch_by_resolution (scene 0)": out["tune"],
        # npt_autotune's second stage: whole-step timing graphs (no chain rule, no Adam) for
        # {scene defaults, per-class choices} x {all lanes, half}; the fastest is kept
        "autotune_in_context_ms": out["tune_in_context"]["autotune_ms"],
        "autotuned_lanes": out["tune_in_context"]["lanes"],
        "graph_nodes": out["nodes_graph"],
        "graph_nodes_autotuned": out["nodes_graph_tuned"],
        "speedup_graph_autotuned_vs_serial": round(out["per_view_serial"] / out["graph_tuned"], 3),
        "speedup_graph": round(out["per_view"] / out["graph"], 3),
        "speedup_graph_autotuned": round(out["per_view"] / out["graph_tuned"], 3),
        "same_losses": out["l1_graph"] == out["l1_per_view"] == out["l1_per_view_serial"],
    }


def host_info(oracle_lib) -> dict:
    """The host CPU the CPU baselines ran on: model name, OpenMP threads used
    and the OMP_NUM_THREADS setting they came from."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cores": int(oracle_lib.ora_num_threads()), "cpu_model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "host_cpus": os.cpu_count()}


def cpu_baseline_ext(s, sh, deg, vp, vi, target, W, H, sample_tiles=0, seed=0):
    """The EXT oracle (FP32, OpenMP on all host cores) on one whole view:
    projection, binning in O(keys) with the per-tile depth sort, forward + L1 +
    backward over every tile, and the per-splat chain rule."""
    from oracle import oracle

    oracle.load()
    t0 = time.perf_counter()
    offs, ids = oracle.ext_bin_all(s, sh, deg, vp, vi, W, H)
    oracle.ext_backward_csr_f32(s, sh, deg, vp, vi, target, W, H, offs, ids)
    dt = time.perf_counter() - t0
    return {
        "value": round(dt * 1000.0, 1),
        "unit": "ms/view (fwd+bwd)",
        **host_info(oracle.load()),
        "kind": "port",
        "sample": f"one whole view (view 0, {len(ids)} tile keys), binning + fwd + L1 + bwd + chain rule",
    }


def cpu_baseline(splats, vp, vi, target, depth, W, H, sample_tiles, seed=0):
    """The oracle (FP32 REF restatement, faithful O(tiles x N) binning) on a
    random sample of tiles of one view, scaled to the whole view."""
    from oracle import oracle

    oracle.load()
    ntiles = ((W + 15) // 16) * ((H + 15) // 16)
    rng = np.random.default_rng(seed)
    tiles = np.sort(rng.choice(ntiles, size=min(sample_tiles, ntiles), replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    oracle.backward_f32(splats, vp, vi, target, depth, W, H, tiles=tiles)
    dt = time.perf_counter() - t0
    ms_view = dt * 1000.0 * ntiles / len(tiles)
    return {
        "value": round(ms_view, 1),
        "unit": "ms/view (fwd+bwd)",
        **host_info(oracle.load()),
        "kind": "port",
        "sample": f"{len(tiles)} random tiles of view 0 ({dt:.1f} s), scaled by {ntiles}/{len(tiles)} tiles",
    }


def bezier_config3(steps=20, warmup=3, cpu_tiles=16, n=100_000, W=1024, H=1024, views=4, small=False):
    """Bezier stroke optimizer (GpuBezierOptimizerManager.dispatch: backward +
    Adam, densification every 50 steps): 100,000 curves (p0 uniform in
    [-1, 1]^3, random-walk control points with N(0, 0.05) steps, width
    log-normal(ln 0.004, 0.3), softness width / 4, opacity sigmoid(N(1, 0.5)),
    RGB uniform), 4 ring views at 1024 x 1024, colour mode, targets = the
    forward image and depth of a seed-1 curve set, black background at depth 1.
    Targets stay in HBM (device buffers of the handle) for the whole run."""
    from nprush_amd import GpuBezierOptimizerManager, camera, scenes

    if small:
        n, W, H = 10_000, 256, 256
    cams = camera.ring_cameras(views, 3.0, W, H)
    curves = scenes.make_bezier_curves(n, 0)
    tcurves = scenes.make_bezier_curves(n, 1)
    m = GpuBezierOptimizerManager(numBeziers=n)
    m.writeMode(1.0)
    bufs = []

    def resident(a):
        ptr = m.dev_alloc(a.nbytes)
        m.dev_upload(ptr, a)
        bufs.append(ptr)
        return ptr
    zeros4 = resident(np.zeros((H, W, 4), np.float32))
    ones1 = resident(np.ones((H, W), np.float32))
    m.write_beziers(tcurves)
    targets, host_targets = [], []
    for vp, vi in cams:
        m.writeVPMatrix(vp)
        m.writeVPInvMatrix(vi)
        m.set_backward_target_dev(zeros4, ones1, zeros4, ones1, zeros4, W, H)
        m.forward()
        C, D, _ = m.read_image()
        t = np.zeros((H, W, 4), np.float32)
        t[..., :3] = C
        t[..., 3] = 1.0
        D = np.ascontiguousarray(D, np.float32)
        host_targets.append((t, D))
        targets.append((resident(t), resident(D)))
    m.write_beziers(curves)
    m.enable_timing(True)

    def one(k):
        vp, vi = cams[k % views]
        tg, td = targets[k % views]
        m.writeVPMatrix(vp)
        m.writeVPInvMatrix(vi)
        m.set_backward_target_dev(tg, td, zeros4, ones1, zeros4, W, H)
        m.dispatch()
        return m.read_timing(), m.read_stats()

    for k in range(warmup):
        one(k)
    m.sync()
    rows = []
    t0 = time.perf_counter()
    for k in range(steps):
        rows.append(one(warmup + k))
    m.sync()
    wall = time.perf_counter() - t0
    # image only (the composite's forward half) on the current state, view 0
    vp, vi = cams[0]
    tg, td = targets[0]
    m.writeVPMatrix(vp)
    m.writeVPInvMatrix(vi)
    m.set_backward_target_dev(tg, td, zeros4, ones1, zeros4, W, H)
    fwd_ms = []
    for _ in range(3):
        m.forward()
        fwd_ms.append(m.read_timing()["composite"])
    m.adc()  # one densification pass (every 50th dispatch), timed on its own
    adc_ms = m.read_timing()["adc"]
    st_adc = m.read_stats()
    # display raster of the trained curves (GpuBezierForwardPipelineManager.dispatch,
    # 8-bit target), view 0, synthetic 256 x 256 brush; wall clock per dispatch,
    # the key-count readback included
    from nprush_amd import GpuBezierForwardPipelineManager

    fm = GpuBezierForwardPipelineManager(numBeziers=n, bezierBuffer=m, brushTexture=scenes.make_brush())
    fm.writeVPMatrix(vp)
    fm.setTarget(None, W, H)
    for _ in range(2):
        fm.dispatch(clear=True, quantize=True)
    m.sync()
    t1 = time.perf_counter()
    for _ in range(5):
        fm.dispatch(clear=True, quantize=True)
    m.sync()
    draw_ms = (time.perf_counter() - t1) * 1000.0 / 5
    fm.close()

    def mean(key, src=0):
        return float(np.mean([r[src][key] for r in rows]))
    bwd, comp, step = mean("backward"), mean("composite"), mean("step")
    fwd_eval, rev_eval = mean("evaluated_wave_entries", 1), mean("reverse_evaluated_wave_entries", 1)
    rev_pairs = mean("reverse_pairs", 1)
    # algorithmic work: every (pixel, curve) pair of the capped tile lists is
    # evaluated once (the reference's forward), every active pair once in reverse
    pairs = mean("pairs", 1)
    flops = pairs * BZ_FLOP_PER_EVAL + rev_pairs * BZ_FLOP_PER_REV_PAIR
    achieved = flops / (comp * 1e-3) / 1e12
    # Adam: 18 parameters x (grad i32, m, v) + 20-float curve + grad_accum, read and written
    adam_gbs = (18 * 12 + 80 + 4) * 2.0 * n / (step * 1e-3) / 1e9
    # HBM bytes per view of forward + reverse from the committed FETCH_SIZE /
    # WRITE_SIZE runs of this workload (calibrated against a streaming kernel)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and not small:
        with open(tpath) as fh:
            tj = json.load(fh).get("k_bz_composite")
        if tj:
            traffic = int(tj["fetch_bytes"] + tj["write_bytes"])
    cpu = None
    if cpu_tiles > 0:
        from oracle import oracle

        oracle.load()
        vp, vi = cams[0]
        tg, td = host_targets[0]
        uni = oracle.bz_uniforms(vp, vi, mode=1.0)
        tex = (tg, td, np.zeros((H, W, 4), np.float32), np.ones((H, W), np.float32),
               np.zeros((H, W, 4), np.float32))
        ntiles = ((W + 15) // 16) * ((H + 15) // 16)
        rng = np.random.default_rng(0)
        tiles = np.sort(rng.choice(ntiles, size=min(cpu_tiles, ntiles), replace=False)).astype(np.int32)
        c1 = time.perf_counter()
        lists, cnt, _ = oracle.bz_bin(curves, uni, W, H, tiles)
        oracle.bz_backward_f32(curves, uni, *tex, W, H, tiles=tiles, lists=lists, counts=cnt)
        dt = time.perf_counter() - c1
        cpu = {"value": round(dt * 1000.0 * ntiles / len(tiles), 1), "unit": "ms/view (backward)",
               **host_info(oracle.load()), "kind": "port",
               "sample": f"{len(tiles)} random tiles of view 0 ({dt:.1f} s): binning + fwd + loss + bwd, "
                         f"scaled by {ntiles}/{len(tiles)} tiles"}
    for ptr in bufs:
        m.dev_free(ptr)
    m.close()
    return {
        "workload": f"Bezier strokes: {n} curves, {views} ring views at {W}x{H}, colour mode, dispatch = "
                    "backward + Adam (densification every 50 steps)",
        "backward_ms_per_view": round(bwd, 3),
        "step_ms": round(step, 3),
        "ms_per_step_wall": round(wall * 1000.0 / steps, 3),
        "steps_per_s": round(steps / wall, 2),
        "densification_ms": round(adc_ms, 3),
        "densification_ms_per_step": round(adc_ms / 50.0, 3),
        "display_raster_ms": round(draw_ms, 3),
        "phases_ms": {k: round(mean(k), 3) for k in ("preprocess", "keys", "sort", "composite", "step")},
        "tile_keys_per_view": int(mean("keys", 1)),
        "pairs_per_view": int(mean("pairs", 1)),
        "evaluated_pair_fraction": round(64.0 * fwd_eval / max(1.0, mean("pairs", 1)), 4),
        "reverse_recompute_pair_fraction": round(64.0 * rev_eval / max(1.0, mean("pairs", 1)), 4),
        "forward_only_composite_ms": round(float(np.median(fwd_ms)), 3),
        "reverse_pairs_per_view": int(rev_pairs),
        "roofline": {"bound": "valu", "compute": "fp32 VALU issue (no matrix work on this path)",
                     "kernel": "k_bz_forward + k_bz_reverse", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic},
        "adam_hbm_gbs": round(adam_gbs, 1),
        "adc_stats": {k: st_adc[k] for k in ("adc_dead", "adc_killed", "adc_split1", "adc_spawned", "adc_split3")},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(cpu["value"] / bwd, 1) if cpu else None,
    }


MLP_FWD_FLOP = 2 * (32 * 256 + 2 * 256 * 256 + 256 * 14)  # 285,696 per sample
PEAK_BF16_TFLOPS = 2500.0


def ext_training_scene(n_live, W, H, V, seed, tseed):
    """EXT training scene of the scale-up configuration's distribution: n_live
    Gaussians (seed) plus as many dead slots (the pool densification fills), SH3,
    V ring views at radius 3 plus a held-out view at the half-step ring angle
    0.4 above the ring; targets = the forward images of a tseed scene,
    quantised to k/255.  Returns (manager, views, held-out view, target slots)."""
    import math

    from nprush_amd import GpuSplatOptimizerManager, camera, scenes

    cap = 2 * n_live
    cams = camera.ring_cameras(V, 3.0, W, H)
    ang = math.pi / V
    held = camera.view_proj((3.0 * math.cos(ang), 0.4, 3.0 * math.sin(ang)), aspect=W / H)
    allc = list(cams) + [held]
    ts, tsh = scenes.make_ext_splats(n_live, tseed, sh_degree=3)
    tm = GpuSplatOptimizerManager(numSplats=n_live, mode="ext", sh_degree=3, splats=ts, sh=tsh, adc_enabled=False)
    tm.setBackwardTarget(np.zeros((H, W, 4), np.float32), np.ones((H, W), np.float32), W, H)
    tg = []
    for vp, vi in allc:
        tm.writeSplatVPMatrix(vp, vi)
        tm.forward()
        c, _ = tm.read_image()
        r = np.ones((H, W, 4), np.float32)
        r[..., :3] = scenes.quantize8(c)
        tg.append(r)
    tm.close()
    s, sh = scenes.make_ext_splats(cap, seed, sh_degree=3)
    s[n_live:, 3] = -10000.0
    m = GpuSplatOptimizerManager(numSplats=cap, mode="ext", sh_degree=3, splats=s, sh=sh, adc_enabled=False)
    m.set_adc_ext(2e-4, 3.3, 0.005, 0.01)
    depth = np.ones((H, W), np.float32)
    slots = [m.upload_target(t, depth) for t in tg]
    views = [(allc[i][0], allc[i][1], slots[i]) for i in range(V)]
    return m, views, (allc[-1][0], allc[-1][1], slots[-1]), slots


# Algorithmic HBM bytes of one densification.
#   REF (adc.wgsl): every slot's opacity and accumulator read, accumulator
#     reset (12 B); every child: parent record read (64), two records written
#     (128), the m/v of both zeroed (256).
#   EXT (k_adc_ext_*): per slot 68 B (opacity 4 + accumulator 8 read, dead and
#     qualifier flags 8 written; two exclusive scans 2 x 12; list pass 16;
#     accumulator reset 8); per placement 968 B (list entries 8, parent record
#     64, SH copied 192 + 192, the dead-slot child's m/v zeroed 512) plus 640
#     for a split (both children written 128, the parent slot's m/v zeroed 512)
#     or 64 for a clone (the child written); per dead slot left unfilled 548 B
#     (list entry 4, record word read and written 32, m/v zeroed 512).
def adc_bytes_ref(n, children):
    return 12 * n + 448 * children


def adc_bytes_ext(n, placed, splits, clones, dead):
    return 68 * n + 968 * placed + 640 * splits + 64 * clones + 548 * max(dead - placed, 0)


def adc_ext_leg(small=False, warm=100):
    """One EXT densification at the scale-up configuration's size (1M live +
    1M dead Gaussians, SH3, 1080p, 8 ring views): `warm` optimizer steps, a
    first densification, `warm` more steps to refill the gradient
    accumulators, then the timed one; HIP-event time and algorithmic bytes
    (adc_bytes_ext)."""
    from nprush_amd import densify

    n_live, W, H, V = (40_000, 480, 272, 8) if small else (1_000_000, 1920, 1080, 8)
    m, views, _, slots = ext_training_scene(n_live, W, H, V, 0, 1)
    # a first densification loads the kernels and sizes the buffers; the second is timed
    densify.train_steps(m, views, 0, warm)
    m.adc()
    densify.train_steps(m, views, warm, warm)
    m.enable_timing(True)
    m.adc()
    ms = m.read_timing()["adc"]
    st = m.read_stats()
    n = 2 * n_live
    # EXT stats: adc_children = placements, adc_rounds = dead slots before the pass
    byts = adc_bytes_ext(n, st["adc_children"], st["adc_splits"], st["adc_clones"], st["adc_rounds"])
    for sl in slots:
        m.free_target(sl)
    m.close()
    return {"workload": f"EXT densification, {n_live} live + {n_live} dead Gaussians SH3, the second one, {warm} steps "
                        f"at {W}x{H}, threshold 2e-4, split above 0.01 x extent 3.3, prune below 0.005",
            "ms": round(ms, 4), "slots": n, "qualifiers": st["adc_qualifiers"], "placed": st["adc_children"],
            "splits": st["adc_splits"], "clones": st["adc_clones"], "dead_slots": st["adc_rounds"],
            "algorithmic_bytes": int(byts), "hbm_gbs": round(byts / (ms * 1e-3) / 1e9, 1),
            "roofline_frac": round(byts / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 3)}



def learned_config4(small=False, events=3, warm=100, window=200, eval_steps=3000, adc_period=100, epochs=10,
                    record_seeds=((0, 1), (4, 5), (6, 7)), eval_every=100):
    """Learned densification on the EXT model (no reference counterpart).
    Pairs are recorded from plain densification on scale-up scenes (1M live
    Gaussians plus 1M dead slots, SH3, 1080p, 8 ring views; (splat seed, target
    seed) in record_seeds, about 0.76M pairs each): `events` times, `warm`
    steps, the parents' features, densification, `window` steps without
    densification, the children's movement.  The MLP (32 -> 256 -> 256 -> 256
    -> 14, GELU, BF16 matrix cores) is trained on them with the project's
    kernels; then a fresh scene (seeds 2 / 3) is trained for eval_steps (the
    scale-up configuration's 3000) with plain and with learned densification
    every adc_period steps (held-out view at the half-step ring angle, 0.4 above
    the ring, as the judge) and the training time to reach plain's final L1 is
    compared."""
    from nprush_amd import densify

    n_live, W, H, V = (40_000, 480, 272, 8) if small else (1_000_000, 1920, 1080, 8)
    cap = 2 * n_live

    def scene(seed, tseed):
        return ext_training_scene(n_live, W, H, V, seed, tseed)

    t0 = time.perf_counter()
    Xs, Ts = [], []
    for seed, tseed in (record_seeds[:1] if small else record_seeds):
        m, views, _, slots = scene(seed, tseed)
        x, t = densify.record_pairs(m, views, events=events, warm=warm, window=window)
        Xs.append(x)
        Ts.append(t)
        for sl in slots:
            m.free_target(sl)
        m.close()
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    rec_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    net, losses, (held_before, held_after) = densify.train_mlp(X, T, epochs=epochs, batch=4096)
    train_s = time.perf_counter() - t1
    res = {}
    decisions = None
    mlp_fwd = None
    for name in ("plain", "learned"):
        m, views, held, slots = scene(2, 3)
        judge = densify.LearnedDensifier(net, held) if name == "learned" else None
        res[name] = densify.run(m, views, eval_steps, adc_period, judge, eval_every=eval_every)
        if judge is not None:
            decisions = [(d, int(k), round(lh, 1), round(lm, 1)) for d, k, lh, lm in judge.decisions]
            judge.close()
            # MLP inference on the device: the recorded features, timed with the handle's sync
            nx = min(len(X), 1 << 20)
            if nx > 0:
                xp = m.dev_alloc(128 * nx)
                yp = m.dev_alloc(56 * nx)
                m.dev_upload(xp, X[:nx].astype(np.float32))
                net.forward_dev(xp, nx, yp)
                net.enable_timing(True)
                ms = []
                for _ in range(5):
                    net.forward_dev(xp, nx, yp)
                    ms.append(net.read_timing()["forward"])
                net.enable_timing(False)
                dt = float(np.mean(ms)) * 1e-3
                tf = nx * MLP_FWD_FLOP / dt / 1e12
                mlp_fwd = {"samples": nx, "ms": round(dt * 1e3, 3),
                           "roofline": {"bound": "mfma", "compute": "bf16 matrix cores, fp32 accumulation",
                                        "kernel": "k_mlp_infer", "achieved": round(tf, 2),
                                        "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                                        "frac": round(tf / PEAK_BF16_TFLOPS, 4), "traffic": None}}
                m.dev_free(xp)
                m.dev_free(yp)
        for sl in slots:
            m.free_target(sl)
        m.close()
    net.close()
    target = res["plain"][-1][1]
    t_plain = res["plain"][-1][0]
    t_learned = densify.time_to_target(res["learned"], target)
    return {
        "workload": f"EXT {n_live} live + {cap - n_live} dead Gaussians SH3, {W}x{H}, {V} ring views + 1 held-out; "
                    f"pairs from {events} plain densifications ({warm}-step lead-in, {window}-step window) on "
                    f"{1 if small else len(record_seeds)} recording scene(s); "
                    f"{eval_steps} steps, densification every {adc_period}",
        "pairs": int(len(X)),
        "eval_steps": eval_steps,
        "record_s": round(rec_s, 2),
        "mlp_train_s": round(train_s, 2),
        "mlp_epochs": epochs,
        "mlp_train_loss_first_last": [round(losses[0], 4), round(losses[-1], 4)] if losses else None,
        "mlp_heldout_loss_untrained_trained": [round(held_before, 4), round(held_after, 4)],
        "mlp_forward": mlp_fwd,
        "plain_curve": [(round(t, 3), round(l, 1)) for t, l in res["plain"]],
        "learned_curve": [(round(t, 3), round(l, 1)) for t, l in res["learned"]],
        "judge_decisions": decisions,
        "target_l1": round(target, 1),
        "plain_train_s": round(t_plain, 3),
        "learned_time_to_target_s": None if t_learned is None else round(t_learned, 3),
        "time_to_target_ratio": None if t_learned is None else round(t_learned / t_plain, 3),
        "final_l1_plain_learned": [round(res["plain"][-1][1], 1), round(res["learned"][-1][1], 1)],
    }


def mlp_bench(n=757_052, steps=20, warmup=3, batch=4096, train_steps=20):
    """The densification MLP alone: inference over n synthetic parent-feature rows
    (the pair count of one learned-densification recording scene) resident in HBM, and
    Adam training steps at the recording's batch size; HIP-event times."""
    from nprush_amd import DensifierMLP, densify

    rng = np.random.default_rng(0)
    x = rng.normal(0.0, 1.0, (n, 32)).astype(np.float32)
    t = rng.normal(0.0, 1.0, (batch, 14)).astype(np.float32)
    net = DensifierMLP(max_batch=batch, params=densify.init_params(0, out_scale=0.5))
    xp = net.dev_alloc(x.nbytes)
    yp = net.dev_alloc(56 * n)
    tp = net.dev_alloc(t.nbytes)
    net.dev_upload(xp, x)
    net.dev_upload(tp, t)
    net.enable_timing(True)
    for _ in range(warmup):
        net.forward_dev(xp, n, yp)
    fwd = []
    c0 = time.perf_counter()
    for _ in range(steps):
        net.forward_dev(xp, n, yp)
        fwd.append(net.read_timing()["forward"])
    wall = time.perf_counter() - c0
    for _ in range(warmup):
        net.train_dev(xp, tp, batch, 1e-4)
    tr = []
    for _ in range(train_steps):
        net.train_dev(xp, tp, batch, 1e-4)
        tr.append(net.read_timing()["train_step"])
    for p_ in (xp, yp, tp):
        net.dev_free(p_)
    net.close()
    fwd_ms = float(np.mean(fwd))
    tr_ms = float(np.mean(tr))
    # HBM bytes per inference launch from the committed FETCH_SIZE / WRITE_SIZE runs
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh).get("k_mlp_infer")
        if tj and tj.get("rows") == n:
            traffic = int(tj["fetch_bytes"] + tj["write_bytes"])
    tf = n * MLP_FWD_FLOP / (fwd_ms * 1e-3) / 1e12
    tf_train = batch * 3 * MLP_FWD_FLOP / (tr_ms * 1e-3) / 1e12
    return {
        "workload": f"densification MLP 32-256-256-256-14 GELU, bf16 x bf16 -> fp32: inference over {n} rows, "
                    f"Adam training steps at batch {batch}",
        "rows": n, "forward_ms": round(fwd_ms, 4), "forward_wall_ms": round(wall * 1e3 / steps, 4),
        "rows_per_s": round(n / (fwd_ms * 1e-3), 1),
        "train_step_ms": round(tr_ms, 4), "train_batch": batch, "train_tflops": round(tf_train, 2),
        "roofline": {"bound": "mfma", "compute": "bf16 matrix cores, fp32 accumulation", "kernel": "k_mlp_infer",
                     "achieved": round(tf, 2), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tf / PEAK_BF16_TFLOPS, 4), "traffic": traffic},
    }


def pt_target_leg(frames=20, small=False, cpu=True):
    """The frame's target generators on a generated scene: the mesh raster (colour
    and encoded depth over the environment) and one path-traced sample per pixel
    at the optimisation resolution of a 16:9 panel (455 x 256), an icosphere of
    81,920 triangles + mirror box + ground (81,964 triangles)."""
    import time as _t

    from nprush_amd import camera, meshes
    from nprush_amd.managers import GpuMeshRenderPipelineManager, GpuPathTracePipelineManager

    W, H = (128, 72) if small else (455, 256)
    ms = meshes.scene(subdiv=3 if small else 6, seed=1)
    env = meshes.environment(512, 256)
    ntri = int(sum(len(i) for _, i in ms))
    eye, tgt, up = (3.0, -2.5, 1.8), (0.2, 0.0, 0.6), (0.0, 0.0, 1.0)
    vp, vi = camera.view_proj(eye, target=tgt, up=up, aspect=W / H)
    view = camera.look_at(eye, tgt, up)
    pt = GpuPathTracePipelineManager(envTexture=env)
    t0 = _t.perf_counter()
    pt.setMeshes(ms)
    build_ms = (_t.perf_counter() - t0) * 1e3
    pt.writeInvViewProjMat(vi)
    pt.setOutputSize(W, H)
    for _ in range(3):
        pt.dispatch()
    pt.read_accum()
    t0 = _t.perf_counter()
    for _ in range(frames):
        pt.dispatch()
    acc = pt.read_accum()
    pt_ms = (_t.perf_counter() - t0) * 1e3 / frames
    mr = GpuMeshRenderPipelineManager(mesh=ms[0], matcapTexture=env)
    mr.setGroundMesh(ms[2])
    mr.writeUniforms(vp, view, vi, 1.0)
    mr.addDraw(None, None, W, H)
    t0 = _t.perf_counter()
    for _ in range(frames):
        mr.addDraw(None, None, W, H)
    raster_ms = (_t.perf_counter() - t0) * 1e3 / frames
    mr.close()
    pt.close()
    out = {"workload": f"path-traced target: {ntri} triangles (icosphere + mirror box + ground), 1 sample per pixel "
                       f"and frame, 8 bounces, at {W}x{H}; mesh raster target of the same scene",
           "width": W, "height": H, "triangles": ntri, "bvh_build_ms_host": round(build_ms, 1),
           "pt_ms_per_frame": round(pt_ms, 3), "paths_per_s": round(W * H / (pt_ms * 1e-3), 1),
           "raster_ms_incl_readback": round(raster_ms, 3), "mean_samples": float(acc[..., 3].mean()),
           "cpu_baseline": None}
    if cpu:
        from oracle import oracle

        v = np.concatenate([m[0] for m in ms])
        offs = np.cumsum([0] + [len(m[0]) for m in ms[:-1]])
        idx = np.concatenate([m[1] + np.uint32(o) for m, o in zip(ms, offs)])
        nodes, tris = oracle.pt_bvh(v, idx)
        t0 = _t.perf_counter()
        oracle.pt_trace(vi, 0, W, H, v, nodes, tris, env)
        dt = _t.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms/frame", "cores": int(oracle.load().ora_num_threads()),
                               "kind": "port", "sample": f"one whole frame at {W}x{H} (oracle, OpenMP)"}
        out["speedup_vs_cpu"] = round(dt * 1e3 / pt_ms, 1)
    return out


def main_mlp(args):
    print("[bench] densification MLP ...", file=sys.stderr, flush=True)
    r = mlp_bench(steps=args.steps, warmup=args.warmup, n=40_000 if args.small else 757_052)
    out = {
        "metric": "MLP inference rows/s", "value": r["rows_per_s"], "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["forward_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic",
        "config": {"workload": r["workload"]}, "roofline": r["roofline"], "cpu_baseline": None, "mlp": r,
    }
    print(json.dumps(out), flush=True)


def main_learned(args):
    print("[bench] learned densification ...", file=sys.stderr, flush=True)
    r = learned_config4(small=args.small)
    out = {
        "metric": "time-to-target-L1 vs plain ADC",
        "value": r["time_to_target_ratio"],
        "unit": "ratio (learned / plain training time to plain's final L1)",
        "n_gpus": 1,
        "steps": r["eval_steps"],
        "warmup": 0,
        "ms_per_step": round(r["plain_train_s"] * 1000.0 / r["eval_steps"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (MLP bf16 x bf16 -> fp32)",
        "data": "synthetic (seeded scenes, targets rendered from other seeds)",
        "config": {"workload": r["workload"]},
        "roofline": (r["mlp_forward"] or {}).get("roofline"),
        "cpu_baseline": None,
        "learned": r,
    }
    print(json.dumps(out), flush=True)


def main_bezier(args):
    print("[bench] Bezier strokes: 100k curves, 4 views at 1024x1024 ...", file=sys.stderr, flush=True)
    r = bezier_config3(steps=args.steps, warmup=args.warmup, cpu_tiles=args.cpu_tiles, small=args.small)
    out = {
        "metric": "Bezier backward ms/view @100k curves 1024x1024",
        "value": r["backward_ms_per_view"],
        "unit": "ms/view",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step_wall"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded curves, targets rendered from a seed-1 curve set)",
        "config": {"workload": r["workload"]},
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu_baseline"],
        "bezier": r,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--small", action="store_true", help="16,384 splats at 256x256 (quick check)")
    ap.add_argument("--mode", choices=("ext", "ref", "bezier", "learned", "mlp"), default="ext")
    ap.add_argument("--no-batched", action="store_true", help="skip the multi-view (64 views, 4 scenes) comparison")
    ap.add_argument("--no-bezier", action="store_true", help="skip the Bezier stroke workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-densification comparison")
    ap.add_argument("--no-ref", action="store_true", help="skip the REF-semantics leg (1M live / 2M slots, 1080p)")
    ap.add_argument("--no-densify", action="store_true", help="skip the timed EXT densification pass")
    ap.add_argument("--no-pt", action="store_true", help="skip the path-traced / raster target leg")
    ap.add_argument("--cpu-tiles", type=int, default=48, help="tiles in the CPU-baseline sample (0 = skip)")
    ap.add_argument("--tile", type=int, default=0,
                    help="EXT: composite tile edge {8,16,32} instead of the autotuner's choice (with --batch)")
    ap.add_argument("--batch", type=int, default=0, help="EXT: composite LDS batch {32,64,128} (with --tile)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one plain streaming kernel (k_debug_math, 16M floats: 64 MiB read + 64 MiB "
                         "written) to calibrate profiler byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is described in "
              "DESIGN.md and is not part of this build.", file=sys.stderr)
        sys.exit(2)

    if args.mode == "ext":
        return main_ext(args)
    if args.mode == "bezier":
        return main_bezier(args)
    if args.mode == "learned":
        return main_learned(args)
    if args.mode == "mlp":
        return main_mlp(args)

    out = ref_leg(args.steps, args.warmup, args.cpu_tiles, small=args.small, calibrate=args.calibrate)
    print(json.dumps(out), flush=True)


def ref_leg(steps, warmup, cpu_tiles, small=False, calibrate=False):
    """The reference-semantics (REF) path at its own headline workload: 1M live
    splats in 2M slots, 1920x1080, 8 ring views; dispatch = backward + Adam.
    Returns the full bench record of that run (the --mode ref line)."""
    from nprush_amd import GpuSplatOptimizerManager

    if small:
        n_live, capacity, W, H, views = 16384, 32768, 256, 256, 8
    else:
        n_live, capacity, W, H, views = 1_000_000, 2_000_000, 1920, 1080, 8
    print(f"[bench] building REF workload: {n_live} live / {capacity} capacity, {W}x{H}, {views} views", file=sys.stderr,
          flush=True)
    cams, targets, depth, splats = build_workload(n_live, capacity, W, H, views, 3.0)

    mgr = GpuSplatOptimizerManager(numSplats=capacity, splats=splats)
    slots = [mgr.upload_target(t, depth) for t in targets]
    mgr.enable_timing(True)

    def one_step(k):
        vp, vi = cams[k % views]
        mgr.writeSplatVPMatrix(vp, vi)
        mgr.set_target_slot(slots[k % views])
        mgr.dispatch()
        t = mgr.read_timing()
        st = mgr.read_stats()
        return t, st

    for k in range(warmup):
        one_step(k)
    mgr.sync()
    print(f"[bench] REF warmup done ({warmup} steps)", file=sys.stderr, flush=True)
    rows = []
    t0 = time.perf_counter()
    for k in range(steps):
        rows.append(one_step(warmup + k))
    mgr.sync()
    wall = time.perf_counter() - t0
    print(f"[bench] REF timed {steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)

    bwd_ms = float(np.mean([t["backward"] for t, _ in rows]))
    comp_ms = float(np.mean([t["composite"] for t, _ in rows]))
    step_ms = float(np.mean([t["step"] for t, _ in rows]))
    pairs = float(np.mean([s["pairs"] for _, s in rows]))
    active = float(np.mean([s["active_pairs"] for _, s in rows]))
    flops = pairs * FWD_FLOP_PER_PAIR + active * BWD_FLOP_PER_ACTIVE_PAIR
    achieved = flops / (comp_ms * 1e-3) / 1e12
    phases = {k: round(float(np.mean([t[k] for t, _ in rows])), 3) for k in
              ("preprocess_keys", "sort", "ranges", "composite", "step")}
    adam_gbs = 520.0 * capacity / (step_ms * 1e-3) / 1e9
    keys = float(np.mean([s["keys"] for _, s in rows]))
    # HBM bytes per composite launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs of this workload (calibrated against a streaming kernel)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and not small:
        with open(tpath) as fh:
            tj = json.load(fh).get("k_composite_ref")
        if tj:
            traffic = int(tj["fetch_bytes"] + tj["write_bytes"])

    # reference ADC (adc.wgsl): a first pass loads the kernels and sizes the
    # buffers; after 25 more steps refill the accumulators the next one is timed
    mgr.adc()
    for k in range(25):
        vp, vi = cams[k % views]
        mgr.writeSplatVPMatrix(vp, vi)
        mgr.set_target_slot(slots[k % views])
        mgr.backward()
        mgr.step()
    mgr.adc()
    adc_ms = mgr.read_timing()["adc"]
    ast = mgr.read_stats()
    adc_b = adc_bytes_ref(capacity, ast["adc_children"])
    adc = {"ms": round(adc_ms, 4), "slots": capacity, "qualifiers": ast["adc_qualifiers"],
           "children": ast["adc_children"], "splits": ast["adc_splits"], "clones": ast["adc_clones"],
           "algorithmic_bytes": int(adc_b), "hbm_gbs": round(adc_b / (adc_ms * 1e-3) / 1e9, 1),
           "roofline_frac": round(adc_b / (adc_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 3)}

    if calibrate:
        from nprush_amd import debug_math

        debug_math(0, np.linspace(-8.0, 8.0, 1 << 24, dtype=np.float32))

    cpu = None
    if cpu_tiles > 0:
        print("[bench] REF CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        vp, vi = cams[0]
        cpu = cpu_baseline(splats, vp, vi, targets[0], depth, W, H, cpu_tiles)

    out = {
        "metric": "fwd+bwd ms/view @1M splats 1080p",
        "value": round(bwd_ms, 3),
        "unit": "ms/view",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(wall * 1000.0 / steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, targets rendered from a seed-1 scene)",
        "config": {"workload": f"REF splat fwd+bwd+Adam, {n_live} live / {capacity} capacity splats, {W}x{H}, "
                               f"{views} ring views, 16x16 tiles, cap 1024/tile",
                   "views": views, "width": W, "height": H, "splats": n_live, "capacity": capacity},
        "roofline": {"bound": "valu", "compute": "fp32 VALU issue (no matrix work on this path)", "kernel": "k_composite_ref", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic},
        "phases_ms": phases,
        "tile_keys_per_view": int(keys),
        "pairs_per_view": int(pairs),
        "active_pairs_per_view": int(active),
        "reverse_lane_efficiency": round(active / max(1.0, 64.0 * float(np.mean([s["reverse_wave_entries"] for _, s in rows]))), 3),
        "forward_eval_fraction": round(64.0 * float(np.mean([s["forward_wave_evals"] for _, s in rows])) / max(pairs, 1.0), 3),
        "adam_hbm_gbs": round(adam_gbs, 1),
        "adc": adc,
        "adam_roofline_frac": round(adam_gbs / PEAK_HBM_GBS, 3),
        "steps_per_s": round(steps / wall, 2),
        "cpu_baseline": cpu,
    }
    for s in slots:
        mgr.free_target(s)
    mgr.close()
    return out



def main_ext(args):
    from nprush_amd import GpuSplatOptimizerManager

    deg = 3
    if args.small:
        n, W, H, views = 16384, 256, 256, 8
    else:
        n, W, H, views = 1_000_000, 1920, 1080, 8
    print(f"[bench] building EXT workload: {n} Gaussians SH{deg}, {W}x{H}, {views} views", file=sys.stderr, flush=True)
    cams, targets, s, sh = build_workload_ext(n, W, H, views, 3.0, deg)
    depth = np.ones((H, W), np.float32)
    pinned = args.tile > 0 and args.batch > 0
    mgr = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=deg, splats=s, sh=sh, adc_period=100,
                                   adc_enabled=False, ext_tile=args.tile if pinned else 16,
                                   ext_batch=args.batch if pinned else 64)
    slots = [mgr.upload_target(t, depth) for t in targets]
    mgr.writeSplatVPMatrix(*cams[0])
    mgr.set_target_slot(slots[0])
    if pinned:
        tile, batch, tune_ms = args.tile, args.batch, None
        print(f"[bench] composite variant pinned: tile {tile}, batch {batch}", file=sys.stderr, flush=True)
    else:
        # composite tile size and LDS batch chosen by the autotuner on view 0
        tile, batch, tune_ms = mgr.autotune()
        print(f"[bench] autotune: tile {tile}, batch {batch}", file=sys.stderr, flush=True)
    mgr.enable_timing(True)

    def one_step(k):
        vp, vi = cams[k % views]
        mgr.writeSplatVPMatrix(vp, vi)
        mgr.set_target_slot(slots[k % views])
        mgr.dispatch()
        return mgr.read_timing(), mgr.read_stats()

    for k in range(args.warmup):
        one_step(k)
    mgr.sync()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    rows = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        rows.append(one_step(args.warmup + k))
    mgr.sync()
    wall = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)

    def mean(key, src=0):
        return float(np.mean([r[src][key] for r in rows]))

    bwd_ms, comp_ms, step_ms = mean("backward"), mean("composite"), mean("step")
    pairs, active = mean("pairs", 1), mean("active_pairs", 1)
    flops = pairs * EXT_FWD_FLOP_PER_PAIR + active * EXT_BWD_FLOP_PER_ACTIVE_PAIR
    achieved = flops / (comp_ms * 1e-3) / 1e12
    phases = {k: round(mean(k), 3) for k in ("preprocess_keys", "sort", "ranges", "composite", "chain", "step")}
    keys = mean("keys", 1)
    # binning (the "sort" phase): 4-byte items written once by the row
    # enumeration, read by the column count, read and written by the column
    # scatter = 16 bytes per tile key; the per-splat (depth, id) sort runs in
    # the preprocess phase and is not counted
    sort_bytes = 16.0 * keys
    # Adam: per splat 64 parameters x (param + grad + m + v), read and written = 2 KiB
    adam_gbs = 2048.0 * n / (step_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and not args.small:
        with open(tpath) as fh:
            tj = json.load(fh).get("k_composite_ext")
        if tj:
            traffic = int(tj["fetch_bytes"] + tj["write_bytes"])
    if args.calibrate:
        from nprush_amd import debug_math

        debug_math(0, np.linspace(-8.0, 8.0, 1 << 24, dtype=np.float32))
    batched = None
    if not args.small and not args.no_batched:
        print("[bench] multi-view step: graph vs per-view launches ...", file=sys.stderr, flush=True)
        batched = batched_config2()
    bezier = None
    if not args.small and not args.no_bezier:
        print("[bench] Bezier strokes: 100k curves, 4 views at 1024x1024 ...", file=sys.stderr, flush=True)
        bezier = bezier_config3(cpu_tiles=1024 if args.cpu_tiles > 0 else 0)
    adc = None
    if not args.no_densify:
        print("[bench] EXT densification ...", file=sys.stderr, flush=True)
        adc = adc_ext_leg(small=args.small)
    learned = None
    if not args.small and not args.no_learned:
        print("[bench] learned densification vs plain ...", file=sys.stderr, flush=True)
        learned = learned_config4()
        learned["mlp_alone"] = mlp_bench()
    ref = None
    if not args.small and not args.no_ref:
        print("[bench] REF path (reference semantics): 1M live / 2M slots, 1920x1080 ...", file=sys.stderr,
              flush=True)
        r = ref_leg(args.steps, args.warmup, args.cpu_tiles)
        ref = {k: r[k] for k in ("metric", "value", "unit", "steps", "warmup", "ms_per_step", "config", "roofline",
                                 "phases_ms", "tile_keys_per_view", "pairs_per_view", "active_pairs_per_view",
                                 "adam_hbm_gbs", "adam_roofline_frac", "adc", "cpu_baseline")}
        ref["speedup_vs_cpu"] = round(r["cpu_baseline"]["value"] / r["value"], 1) if r["cpu_baseline"] else None
    ptleg = None
    if not args.no_pt:
        print("[bench] path-traced and raster target ...", file=sys.stderr, flush=True)
        ptleg = pt_target_leg(small=args.small, cpu=args.cpu_tiles > 0)
    cpu = None
    if args.cpu_tiles > 0:
        print("[bench] CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        vp, vi = cams[0]
        cpu = cpu_baseline_ext(s, sh, deg, vp, vi, targets[0], W, H, args.cpu_tiles)
    out = {
        "metric": "fwd+bwd ms/view @1M splats 1080p",
        "value": round(bwd_ms, 3),
        "unit": "ms/view",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / args.steps, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, targets rendered from a seed-1 scene)",
        "config": {"workload": f"EXT 3D Gaussian fwd+bwd+Adam, {n} Gaussians, SH degree {deg}, {W}x{H}, "
                               f"{views} ring views, {tile}x{tile} tiles, depth-ordered uncapped lists, L1",
                   "views": views, "width": W, "height": H, "splats": n, "sh_degree": deg,
                   "tile": tile, "lds_batch": batch},
        "autotune_ms": None if tune_ms is None else {f"tile{t}_batch{b}": round(float(tune_ms[i, j]), 3)
                                                     for i, t in enumerate((8, 16, 32))
                                                     for j, b in enumerate((32, 64, 128))},
        "roofline": {"bound": "valu", "compute": "fp32 VALU issue (no matrix work on this path)", "kernel": "k_composite_ext", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic},
        "phases_ms": phases,
        "tile_keys_per_view": int(keys),
        "pairs_per_view": int(pairs),
        "active_pairs_per_view": int(active),
        "reverse_lane_efficiency": round(active / max(1.0, 64.0 * mean("reverse_wave_entries", 1)), 3),
        "binning_hbm_gbs": round(sort_bytes / (mean("sort") * 1e-3) / 1e9, 1),
        "adam_hbm_gbs": round(adam_gbs, 1),
        "adam_roofline_frac": round(adam_gbs / PEAK_HBM_GBS, 3),
        "steps_per_s": round(args.steps / wall, 2),
        "batched_multiview": batched,
        "bezier_strokes": bezier,
        "densification": adc,
        "learned_densification": learned,
        "ref": ref,
        "target_generators": ptleg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for sl in slots:
        mgr.free_target(sl)
    mgr.close()


if __name__ == "__main__":
    main()
