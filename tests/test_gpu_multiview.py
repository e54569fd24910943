"""Multi-view step (npt_*): the captured hipGraph (views in concurrent lanes,
then Adam) gives bit-identical splats, SH, Adam state and per-view losses to
view-after-view launches, over several steps with densification in the
cadence, and after the scene outgrows the captured key buffers."""
import numpy as np
import pytest

from nprush_amd import GpuMultiViewStep, GpuSplatOptimizerManager, camera, scenes

pytestmark = pytest.mark.gpu


def _scene(n, seed, cap):
    s, sh = scenes.make_ext_splats(n, seed, sh_degree=2)
    pool = np.zeros((cap, 16), np.float32)
    pool[:n] = s
    pool[n:, 3] = -10000.0
    pool[n:, 4:7] = -4.0
    pool[n:, 8] = 1.0
    shp = np.zeros((cap, 48), np.float32)
    shp[:n] = sh
    return pool, shp


def _views(seed):
    rng = np.random.default_rng(seed)
    sizes = [(256, 256), (200, 120), (512, 384), (96, 64), (320, 240)]
    out = []
    for k, (w, h) in enumerate(sizes):
        ang = rng.uniform(0, 2 * np.pi)
        eye = (3.0 * np.cos(ang), rng.uniform(-0.5, 0.5), 3.0 * np.sin(ang))
        vp, vi = camera.view_proj(eye, aspect=w / h)
        tgt = np.ones((h, w, 4), np.float32)
        tgt[..., :3] = rng.random((h, w, 3)).astype(np.float32)
        out.append((k % 2, vp, vi, tgt))
    return out


def _build(views, scenes_data):
    mgrs = []
    for pool, shp in scenes_data:
        m = GpuSplatOptimizerManager(numSplats=len(pool), mode="ext", sh_degree=2, splats=pool, sh=shp,
                                     adc_period=2)
        m.set_adc_ext(2e-4, 3.3, 0.005, 0.01)
        mgrs.append(m)
    step = GpuMultiViewStep(mgrs, lanes=3)
    for sc, vp, vi, tgt in views:
        step.add_view(sc, vp, vi, tgt)
    return mgrs, step


def test_graph_step_equals_per_view_step():
    data = [_scene(3000, 0, 3600), _scene(6000, 1, 7000)]
    views = _views(4)
    ma, sa = _build(views, data)
    mb, sb = _build(views, data)
    for k in range(4):
        if k == 3:   # every splat grows: more keys than the captured buffers hold
            for mgr in ma + mb:
                s = mgr.read_splats()
                s[:, 4:7] += 1.0
                mgr.write_splats(s)
        sa.step(graph=False)
        sb.step(graph=True)
        assert np.array_equal(sa.read_losses(), sb.read_losses()), k
        for x, y in zip(ma, mb):
            assert np.array_equal(x.read_splats(), y.read_splats()), k
            assert np.array_equal(x.read_sh(), y.read_sh()), k
            xm, xv, xt = x.read_adam()
            ym, yv, yt = y.read_adam()
            assert np.array_equal(xm, ym) and np.array_equal(xv, yv) and xt == yt, k
    assert sa.read_losses()[:, 1].min() > 0
    sa.close(); sb.close()
    for mgr in ma + mb:
        mgr.close()


def test_graph_overflow_keeps_densification_statistics():
    """A step whose keys overflow the captured buffers is redone view by view;
    the graph's chain-rule kernels must then add nothing, so the densification
    accumulators (summed |dL/d ndc| and view counts) equal the per-view run's
    after the overflow step and after one more graph step (no densification in
    between: its period is longer than the test)."""
    data = [_scene(3000, 0, 3600), _scene(6000, 1, 7000)]
    views = _views(4)

    def build():
        mgrs = []
        for pool, shp in data:
            m = GpuSplatOptimizerManager(numSplats=len(pool), mode="ext", sh_degree=2, splats=pool, sh=shp,
                                         adc_period=1000)
            m.set_adc_ext(2e-4, 3.3, 0.005, 0.01)
            mgrs.append(m)
        st = GpuMultiViewStep(mgrs, lanes=3)
        for sc, vp, vi, tgt in views:
            st.add_view(sc, vp, vi, tgt)
        return mgrs, st

    ma, sa = build()
    mb, sb = build()
    for k in range(3):
        if k == 1:   # every splat grows: more keys than the captured buffers hold
            for mgr in ma + mb:
                s = mgr.read_splats()
                s[:, 4:7] += 1.0
                mgr.write_splats(s)
        sa.step(graph=False)
        sb.step(graph=True)
        assert np.array_equal(sa.read_losses(), sb.read_losses()), k
        for x, y in zip(ma, mb):
            ax, ay = x.read_accum(), y.read_accum()
            assert ax[:, 1].max() > 0, k
            assert np.array_equal(ax, ay), (k, np.abs(ax - ay).max())
            assert np.array_equal(x.read_splats(), y.read_splats()), k
    sa.close(); sb.close()
    for mgr in ma + mb:
        mgr.close()


def test_autotuned_graph_step_matches_per_view():
    data = [_scene(2500, 2, 3000), _scene(4000, 3, 4500)]
    views = _views(7)
    ma, sa = _build(views, data)
    mb, sb = _build(views, data)
    tune = sb.autotune()
    assert tune.shape == (len(views), 2)
    assert set(tune[:, 0]) <= {8, 16, 32} and set(tune[:, 1]) <= {32, 64, 128}
    for k in range(2):
        sa.step(graph=False)
        sb.step(graph=True)
        la, lb = sa.read_losses(), sb.read_losses()
        # same images; the FP64 loss sums group pixels by tile, so only their rounding may differ
        assert np.allclose(la, lb, rtol=1e-12, atol=0), k
        for x, y in zip(ma, mb):
            gx, gy = x.read_splats(), y.read_splats()
            assert np.allclose(gx, gy, rtol=1e-5, atol=1e-6), k
    sa.close(); sb.close()
    for mgr in ma + mb:
        mgr.close()


def test_serial_per_view_step_equals_per_view_step():
    """Mode 2 (each view completed before the next is launched) gives the same
    losses, splats, SH and Adam state as mode 0 (views of different scenes
    overlapping), bit for bit, over 3 steps with densification every 2."""
    data = [_scene(3000, 4, 3600), _scene(5000, 5, 6000)]
    views = _views(5)
    ma, sa = _build(views, data)
    mb, sb = _build(views, data)
    for k in range(3):
        sa.step(graph=False)
        sb.step(graph=False, serial=True)
        assert np.array_equal(sa.read_losses(), sb.read_losses()), k
        for x, y in zip(ma, mb):
            assert np.array_equal(x.read_splats(), y.read_splats()), k
            assert np.array_equal(x.read_sh(), y.read_sh()), k
            xm, xv, xt = x.read_adam()
            ym, yv, yt = y.read_adam()
            assert np.array_equal(xm, ym) and np.array_equal(xv, yv) and xt == yt, k
    sa.close(); sb.close()
    for mgr in ma + mb:
        mgr.close()


def test_autotune_in_context_keeps_fastest_and_results():
    """npt_autotune times whole-step graphs for its candidates and keeps the
    fastest; the step it chooses gives the per-view results (the autotuner
    itself leaves the scenes untouched)."""
    data = [_scene(2500, 6, 3000), _scene(4000, 7, 4500)]
    views = _views(9)
    ma, sa = _build(views, data)
    mb, sb = _build(views, data)
    before = [m.read_splats() for m in mb]
    sb.autotune()
    for m, b in zip(mb, before):
        assert np.array_equal(m.read_splats(), b), "autotuning changed a scene"
    info = sb.graph_info()
    ms = [x for x in info["autotune_ms"] if x >= 0]
    assert len(ms) >= 2 and min(ms) > 0
    print("autotune candidates (ms):", info["autotune_ms"], "lanes", info["lanes"])
    sa.step(graph=False)
    sb.step(graph=True)
    assert sb.graph_info()["nodes"] > 0
    np.testing.assert_allclose(sa.read_losses(), sb.read_losses(), rtol=1e-12, atol=0)
    for x, y in zip(ma, mb):
        np.testing.assert_allclose(x.read_splats(), y.read_splats(), rtol=1e-5, atol=1e-6)
    sa.close(); sb.close()
    for mgr in ma + mb:
        mgr.close()


def _many_views(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        w, h = (64, 48) if k % 2 else (48, 64)
        ang = rng.uniform(0, 2 * np.pi)
        eye = (3.0 * np.cos(ang), rng.uniform(-0.5, 0.5), 3.0 * np.sin(ang))
        vp, vi = camera.view_proj(eye, aspect=w / h)
        tgt = np.ones((h, w, 4), np.float32)
        tgt[..., :3] = rng.random((h, w, 3)).astype(np.float32)
        out.append((k % 2, vp, vi, tgt))
    return out


def test_large_graph_near_the_node_limit_and_refusal_above_it():
    """A step graph close to the node limit is captured, replayed and equals
    per-view launches; one above it is refused (NPS_ERR_CAPACITY) and per-view
    mode still works."""
    from nprush_amd import _lib

    data = [_scene(1500, 8, 1800), _scene(1500, 9, 1800)]
    # nodes per view from two small graphs, then the view count that brings the step near the limit
    counts = []
    for nv in (4, 8):
        probe_m, probe = _build(_many_views(nv, 10), data)
        probe.step(graph=True)
        counts.append(probe.graph_info()["nodes"])
        probe.close()
        for m in probe_m:
            m.close()
    per_view = (counts[1] - counts[0]) / 4.0
    fixed = counts[0] - 4 * per_view
    near = int((0.985 * 8192 - fixed) / per_view)
    views = _many_views(near, 11)
    ma, sa = _build(views, data)
    mb, sb = _build(views, data)
    sa.step(graph=False)
    sb.step(graph=True)
    nodes = sb.graph_info()["nodes"]
    print("graph nodes:", nodes, "views:", near)
    assert 0.95 * 8192 <= nodes <= 8192
    sb.step(graph=True)   # replayed
    sa.step(graph=False)
    assert np.array_equal(sa.read_losses(), sb.read_losses())
    for x, y in zip(ma, mb):
        assert np.array_equal(x.read_splats(), y.read_splats())
    sa.close(); sb.close()
    for mgr in ma + mb:
        mgr.close()
    # above the limit
    views = _many_views(near + 40, 12)
    mc, sc = _build(views, data)
    with pytest.raises(_lib.NpsError, match="nodes"):
        sc.step(graph=True)
    sc.step(graph=False)
    assert sc.read_losses()[:, 1].min() > 0
    sc.close()
    for mgr in mc:
        mgr.close()
