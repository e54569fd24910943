"""The batched uneven workload (4 scenes of the scale-up distribution, views of
256^2, 512^2, 1024x768 and 1920x1080 on random ring positions) stepped as one
captured multi-view graph, against the EXT oracle: per view binning + forward +
L1 + backward + chain rule, gradients summed per scene, then the oracle's Adam.

Compared after one step from zero Adam state:
  * per-view L1 and L2 losses: the images are bit-exact, so only the FP64
    summation order differs (relative 1e-9);
  * Adam's first moment m = (1 - 0.9) g, i.e. the GPU's summed gradients: per
    parameter group |m_gpu - m_ora| <= 1e-4 max|m_ora| + 1e-5 |m_ora| (the GPU
    sums per-pair FP32 terms in FP32 wave sums and FP64 across waves, the oracle
    in FP64);
  * parameters: Adam's first step is lr * sign(g) up to rounding, so every
    parameter whose gradient is above that tolerance moves exactly as the
    oracle's (a few FP32 ULP); those below may differ by at most 2 lr.
"""
import numpy as np
import pytest

from nprush_amd import GpuMultiViewStep, GpuSplatOptimizerManager, scenes

pytestmark = pytest.mark.gpu

GROUPS = {"means": ([0, 1, 2], 1.6e-4), "opacity": ([3], 5e-2), "scales": ([4, 5, 6], 5e-3),
          "quat": ([8, 9, 10, 11], 1e-3), "sh_dc": (list(range(16, 19)), 2.5e-3),
          "sh_rest": (list(range(19, 64)), 1.25e-4)}


def _targets(views, counts):
    out = {}
    for sidx, n in enumerate(counts):
        ts, tsh = scenes.make_ext_splats(n, 100 + sidx, sh_degree=3)
        tm = GpuSplatOptimizerManager(numSplats=n, mode="ext", sh_degree=3, splats=ts, sh=tsh, adc_enabled=False)
        for k, (sc, w, h, vp, vi) in enumerate(views):
            if sc != sidx:
                continue
            tm.setBackwardTarget(np.zeros((h, w, 4), np.float32), np.ones((h, w), np.float32), w, h)
            tm.writeSplatVPMatrix(vp, vi)
            tm.forward()
            c, _ = tm.read_image()
            rgba = np.ones((h, w, 4), np.float32)
            rgba[..., :3] = scenes.quantize8(c)
            out[k] = rgba
        tm.close()
    return out


def _oracle_step(oracle, views, tgts, data):
    res = []
    losses = np.zeros((len(views), 2))
    for sidx, (s, sh) in enumerate(data):
        g = np.zeros((len(s), 64))
        for k, (sc, w, h, vp, vi) in enumerate(views):
            if sc != sidx:
                continue
            offs, ids = oracle.ext_bin_all(s, sh, 3, vp, vi, w, h)
            gg, gs, _, loss = oracle.ext_backward_csr_f32(s, sh, 3, vp, vi, tgts[k], w, h, offs, ids)
            g[:, :16] += gg
            g[:, 16:] += gs
            losses[k] = (loss[1], loss[0])
        z = np.zeros((len(s), 64), np.float32)
        res.append(oracle.ext_adam_f32(s, sh, g.astype(np.float32), z, z, 3, 0.0))
    return res, losses


def _check(mgrs, step, ora, ora_losses):
    got = step.read_losses()
    np.testing.assert_allclose(got, ora_losses, rtol=1e-9, atol=0)
    for sidx, (mgr, (s1, sh1, m1, v1, t1)) in enumerate(zip(mgrs, ora)):
        m, v, t = mgr.read_adam()
        assert t == t1 == 1.0
        m = m.reshape(len(s1), 64)
        p_gpu = np.concatenate([mgr.read_splats(), mgr.read_sh()], axis=1).astype(np.float64)
        p_ora = np.concatenate([s1, sh1], axis=1).astype(np.float64)
        for name, (cols, lr) in GROUPS.items():
            ref, mg = m1[:, cols].astype(np.float64), m[:, cols].astype(np.float64)
            scale = np.abs(ref).max()
            tol = 1e-4 * scale + 1e-5 * np.abs(ref)
            err = np.abs(mg - ref)
            assert np.all(err <= tol), (sidx, name, err.max(), scale)
            firm = np.abs(ref) > 2 * tol
            d = np.abs(p_gpu[:, cols] - p_ora[:, cols])
            ulp = np.spacing(np.abs(p_ora[:, cols]).astype(np.float32)).astype(np.float64)
            assert np.all(d[firm] <= 4 * ulp[firm] + 1e-3 * lr), (sidx, name, d[firm].max())
            assert np.all(d <= 2 * lr * (1 + 1e-3) + 4 * ulp), (sidx, name, d.max())
        assert np.array_equal(p_gpu[:, [7, 12, 13, 14, 15]], p_ora[:, [7, 12, 13, 14, 15]])


def _run(oracle, views, counts, autotune):
    data = [scenes.make_ext_splats(n, sidx, sh_degree=3) for sidx, n in enumerate(counts)]
    tgts = _targets(views, counts)
    ora, ora_losses = _oracle_step(oracle, views, tgts, data)
    mgrs = [GpuSplatOptimizerManager(numSplats=len(s), mode="ext", sh_degree=3, splats=s, sh=sh, adc_enabled=False)
            for s, sh in data]
    step = GpuMultiViewStep(mgrs, lanes=8)
    for k, (sc, w, h, vp, vi) in enumerate(views):
        step.add_view(sc, vp, vi, tgts[k])
    if autotune:
        tune = step.autotune()
        print("autotuned (tile, batch) per view:", sorted({(int(a), int(b)) for a, b in tune}))
    step.step(graph=True)
    _check(mgrs, step, ora, ora_losses)
    step.close()
    for m in mgrs:
        m.close()


@pytest.mark.parametrize("autotune", [False, True])
def test_config2_views_reduced_scenes_graph_vs_oracle(oracle, autotune):
    """All 64 views of the workload at their resolutions; scenes at a tenth of
    their Gaussians (5k/15k/30k/50k) so the oracle finishes in seconds."""
    views = scenes.config2_views()
    counts = [n // 10 for n in scenes.CONFIG2_COUNTS]
    _run(oracle, views, counts, autotune)


def test_config2_full_scenes_view_subset_graph_vs_oracle(oracle):
    """The full 50k/150k/300k/500k scenes with two of each scene's views (8
    views: every resolution class occurs)."""
    views = scenes.config2_views()
    keep, seen = [], {}
    for v in views:
        if seen.get(v[0], 0) < 2:
            keep.append(v)
            seen[v[0]] = seen.get(v[0], 0) + 1
    assert len(keep) == 8
    _run(oracle, keep, scenes.CONFIG2_COUNTS, False)
