"""Learned densification on the GPU: parent features, placement order, child
deltas, device snapshots and the held-out-view judge (nps_adc_features,
nps_adc_children, nps_adc_apply_deltas(_dev), nps_snapshot, densify.LearnedDensifier)."""
import numpy as np
import pytest

from nprush_amd import DensifierMLP, GpuSplatOptimizerManager, camera, densify, scenes

pytestmark = pytest.mark.gpu

W, H = 96, 64


def _scene(n_live=3000, cap=4000, seed=0):
    s, sh = scenes.make_ext_splats(cap, seed, sh_degree=3)
    s[n_live:, 3] = -10000.0
    ts, tsh = scenes.make_ext_splats(n_live, seed + 1, sh_degree=3)
    cams = camera.ring_cameras(3, 3.0, W, H)
    tm = GpuSplatOptimizerManager(numSplats=n_live, mode="ext", sh_degree=3, splats=ts, sh=tsh, adc_enabled=False)
    tm.setBackwardTarget(np.zeros((H, W, 4), np.float32), np.ones((H, W), np.float32), W, H)
    targets = []
    for vp, vi in cams:
        tm.writeSplatVPMatrix(vp, vi)
        tm.forward()
        c, _ = tm.read_image()
        t = np.ones((H, W, 4), np.float32)
        t[..., :3] = c
        targets.append(t)
    tm.close()
    m = GpuSplatOptimizerManager(numSplats=cap, mode="ext", sh_degree=3, splats=s, sh=sh, adc_enabled=False)
    m.set_adc_ext(1e-5, 3.3, 0.005, 0.01)
    depth = np.ones((H, W), np.float32)
    slots = [m.upload_target(t, depth) for t in targets]
    views = [(cams[i][0], cams[i][1], slots[i]) for i in range(3)]
    return m, views


def _host_features(s, sh, acc, m, v, idx, split_scale, cam):
    f = np.zeros((len(idx), 32))
    p = s[idx].astype(np.float64)
    f[:, 0:3] = p[:, 0:3]
    q = p[:, 8:12] / np.linalg.norm(p[:, 8:12], axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                  np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                  np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    sc = np.exp(p[:, 4:7])
    cov = np.einsum("nra,nca,na->nrc", R, R, sc * sc)
    f[:, 3:9] = np.stack([cov[:, 0, 0], cov[:, 0, 1], cov[:, 0, 2], cov[:, 1, 1], cov[:, 1, 2], cov[:, 2, 2]], -1)
    f[:, 9] = 1 / (1 + np.exp(-p[:, 3]))
    f[:, 10:13] = sh[idx, 0:3] * 0.28209479177387814 + 0.5
    a = acc[idx].astype(np.float64)
    f[:, 13] = np.where(a[:, 1] > 0, a[:, 0] / np.maximum(a[:, 1], 1e-30), 0)
    f[:, 14] = a[:, 1]
    f[:, 15:18] = p[:, 4:7]
    smax = sc.max(1)
    f[:, 18] = smax
    f[:, 19] = smax / split_scale
    f[:, 20] = smax > split_scale
    f[:, 21] = p[:, 3]
    f[:, 22] = a[:, 0]
    d = np.linalg.norm(p[:, 0:3] - cam, axis=1)
    f[:, 23] = d
    f[:, 24] = 1 / np.maximum(d, 1e-6)
    cols = densify.CHILD_COLS
    mm, vv = m[idx][:, cols].astype(np.float64), v[idx][:, cols].astype(np.float64)
    f[:, 25:32] = mm / (np.sqrt(vv) + 1e-12)
    return f


def test_features_placements_deltas_snapshot():
    m, views = _scene()
    densify.train_steps(m, views, 0, 12)
    s, sh, acc = m.read_splats(), m.read_sh(), m.read_accum()
    am, av, _ = m.read_adam()
    feats = m.adc_features()
    with np.errstate(over="ignore"):  # parked slots hold logit -1e4
        op = 1 / (1 + np.exp(-s[:, 3].astype(np.float64)))
    qual = np.where((op >= 0.005) & (acc[:, 1] > 0) & (acc[:, 0] / np.maximum(acc[:, 1], 1e-30) >= 1e-5))[0]
    dead = np.where(op < 0.005)[0]
    assert len(feats) == len(qual) > 0
    vi = np.asarray(views[-1][1], np.float64)
    cam = np.array([vi[8] / vi[11], vi[9] / vi[11], vi[10] / vi[11]])
    want = _host_features(s, sh, acc, am, av, qual, 0.01 * 3.3, cam)
    err = np.abs(feats - want) / (np.abs(want) * 2e-5 + 1e-6)
    bad = np.argwhere(err > 1.0)
    rows = sorted(set(int(r) for r, _ in bad))
    assert len(bad) == 0, (rows[:10], len(rows), len(qual), int(qual[0]), feats[0, :4].tolist(), want[0, :4].tolist(),
                           feats[1, :4].tolist(), want[1, :4].tolist(), s[qual[0], :4].tolist())
    m.adc()
    par, ch = m.adc_children()
    P = min(len(qual), len(dead))
    assert np.array_equal(par, qual[:P]) and np.array_equal(ch, dead[:P])
    heur = m.read_splats()
    m.snapshot()
    rng = np.random.default_rng(0)
    d = rng.normal(0, 0.01, (P, 14)).astype(np.float32)
    m.adc_apply_deltas(d)
    got = m.read_splats()
    exp = heur.copy()
    cols = densify.CHILD_COLS
    for c, slots in enumerate((par, ch)):
        for j, col in enumerate(cols):
            exp[slots, col] = heur[slots, col] + d[:, 7 * c + j]
    assert np.array_equal(got, exp)
    m.restore()
    assert np.array_equal(m.read_splats(), heur)
    m.close()


def test_snapshot_undoes_training_steps():
    """A restore after training steps gives back the whole optimizer state,
    Adam's step counter included: snapshot, 5 steps, restore, 5 steps equals a
    twin's 5 steps bit for bit (splats, SH, Adam m/v and t)."""
    m, views = _scene()
    twin, tviews = _scene()
    m.snapshot()
    densify.train_steps(m, views, 0, 5)
    m.restore()
    densify.train_steps(m, views, 0, 5)
    densify.train_steps(twin, tviews, 0, 5)
    assert np.array_equal(m.read_splats(), twin.read_splats())
    assert np.array_equal(m.read_sh(), twin.read_sh())
    am, av, at = m.read_adam()
    bm, bv, bt = twin.read_adam()
    assert np.array_equal(am, bm) and np.array_equal(av, bv) and at == bt == 5.0
    m.close()
    twin.close()


def test_lookahead_bound_with_no_move_is_the_plain_run():
    """run_lookahead_bound trains ahead from a snapshot at every densification;
    moving the children 0 of the way there must leave the plain trajectory
    untouched (same losses, bit for bit), so the bound's gains come from the
    children's placement only."""
    m, views = _scene(seed=2)
    plain = densify.run(m, views, 90, 30, None, eval_every=30)
    m.close()
    m, views = _scene(seed=2)
    bound0 = densify.run_lookahead_bound(m, views, 90, 30, horizon=15, eval_every=30, fraction=0.0)
    m.close()
    assert [l for _, l in plain] == [l for _, l in bound0]


@pytest.mark.parametrize("bad", [True, False])
def test_judge_keeps_heuristic_on_worse_proposal(bad):
    m, views = _scene()
    twin, tviews = _scene()
    densify.train_steps(m, views, 0, 12)
    densify.train_steps(twin, tviews, 0, 12)
    net = DensifierMLP(max_batch=4096, params=densify.init_params(0))
    if bad:  # a proposal that pushes every child far away
        net.set_norm(np.zeros(32), np.ones(32), np.full(14, 5.0), np.ones(14))
    judge = densify.LearnedDensifier(net, views[0])
    judge.densify(m)
    twin.adc()
    decision = judge.decisions[-1][0]
    if bad:
        assert decision == "heuristic"
        assert np.array_equal(m.read_splats(), twin.read_splats())
    else:  # zero deltas: same loss, accepted, identical state
        assert decision == "learned" and judge.decisions[-1][2] == judge.decisions[-1][3]
        assert np.array_equal(m.read_splats(), twin.read_splats())
    judge.close()
    for x in (m, twin, net):
        x.close()


def test_recorded_pairs_train_and_densify():
    """The whole loop at small scale: pairs from plain runs, the MLP trained on
    them, then plain vs learned densification on a fresh scene.

    Asserted: training lowers the training loss; the shrink factors lie in
    [0, 1]; every accepted proposal lowered (or kept) the held-out view's loss;
    if the judge accepted none, the learned run is the plain run bit for bit
    (same final loss); otherwise its final loss is within 2 % of plain's."""
    m, views = _scene(seed=0)
    X, T = densify.record_pairs(m, views, events=2, warm=20, window=40)
    m.close()
    assert X.shape[0] > 100 and T.shape == (X.shape[0], 14)
    net, losses, (before, after) = densify.train_mlp(X, T, epochs=10, batch=1024)
    assert losses[-1] < losses[0]
    assert np.all((net.shrink >= 0) & (net.shrink <= 1))
    finals = {}
    accepted = 0
    for name in ("plain", "learned"):
        m, views = _scene(seed=4)
        judge = densify.LearnedDensifier(net, views[0]) if name == "learned" else None
        curve = densify.run(m, views, 120, 40, judge, eval_every=40)
        finals[name] = curve[-1][1]
        if judge is not None:
            assert all(lm <= lh for dec, _, lh, lm in judge.decisions if dec == "learned")
            accepted = sum(1 for d in judge.decisions if d[0] == "learned")
            judge.close()
        m.close()
    print("held-out R2 per output", np.round(net.heldout_r2, 3).tolist(), "shrink", np.round(net.shrink, 3).tolist())
    print("final L1 plain %.4f learned %.4f, proposals accepted %d" % (finals["plain"], finals["learned"], accepted))
    if accepted == 0:
        assert finals["learned"] == finals["plain"]
    else:
        assert finals["learned"] <= 1.02 * finals["plain"]
    net.close()
