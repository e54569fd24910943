"""Learned densification for the EXT model (no reference counterpart).

The heuristic (3D Gaussian splatting's clone/split, ``nps_adc``) places the
children of every qualifying parent.  The MLP (``DensifierMLP``: 32 parent
features -> 14 deltas) proposes corrections to those children: position,
log-scale and opacity logit of the child kept in the parent's slot and of the
child placed in a dead slot.  It is trained on pairs this module records from
plain runs: the parent's features at a densification, and how far each child
moved in the following ``window`` steps without further densification.

At densification time the heuristic children are computed first; the proposal
is applied on top of them and kept only if the loss of a held-out view is not
worse than with the heuristic children alone (the heuristic is the fallback
and the judge).
"""
from __future__ import annotations

import time

import numpy as np

from .managers import DensifierMLP, GpuSplatOptimizerManager

# child parameters the deltas address, in EXT splat columns: position, log-scale, opacity logit
CHILD_COLS = [0, 1, 2, 4, 5, 6, 3]


def init_params(seed: int = 0, out_scale: float = 0.0) -> np.ndarray:
    """He-normal weights, zero biases; the output layer scaled by out_scale
    (0: the untrained network proposes no change)."""
    rng = np.random.default_rng(seed)
    parts = []
    dims = [(32, 256), (256, 256), (256, 256), (256, 14)]
    for k, (i, o) in enumerate(dims):
        W = rng.normal(0.0, np.sqrt(2.0 / i), (o, i))
        if k == len(dims) - 1:
            W *= out_scale
        parts += [W.reshape(-1), np.zeros(o)]
    return np.concatenate(parts).astype(np.float32)


def _set_view(mgr: GpuSplatOptimizerManager, view):
    vp, vi, slot = view
    mgr.writeSplatVPMatrix(vp, vi)
    mgr.set_target_slot(slot)


def train_steps(mgr: GpuSplatOptimizerManager, views, k0: int, n: int):
    """n optimizer steps (backward + Adam) cycling over views, no densification."""
    for k in range(n):
        _set_view(mgr, views[(k0 + k) % len(views)])
        mgr.backward()
        mgr.step()


def _children(s: np.ndarray, par: np.ndarray, ch: np.ndarray) -> np.ndarray:
    return np.concatenate([s[par][:, CHILD_COLS], s[ch][:, CHILD_COLS]], 1)


def record_pairs(mgr: GpuSplatOptimizerManager, views, events: int, warm: int = 100, window: int = 200):
    """Plain densification, `events` times: `warm` steps, features of the parents,
    densification, `window` steps without densification, then the children's
    movement.  Returns (X [P, 32], T [P, 14])."""
    X, T = [], []
    k = 0
    for _ in range(events):
        train_steps(mgr, views, k, warm)
        k += warm
        feats = mgr.adc_features()
        mgr.adc()
        par, ch = mgr.adc_children()
        if len(par) == 0:
            continue
        before = _children(mgr.read_splats(), par, ch)
        train_steps(mgr, views, k, window)
        k += window
        after = _children(mgr.read_splats(), par, ch)
        X.append(feats[:len(par)])
        T.append(after - before)
    if not X:
        return np.zeros((0, 32), np.float32), np.zeros((0, 14), np.float32)
    return np.concatenate(X).astype(np.float32), np.concatenate(T).astype(np.float32)


def train_mlp(X: np.ndarray, T: np.ndarray, epochs: int = 30, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
              holdout: float = 0.1, shrink: bool = True):
    """Trains a fresh DensifierMLP on standardised (X, T); returns (net, train
    losses per epoch, held-out loss of the untrained and trained network).

    shrink: each of the 14 outputs is then scaled by its least-squares factor on
    the held-out pairs, beta = <prediction, target> / <prediction, prediction>
    clipped to [0, 1] (the factor that minimises the held-out squared error of
    beta * prediction, shrinking weak predictions toward "no change"); the
    network's de-standardisation carries it.  ``net.shrink`` and
    ``net.heldout_r2`` keep the factors and the held-out R^2 per output."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    perm = rng.permutation(n)
    nv = max(1, int(n * holdout)) if n > 10 else 0
    va, tr = perm[:nv], perm[nv:]
    in_mean, in_std = X[tr].mean(0), X[tr].std(0) + 1e-6
    out_mean, out_std = T[tr].mean(0), T[tr].std(0) + 1e-9
    net = DensifierMLP(max_batch=batch, params=init_params(seed))
    net.set_norm(in_mean, in_std, out_mean, out_std)

    def held():
        if nv == 0:
            return float("nan")
        y = net.forward(X[va])
        return float(np.mean(((y - T[va]) / out_std) ** 2))
    before = held()
    losses = []
    for _ in range(epochs):
        rng.shuffle(tr)
        tot, cnt = 0.0, 0
        for b in range(0, len(tr), batch):
            idx = tr[b:b + batch]
            if len(idx) == 0:
                continue
            tot += net.train(X[idx], T[idx], lr) * len(idx)
            cnt += len(idx)
        losses.append(tot / max(cnt, 1))
    after = held()
    net.shrink = np.ones(14, np.float32)
    net.heldout_r2 = np.full(14, np.nan, np.float32)
    if nv > 1:
        y, t = net.forward(X[va]).astype(np.float64), T[va].astype(np.float64)
        tc = t - t.mean(0)
        net.heldout_r2 = (1.0 - ((y - t) ** 2).sum(0) / np.maximum((tc ** 2).sum(0), 1e-300)).astype(np.float32)
        if shrink:
            beta = np.clip((y * t).sum(0) / np.maximum((y * y).sum(0), 1e-300), 0.0, 1.0)
            net.shrink = beta.astype(np.float32)
            net.set_norm(in_mean, in_std, beta * out_mean, beta * out_std + 1e-12)
    return net, losses, (before, after)


def view_loss(mgr: GpuSplatOptimizerManager, view) -> float:
    """L1 of one view (image only)."""
    _set_view(mgr, view)
    mgr.forward()
    return mgr.read_loss()[1]


class LearnedDensifier:
    """Densification with MLP-proposed children; the heuristic children are kept
    whenever the proposal's held-out-view L1 is worse."""

    def __init__(self, net: DensifierMLP, heldout_view):
        self.net = net
        self.heldout = heldout_view
        self.decisions = []
        self._delta = (None, 0, None)  # (device buffer, rows, owner)

    def _deltas_dev(self, mgr, rows):
        ptr, cap, owner = self._delta
        if ptr is None or cap < rows or owner is not mgr:
            if ptr is not None:
                owner.dev_free(ptr)
            cap = max(rows, 1024)
            ptr = mgr.dev_alloc(56 * cap)
            self._delta = (ptr, cap, mgr)
        return ptr

    def close(self):
        ptr, _, owner = self._delta
        if ptr is not None:
            owner.dev_free(ptr)
        self._delta = (None, 0, None)

    def densify(self, mgr: GpuSplatOptimizerManager):
        # features, proposal and application stay in device memory
        feat_ptr, nq = mgr.adc_features_dev()
        rows = 0
        if nq > 0:
            dptr = self._deltas_dev(mgr, nq)
            self.net.forward_dev(feat_ptr, nq, dptr)
        mgr.adc()
        par, _ = mgr.adc_children()
        if len(par) == 0:
            self.decisions.append(("none", 0, 0.0, 0.0))
            return
        rows = len(par)
        lh = view_loss(mgr, self.heldout)
        mgr.snapshot()
        mgr.adc_apply_deltas_dev(self._delta[0], rows)
        lm = view_loss(mgr, self.heldout)
        if lm > lh:
            mgr.restore()
            self.decisions.append(("heuristic", len(par), lh, lm))
        else:
            self.decisions.append(("learned", len(par), lh, lm))


def run(mgr: GpuSplatOptimizerManager, views, steps: int, adc_period: int, densifier: LearnedDensifier | None = None,
        eval_every: int = 50):
    """Training with densification every adc_period steps (plain, or learned
    when a densifier is given).  Returns [(training seconds, mean L1 over the
    views)] every eval_every steps; evaluation time is not counted."""
    curve = []
    t_train = 0.0
    for k in range(steps):
        t0 = time.perf_counter()
        _set_view(mgr, views[k % len(views)])
        mgr.backward()
        mgr.step()
        if (k + 1) % adc_period == 0:
            if densifier is None:
                mgr.adc()
            else:
                densifier.densify(mgr)
        mgr.sync()
        t_train += time.perf_counter() - t0
        if (k + 1) % eval_every == 0:
            curve.append((t_train, float(np.mean([view_loss(mgr, v) for v in views]))))
    return curve


def run_lookahead_bound(mgr: GpuSplatOptimizerManager, views, steps: int, adc_period: int, horizon: int,
                        eval_every: int = 100, fraction: float = 1.0):
    """Upper bound of what any child-placement proposal can gain: training with
    plain densification where, after each densification, the children are
    moved `fraction` of the way to the state they reach `horizon` steps later
    (a look-ahead run from a snapshot, then restored).  fraction 1 is a perfect
    prediction of their future; a fraction f removes 1 - (1 - f)^2 of the
    squared placement error (a predictor of that R^2).  Only the main
    trajectory is timed.  Returns the run's [(seconds, mean L1)]."""
    curve = []
    t_train = 0.0
    for k in range(steps):
        t0 = time.perf_counter()
        _set_view(mgr, views[k % len(views)])
        mgr.backward()
        mgr.step()
        if (k + 1) % adc_period == 0:
            mgr.adc()
            mgr.sync()
            t_train += time.perf_counter() - t0
            par, ch = mgr.adc_children()
            if len(par):
                mgr.snapshot()
                train_steps(mgr, views, k + 1, horizon)
                ahead = mgr.read_splats()
                mgr.restore()
                s = mgr.read_splats()
                for slots in (par, ch):
                    now, fut = s[slots[:, None], CHILD_COLS], ahead[slots[:, None], CHILD_COLS]
                    s[slots[:, None], CHILD_COLS] = fut if fraction == 1.0 else now + np.float32(fraction) * (fut - now)
                mgr.write_splats(s)
            t0 = time.perf_counter()
        mgr.sync()
        t_train += time.perf_counter() - t0
        if (k + 1) % eval_every == 0:
            curve.append((t_train, float(np.mean([view_loss(mgr, v) for v in views]))))
    return curve


def time_to_target(curve, target: float):
    """Training seconds of the first evaluation at or below target (None if never)."""
    for t, l in curve:
        if l <= target:
            return t
    return None
