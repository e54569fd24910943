"""Densification MLP on the GPU (nml_*): BF16 matrix-core forward and backward
against the CPU oracle (which rounds to bfloat16 at the same points), and
training on a small regression problem."""
import numpy as np
import pytest

from nprush_amd import DensifierMLP
from oracle import oracle_mlp as om

pytestmark = pytest.mark.gpu


def _norm():
    rng = np.random.default_rng(5)
    return (rng.normal(0, 1, 32).astype(np.float32), rng.uniform(0.5, 2, 32).astype(np.float32),
            rng.normal(0, 1, 14).astype(np.float32), rng.uniform(0.5, 2, 14).astype(np.float32))


@pytest.mark.parametrize("n", [1, 37, 512, 3000])
def test_forward_matches_oracle(n):
    rng = np.random.default_rng(n)
    p = om.init_params(3, out_scale=0.7)
    nrm = _norm()
    net = DensifierMLP(max_batch=1024, params=p)
    net.set_norm(*nrm)
    x = rng.normal(0, 1.5, (n, 32)).astype(np.float32)
    got = net.forward(x)
    ys, _ = om.forward(p, x, nrm, bf=True)
    want = ys * nrm[3] + nrm[2]
    err = np.abs(got - want)
    assert err.max() <= 2e-3 * np.abs(want).max() + 1e-5, err.max()
    net.close()


@pytest.mark.parametrize("n", [5, 64, 1000])
def test_gradients_match_oracle(n):
    rng = np.random.default_rng(10 + n)
    p = om.init_params(4, out_scale=0.7)
    nrm = _norm()
    net = DensifierMLP(max_batch=1024, params=p)
    net.set_norm(*nrm)
    x = rng.normal(0, 1.5, (n, 32)).astype(np.float32)
    t = rng.normal(0, 1.5, (n, 14)).astype(np.float32)
    loss, g = net.grads(x, t)
    wl, wg = om.grads(p, x, t, nrm, bf=True)
    assert abs(loss - wl) < 1e-3 * wl
    # per layer: within 1 % of the layer's gradient norm
    off = 0
    for i, o in om.DIMS:
        for size in (i * o, o):
            a, b = g[off:off + size], wg[off:off + size]
            assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-9, (off, size)
            off += size
    # and against the FP64 network (no bfloat16): the rounding costs a few percent at most
    _, g64 = om.grads(p, x, t, nrm, bf=False)
    assert np.linalg.norm(g - g64) < 0.05 * np.linalg.norm(g64)
    net.close()


def _oracle_adam_run(x, t, epochs, batch, lr):
    """The same schedule with the oracle's bfloat16 gradients and float64 Adam."""
    p = om.init_params(0, out_scale=0.0).astype(np.float64)
    m, v = np.zeros_like(p), np.zeros_like(p)
    k = 0
    for _ in range(epochs):
        for b in range(0, len(x), batch):
            _, g = om.grads(p.astype(np.float32), x[b:b + batch], t[b:b + batch], bf=True)
            k += 1
            m = 0.9 * m + 0.1 * g
            v = 0.999 * v + 0.001 * g * g
            p = p - lr * (m / (1 - 0.9 ** k)) / (np.sqrt(v / (1 - 0.999 ** k)) + 1e-8)
    return p


def test_training_follows_the_oracle_trajectory():
    """40 Adam steps on a smooth regression: the fit drops below a fifth of the
    target's energy and lands within 2 % of the oracle's own Adam run."""
    rng = np.random.default_rng(0)
    n, batch, lr, epochs = 4096, 1024, 1e-3, 10
    x = rng.normal(0, 1, (n, 32)).astype(np.float32)
    A = rng.normal(0, 0.3, (32, 14))
    t = (np.tanh(x @ A) + 0.1 * np.sin(3 * x[:, :14])).astype(np.float32)
    net = DensifierMLP(max_batch=batch, params=om.init_params(0, out_scale=0.0))
    for _ in range(epochs):
        for b in range(0, n, batch):
            net.train(x[b:b + batch], t[b:b + batch], lr)
    mse = float(np.mean((net.forward(x) - t) ** 2))
    p = _oracle_adam_run(x, t, epochs, batch, lr)
    want = float(np.mean((om.forward(p.astype(np.float32), x, bf=True)[0] - t) ** 2))
    assert mse < 0.2 * float(np.mean(t ** 2)), mse
    assert abs(mse - want) < 0.02 * want, (mse, want)
    net.close()
