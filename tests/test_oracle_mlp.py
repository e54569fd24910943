"""Pins of the densification MLP's CPU oracle (oracle/oracle_mlp.py)."""
import numpy as np

from oracle import oracle_mlp as om


def test_param_count_and_layout():
    assert om.NUM_PARAMS == 143630
    p = np.arange(om.NUM_PARAMS, dtype=np.float64)
    layers = om.split(p)
    assert [W.shape for W, _ in layers] == [(256, 32), (256, 256), (256, 256), (14, 256)]
    assert layers[0][1][0] == 256 * 32 and layers[3][1][-1] == om.NUM_PARAMS - 1


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 3.0, -3.0])
    # tanh form of GELU: 0.5 z (1 + tanh(sqrt(2/pi)(z + 0.044715 z^3)))
    want = np.array([0.0, 0.8411920, -0.1588080, 2.9963627, -0.0036373])
    assert np.abs(om.gelu(z) - want).max() < 2e-6
    h = 1e-6
    fd = (om.gelu(z + h) - om.gelu(z - h)) / (2 * h)
    assert np.abs(om.gelu_d(z) - fd).max() < 1e-8


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, np.inf], np.float32)
    got = om.bf16(x)
    # 1 + 2^-8 is a tie: rounds to even (1.0); 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7 and got[3] == -2.5
    assert np.isinf(got[5])


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = om.init_params(1, out_scale=0.5).astype(np.float64)
    x = rng.normal(0, 1, (9, 32))
    t = rng.normal(0, 1, (9, 14))
    loss, g = om.grads(p, x, t, bf=False)
    idx = np.concatenate([rng.choice(om.NUM_PARAMS, 60, replace=False), [om.NUM_PARAMS - 1, 0, 256 * 32]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (om.grads(pp, x, t, bf=False)[0] - om.grads(pm, x, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), i


def test_bf16_path_close_to_fp64():
    rng = np.random.default_rng(1)
    p = om.init_params(2, out_scale=0.5)
    x = rng.normal(0, 1, (64, 32))
    t = rng.normal(0, 1, (64, 14))
    l64, g64 = om.grads(p, x, t, bf=False)
    l16, g16 = om.grads(p, x, t, bf=True)
    assert abs(l16 - l64) < 0.02 * l64
    assert np.linalg.norm(g16 - g64) < 0.03 * np.linalg.norm(g64)
