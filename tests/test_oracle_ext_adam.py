"""Pins of the EXT Adam oracle (CPU), independent of the GPU: the first step's
closed form per parameter group, the value of epsilon, an FP64 restatement of
the m/v recursion with bias correction at t = 7, and the parameters Adam must
leave alone (record pads, SH coefficients above the degree).

The EXT model has no reference function; its learning rates are the ones the
configurations name (means 1.6e-4, opacity 5e-2, scales 5e-3, rotation 1e-3,
SH DC 2.5e-3, higher SH bands 2.5e-3/20) with beta 0.9/0.999 and eps 1e-15."""
import numpy as np
import pytest

# (record / SH column indices, learning rate); SH columns are 16 + 3k + c
GROUPS = {
    "means": (list(range(0, 3)), 1.6e-4),
    "opacity": ([3], 5e-2),
    "log_scales": (list(range(4, 7)), 5e-3),
    "rotation": (list(range(8, 12)), 1e-3),
    "sh_dc": (list(range(16, 19)), 2.5e-3),
    "sh_higher": (list(range(19, 64)), 1.25e-4),
}
PADS = [7, 12, 13, 14, 15]


def _state(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.normal(0.0, 1.0, (n, 16)).astype(np.float32)
    sh = rng.normal(0.0, 1.0, (n, 48)).astype(np.float32)
    return s, sh


def _cols(s, sh):
    return np.concatenate([s, sh], axis=1).astype(np.float64)


@pytest.mark.parametrize("deg", [0, 1, 3])
def test_first_step_is_minus_lr_sign_g_per_group(oracle, deg):
    n = 32
    s, sh = _state(n, 1)
    rng = np.random.default_rng(2)
    g = (rng.choice([-1.0, 1.0], (n, 64)) * rng.uniform(0.5, 4.0, (n, 64))).astype(np.float32)
    z = np.zeros((n, 64), np.float32)
    s1, sh1, m1, v1, t1 = oracle.ext_adam_f32(s, sh, g, z, z, deg, 0.0)
    assert t1 == 1.0
    d = _cols(s1, sh1) - _cols(s, sh)
    nk = (deg + 1) ** 2
    for name, (cols, lr) in GROUPS.items():
        live = [c for c in cols if c < 16 or (c - 16) // 3 < nk]
        dead = [c for c in cols if c not in live]
        if live:
            want = -lr * np.sign(g[:, live].astype(np.float64))
            # FP32 rounding of p - upd only: |p| <= ~5, so 1 ULP of p is <= 4.8e-7
            np.testing.assert_allclose(d[:, live], want, rtol=0, atol=1e-6 * max(1.0, lr * 1e3), err_msg=name)
            # the update itself (before rounding into p) is lr*sign(g) to a few FP32 ULP of lr
            upd = np.float32(lr) * np.sign(g[:, live])
            got = (_cols(s, sh) - _cols(s1, sh1))[:, live]
            assert np.all(np.abs(got - upd) <= 4 * np.spacing(np.abs(_cols(s, sh)[:, live]).astype(np.float32)) +
                          1e-3 * lr), name
        if dead:
            assert np.array_equal(_cols(s1, sh1)[:, dead], _cols(s, sh)[:, dead]), f"{name} above degree {deg} moved"
            assert not m1[:, dead].any() and not v1[:, dead].any()
    assert np.array_equal(s1[:, PADS], s[:, PADS]), "record pads moved"
    assert not m1[:, PADS].any() and not v1[:, PADS].any()


def test_group_rates_are_distinct_and_exact(oracle):
    """A wrong rate in any one group (e.g. higher SH bands at lr instead of lr/20)
    changes that group's first step by a factor, so check each rate to 1e-5."""
    n = 4
    s = np.zeros((n, 16), np.float32)
    sh = np.zeros((n, 48), np.float32)
    g = np.ones((n, 64), np.float32)
    z = np.zeros((n, 64), np.float32)
    s1, sh1, _, _, _ = oracle.ext_adam_f32(s, sh, g, z, z, 3, 0.0)
    d = -_cols(s1, sh1)
    for name, (cols, lr) in GROUPS.items():
        np.testing.assert_allclose(d[:, cols], lr, rtol=1e-5, err_msg=name)


def test_epsilon_is_1e_minus_15(oracle):
    """With |g| = 1e-14 the first step is lr*g/(|g| + eps) = lr/1.1 for eps = 1e-15;
    eps = 1e-8 (the reference's REF value) would give ~lr*1e-6."""
    n = 2
    s = np.zeros((n, 16), np.float32)
    sh = np.zeros((n, 48), np.float32)
    g = np.full((n, 64), 1e-14, np.float32)
    z = np.zeros((n, 64), np.float32)
    s1, _, _, _, _ = oracle.ext_adam_f32(s, sh, g, z, z, 0, 0.0)
    np.testing.assert_allclose(-s1[:, 0], 1.6e-4 * 1e-14 / (1e-14 + 1e-15), rtol=1e-4)
    np.testing.assert_allclose(-s1[:, 3], 5e-2 * 1e-14 / (1e-14 + 1e-15), rtol=1e-4)


def test_seven_steps_follow_fp64_recursion(oracle):
    """Seven steps from zero state with varying gradients: m, v, t and every
    parameter against an FP64 restatement of Adam with bias correction
    (m_t = b1 m + (1-b1) g, v_t = b2 v + (1-b2) g^2,
     p -= lr (m_t/(1-b1^t)) / (sqrt(v_t/(1-b2^t)) + eps))."""
    n, deg = 24, 2
    s, sh = _state(n, 7)
    rng = np.random.default_rng(8)
    m = np.zeros((n, 64), np.float32)
    v = np.zeros((n, 64), np.float32)
    t = 0.0
    P = _cols(s, sh)
    M = np.zeros((n, 64))
    V = np.zeros((n, 64))
    lr = np.zeros(64)
    for cols, r in GROUPS.values():
        lr[cols] = r
    nk = (deg + 1) ** 2
    lr[16 + 3 * nk:] = 0.0
    for step in range(1, 8):
        g = rng.normal(0.0, 1.0, (n, 64)).astype(np.float32) * (1.0 + step)
        s, sh, m, v, t = oracle.ext_adam_f32(s, sh, g, m, v, deg, t)
        G = g.astype(np.float64)
        on = lr > 0
        M[:, on] = 0.9 * M[:, on] + 0.1 * G[:, on]
        V[:, on] = 0.999 * V[:, on] + 0.001 * G[:, on] ** 2
        upd = lr * (M / (1 - 0.9 ** step)) / (np.sqrt(V / (1 - 0.999 ** step)) + 1e-15)
        P = P - upd
        assert t == float(step)
    np.testing.assert_allclose(m[:, lr > 0], M[:, lr > 0], rtol=2e-5, atol=1e-6)
    np.testing.assert_allclose(v[:, lr > 0], V[:, lr > 0], rtol=2e-5, atol=1e-6)
    got = _cols(s, sh)
    # accumulated FP32 rounding of 7 updates of size <= lr into |p| <= ~5
    np.testing.assert_allclose(got, P, rtol=0, atol=4e-6)
    assert np.array_equal(got[:, lr == 0], P[:, lr == 0])
    # the parameters did move by many ULP, so the tolerance above is meaningful (a few
    # random walks of +-lr steps return to their start)
    moved = np.abs(got[:, lr > 0] - _cols(*_state(n, 7))[:, lr > 0])
    assert np.median(moved) > 1e-4 and (moved > 1e-5).mean() > 0.9


def test_bias_correction_at_t7_closed_form(oracle):
    """A constant gradient g for 7 steps: m_hat = g and v_hat = g^2 exactly in
    real arithmetic, so every step moves by lr*sign(g) and after 7 steps the
    displacement is 7*lr*sign(g)."""
    n = 8
    s = np.zeros((n, 16), np.float32)
    sh = np.zeros((n, 48), np.float32)
    g = np.where(np.arange(64) % 2 == 0, 3.0, -0.25).astype(np.float32)[None, :].repeat(n, 0)
    m = np.zeros((n, 64), np.float32)
    v = np.zeros((n, 64), np.float32)
    t = 0.0
    for _ in range(7):
        s, sh, m, v, t = oracle.ext_adam_f32(s, sh, g, m, v, 3, t)
    d = _cols(s, sh)
    for name, (cols, lr) in GROUPS.items():
        np.testing.assert_allclose(d[:, cols], -7 * lr * np.sign(g[:, cols]), rtol=2e-5, err_msg=name)
    assert not d[:, PADS].any()
