"""Synthetic scenes of the benchmark configurations (seeded, numpy PCG64).

Splat layout (16 float32, splat_backward.wgsl:1-6):
``[x y z sx] [r g b opacity] [qw qx qy qz] [sy shape_a shape_b sz]``.

Distribution (all configurations): means uniform in [-1, 1]^3, per-axis scales
exp(N(-4, 0.5)), uniform unit quaternions (Shoemake's method), opacity
clamp(sigmoid(N(0, 1)), 0.01, 0.99) (the reference stores linear opacity
clamped to [0.01, 0.99], splat_step.wgsl:85-86), RGB uniform in [0, 1],
shape_a = 2, shape_b = 0.5 (the reference's initial shape,
GpuSplatOptimizerManager.ts:83-84, which makes exp(-b r^a) a Gaussian).
"""
from __future__ import annotations

import math

import numpy as np

from . import camera

SPLAT_FLOATS = 16


def make_splats(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = np.zeros((n, SPLAT_FLOATS), dtype=np.float32)
    means = rng.uniform(-1.0, 1.0, size=(n, 3))
    scales = np.exp(rng.normal(-4.0, 0.5, size=(n, 3)))
    u = rng.random((n, 3))
    q = np.stack([
        np.sqrt(1.0 - u[:, 0]) * np.sin(2.0 * math.pi * u[:, 1]),
        np.sqrt(1.0 - u[:, 0]) * np.cos(2.0 * math.pi * u[:, 1]),
        np.sqrt(u[:, 0]) * np.sin(2.0 * math.pi * u[:, 2]),
        np.sqrt(u[:, 0]) * np.cos(2.0 * math.pi * u[:, 2]),
    ], axis=1)
    logit = rng.normal(0.0, 1.0, size=n)
    opacity = np.clip(1.0 / (1.0 + np.exp(-logit)), 0.01, 0.99)
    rgb = rng.uniform(0.0, 1.0, size=(n, 3))
    s[:, 0:3] = means
    s[:, 3] = scales[:, 0]
    s[:, 4:7] = rgb
    s[:, 7] = opacity
    s[:, 8:12] = q
    s[:, 12] = scales[:, 1]
    s[:, 13] = 2.0
    s[:, 14] = 0.5
    s[:, 15] = scales[:, 2]
    return s


def with_dead_capacity(live: np.ndarray, capacity: int) -> np.ndarray:
    """Pads a live set with dead slots (opacity 0) up to `capacity` splats, as the
    reference's fixed-capacity pool holds them (GpuSplatOptimizerManager.ts:71-75)."""
    out = np.zeros((capacity, SPLAT_FLOATS), dtype=np.float32)
    out[: live.shape[0]] = live
    dead = out[live.shape[0]:]
    dead[:, 3] = dead[:, 12] = dead[:, 15] = 0.1
    dead[:, 8] = 1.0
    dead[:, 13] = 2.0
    dead[:, 14] = 0.5
    return out


def quantize8(img: np.ndarray) -> np.ndarray:
    """Values as an rgba8unorm texture returns them: k/255."""
    return (np.rint(np.clip(img, 0.0, 1.0) * 255.0) / 255.0).astype(np.float32)


def config0_camera():
    """One pinhole camera at (0,0,3) looking at the origin, fov 60 deg, square."""
    return camera.view_proj((0.0, 0.0, 3.0), aspect=1.0, fovy=math.pi / 3)


CONFIG0 = dict(n=16384, width=256, height=256, seed=0, target_seed=1, steps=100)
CONFIG1 = dict(n=1_000_000, capacity=2_000_000, width=1920, height=1080, views=8, radius=3.0, seed=0,
               target_seed=1)


CONFIG2_SIZES = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]
CONFIG2_COUNTS = [50_000, 150_000, 300_000, 500_000]


def config2_views(num_views: int = 64, seed: int = 4):
    """Views of the batched uneven workload: view k belongs to scene k % 4, its
    resolution drawn uniformly (seed 4) from {256^2, 512^2, 1024x768,
    1920x1080}, its eye on a radius-3 circle at a random angle and height in
    [-1, 1], looking at the origin.  Returns [(scene, W, H, vp, vp_inv)]."""
    rng = np.random.default_rng(seed)
    views = []
    for k in range(num_views):
        w, h = CONFIG2_SIZES[int(rng.integers(0, 4))]
        ang = rng.uniform(0.0, 2.0 * np.pi)
        eye = (3.0 * np.cos(ang), rng.uniform(-1.0, 1.0), 3.0 * np.sin(ang))
        views.append((k % 4, w, h) + camera.view_proj(eye, aspect=w / h))
    return views


SH_C0 = 0.28209479177387814


def make_ext_splats(n: int, seed: int, sh_degree: int = 3, higher_std: float = 0.05):
    """EXT-layout splats (config-1 distribution): [x y z opacity_logit]
    [log_sx log_sy log_sz 0] [qw qx qy qz] [0 0 0 0] plus SH coefficients
    (n, 48): DC from RGB uniform in [0, 1] (c0 = (rgb - 0.5) / C0), higher
    bands N(0, higher_std) up to sh_degree, zero beyond."""
    rng = np.random.default_rng(seed)
    s = np.zeros((n, SPLAT_FLOATS), dtype=np.float32)
    s[:, 0:3] = rng.uniform(-1.0, 1.0, size=(n, 3))
    s[:, 3] = rng.normal(0.0, 1.0, size=n)
    s[:, 4:7] = rng.normal(-4.0, 0.5, size=(n, 3))
    u = rng.random((n, 3))
    s[:, 8] = np.sqrt(1.0 - u[:, 0]) * np.sin(2.0 * math.pi * u[:, 1])
    s[:, 9] = np.sqrt(1.0 - u[:, 0]) * np.cos(2.0 * math.pi * u[:, 1])
    s[:, 10] = np.sqrt(u[:, 0]) * np.sin(2.0 * math.pi * u[:, 2])
    s[:, 11] = np.sqrt(u[:, 0]) * np.cos(2.0 * math.pi * u[:, 2])
    sh = np.zeros((n, 48), dtype=np.float32)
    rgb = rng.uniform(0.0, 1.0, size=(n, 3))
    sh[:, 0:3] = (rgb - 0.5) / SH_C0
    nk = (sh_degree + 1) ** 2
    if nk > 1:
        sh[:, 3:3 * nk] = rng.normal(0.0, higher_std, size=(n, 3 * nk - 3))
    return s, sh


BEZIER_FLOATS = 20


def make_bezier_curves(n: int, seed: int, radius_median: float = 0.004, radius_sigma: float = 0.3,
                       walk_std: float = 0.05, soft_ratio: float = 0.25) -> np.ndarray:
    """Cubic Bezier curves of the stroke benchmark (layout: p0.xyz width |
    p1.xyz softness | p2.xyz 0 | p3.xyz 0 | rgba, bezier_backward.wgsl:4-10).

    p0 uniform in [-1, 1]^3, p1..p3 a random walk with N(0, walk_std) steps,
    width log-normal(ln radius_median, radius_sigma), softness = soft_ratio *
    width (the reference's initial 0.02 / 0.005), opacity sigmoid(N(1, 0.5))
    capped at 0.99 (the step's clamp), RGB uniform in [0, 1]."""
    rng = np.random.default_rng(seed)
    c = np.zeros((n, BEZIER_FLOATS), dtype=np.float32)
    p = rng.uniform(-1.0, 1.0, size=(n, 3))
    for k in range(4):
        if k:
            p = p + rng.normal(0.0, walk_std, size=(n, 3))
        c[:, 4 * k:4 * k + 3] = p
    width = np.exp(rng.normal(math.log(radius_median), radius_sigma, size=n))
    c[:, 3] = width
    c[:, 7] = width * soft_ratio
    c[:, 16:19] = rng.uniform(0.0, 1.0, size=(n, 3))
    c[:, 19] = np.minimum(1.0 / (1.0 + np.exp(-rng.normal(1.0, 0.5, size=n))), 0.99)
    return c


def make_brush(size: int = 256, seed: int = 0) -> np.ndarray:
    """A synthetic bristle brush of the reference asset's shape (256 x 256 rgba8,
    red channel sampled): streaks along u with random strengths, a soft falloff
    towards both edges in v.  Rows are v, columns u."""
    rng = np.random.default_rng(seed)
    v = (np.arange(size) + 0.5) / size
    across = np.clip(1.0 - np.abs(2.0 * v - 1.0) ** 3, 0.0, 1.0)           # soft edges in v
    bristle = 0.6 + 0.4 * rng.random(size)                                   # per-row strength
    along = 0.85 + 0.15 * np.sin(np.linspace(0.0, 6.0 * np.pi, size))       # slow variation in u
    r = np.clip(across[:, None] * bristle[:, None] * along[None, :], 0.0, 1.0)
    b = np.zeros((size, size, 4), np.uint8)
    b[..., 0] = np.rint(r * 255.0).astype(np.uint8)
    b[..., 1] = b[..., 0]
    b[..., 2] = b[..., 0]
    b[..., 3] = 255
    return b
