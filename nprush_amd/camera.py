"""Camera matrices in the reference's conventions.

The reference builds ``viewProjMat = perspective(fov, aspect, 0.01, 100) * view``
and its inverse with wgpu-matrix 3.4.2 (src/lib/viewer/Camera.svelte.ts:18-33,
deno.lock:14,782), storing every matrix as a column-major Float32Array:
WebGPU clip space (z in [0, 1]), ``w_clip = -z_view``.  wgpu-matrix is not
vendored in the reference, so this module restates the closed forms:

* perspective: f = tan(pi/2 - fovy/2); m[0] = f/aspect, m[5] = f,
  m[10] = far/(near-far), m[11] = -1, m[14] = far*near/(near-far);
* look_at (view matrix): z = normalize(eye - target), x = normalize(cross(up, z)),
  y = normalize(cross(z, x)); rows x, y, z with translation -dot(axis, eye);
* mul: double-precision sums of float32 entries, rounded to float32;
* inverse: computed in float64 from the float32 entries, rounded to float32.
"""
from __future__ import annotations

import math

import numpy as np


def _f32(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float64).astype(np.float32)


def perspective(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = math.tan(math.pi * 0.5 - 0.5 * fovy)
    m = np.zeros(16, dtype=np.float64)
    m[0] = f / aspect
    m[5] = f
    range_inv = 1.0 / (near - far)
    m[10] = far * range_inv
    m[11] = -1.0
    m[14] = far * near * range_inv
    return _f32(m)


def _normalize(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64)
    n = math.sqrt(float(v @ v))
    return _f32(v / n if n > 0.00001 else np.zeros(3)).astype(np.float64)


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye = _f32(eye).astype(np.float64)
    target = _f32(target).astype(np.float64)
    up = _f32(up).astype(np.float64)
    z = _normalize(eye - target)
    x = _normalize(np.cross(up, z))
    y = _normalize(np.cross(z, x))
    m = np.zeros(16, dtype=np.float64)
    m[0], m[1], m[2], m[3] = x[0], y[0], z[0], 0.0
    m[4], m[5], m[6], m[7] = x[1], y[1], z[1], 0.0
    m[8], m[9], m[10], m[11] = x[2], y[2], z[2], 0.0
    m[12] = -(x[0] * eye[0] + x[1] * eye[1] + x[2] * eye[2])
    m[13] = -(y[0] * eye[0] + y[1] * eye[1] + y[2] * eye[2])
    m[14] = -(z[0] * eye[0] + z[1] * eye[1] + z[2] * eye[2])
    m[15] = 1.0
    return _f32(m)


def mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a @ b for column-major float32[16] matrices (sums in float64)."""
    A = np.asarray(a, dtype=np.float64).reshape(4, 4).T
    B = np.asarray(b, dtype=np.float64).reshape(4, 4).T
    out = np.zeros((4, 4), dtype=np.float64)
    for r in range(4):
        for c in range(4):
            out[r, c] = ((A[r, 0] * B[0, c] + A[r, 1] * B[1, c]) + A[r, 2] * B[2, c]) + A[r, 3] * B[3, c]
    return _f32(out.T.reshape(16))


def inverse(m: np.ndarray) -> np.ndarray:
    M = np.asarray(m, dtype=np.float64).reshape(4, 4).T
    return _f32(np.linalg.inv(M).T.reshape(16))


def view_proj(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovy=math.pi / 3, aspect=1.0,
              near=0.01, far=100.0):
    """Returns (vp, vp_inv) as column-major float32[16]."""
    vp = mul(perspective(fovy, aspect, near, far), look_at(eye, target, up))
    return vp, inverse(vp)


def ring_cameras(count: int, radius: float, width: int, height: int, fovy=math.pi / 3):
    """Cameras on a horizontal ring, eye = radius*(cos 2pi k/count, 0, sin 2pi k/count),
    looking at the origin with +y up."""
    cams = []
    for k in range(count):
        ang = 2.0 * math.pi * k / count
        eye = (radius * math.cos(ang), 0.0, radius * math.sin(ang))
        cams.append(view_proj(eye, aspect=width / height, fovy=fovy))
    return cams
