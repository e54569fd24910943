"""Synthetic mesh scenes and environment maps for the path-traced target.

The reference loads its scene from .glb assets and an .hdr panorama
(src/lib/gpu/loadGlb.ts, loadHdrTexture.ts); those assets are not shipped with
it, so scenes here are generated: an icosphere (diffuse, coloured), a mirror
box (vertex alpha 0, path_trace.wgsl:47 material rule) and a ground quad, in
the 12-float vertex layout of path_trace.wgsl:4 [p xyz, n xyz, rgba, uv], and
an equirectangular environment (Z up, envmap.wgsl) with a sky gradient and a
sun.  Everything is seeded and deterministic.
"""
from __future__ import annotations

import numpy as np


def _vert(p, n, c, uv=(0.0, 0.0)):
    return [p[0], p[1], p[2], n[0], n[1], n[2], c[0], c[1], c[2], c[3], uv[0], uv[1]]


def icosphere(subdiv: int = 2, radius: float = 1.0, centre=(0.0, 0.0, 0.0), color=(0.8, 0.3, 0.2, 1.0)):
    """Unit icosahedron subdivided `subdiv` times, scaled and moved; smooth normals,
    counter-clockwise faces seen from outside.  20 * 4^subdiv triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    cen = np.asarray(centre, np.float64)
    verts = np.array([_vert(cen + radius * p, p, color) for p in v], np.float32)
    return verts, np.array(f, np.uint32)


def box(lo=(-0.5, -0.5, 0.0), hi=(0.5, 0.5, 1.0), color=(1.0, 1.0, 1.0, 0.0)):
    """Axis-aligned box, 12 triangles, flat normals, outward ccw faces.  Alpha 0
    (the default) makes it a mirror."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    verts, idx = [], []
    for axis in range(3):
        for side in (0, 1):
            n = np.zeros(3)
            n[axis] = 1.0 if side else -1.0
            u, w = (axis + 1) % 3, (axis + 2) % 3
            quad = []
            for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
                p = np.empty(3)
                p[axis] = hi[axis] if side else lo[axis]
                p[u] = hi[u] if a else lo[u]
                p[w] = hi[w] if b else lo[w]
                quad.append(p)
            base = len(verts)
            for p in quad:
                verts.append(_vert(p, n, color))
            tri = [(0, 1, 2), (0, 2, 3)]
            if not side:
                tri = [(0, 2, 1), (0, 3, 2)]
            for t in tri:
                idx.append((base + t[0], base + t[1], base + t[2]))
    return np.array(verts, np.float32), np.array(idx, np.uint32)


def ground(size: float = 10.0, z: float = 0.0, color=(0.6, 0.6, 0.55, 1.0), grid: int = 1):
    """Square ground at height z facing +z, split into grid x grid quads."""
    verts, idx = [], []
    xs = np.linspace(-size, size, grid + 1)
    for j in range(grid + 1):
        for i in range(grid + 1):
            verts.append(_vert((xs[i], xs[j], z), (0.0, 0.0, 1.0), color, (i / grid, j / grid)))
    for j in range(grid):
        for i in range(grid):
            a = j * (grid + 1) + i
            idx += [(a, a + 1, a + grid + 2), (a, a + grid + 2, a + grid + 1)]
    return np.array(verts, np.float32), np.array(idx, np.uint32)


def scene(subdiv: int = 3, seed: int = 0):
    """The default target scene: a coloured icosphere, a mirror box and a ground,
    as a list of (vertices, indices) meshes in setMeshes order."""
    rng = np.random.default_rng(seed)
    col = tuple(rng.uniform(0.2, 0.9, 3)) + (1.0,)
    return [icosphere(subdiv, 0.8, (0.0, 0.0, 0.8), col),
            box((0.9, -0.4, 0.0), (1.5, 0.4, 1.2)),
            ground(6.0, 0.0, grid=4)]


def environment(width: int = 256, height: int = 128, sun=(0.5, 0.3, 0.8), sun_power: float = 20.0):
    """Equirectangular HDR environment, Z up (envmap.wgsl mapping: u = atan2(y, x) /
    2 pi + 1/2, v = 1/2 - asin(z) / pi): a blue-to-white sky over a brown floor and
    a small sun disc around the unit direction `sun`."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    lon = (u - 0.5) * 2.0 * np.pi
    lat = (0.5 - v) * np.pi
    lo, la = np.meshgrid(lon, lat)
    d = np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], -1)
    up = np.clip(d[..., 2], 0.0, 1.0)[..., None]
    sky = (1.0 - up) * np.array([0.9, 0.95, 1.0]) + up * np.array([0.25, 0.45, 0.9])
    floor = np.array([0.3, 0.22, 0.15])
    col = np.where(d[..., 2:3] >= 0.0, sky, floor)
    s = np.asarray(sun, np.float64)
    s = s / np.linalg.norm(s)
    col = col + sun_power * (np.einsum("hwk,k->hw", d, s) > 0.995)[..., None]
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = col
    return out


def constant_environment(value=(1.0, 1.0, 1.0), width: int = 8, height: int = 4):
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = np.asarray(value, np.float32)
    return out
