"""Pins of the path-trace oracle (oracle/oracle_pt.c restating path_trace.wgsl,
bvh.ts, path_trace_resolve.wgsl and envmap.wgsl) by closed forms, brute force
and the mathematics:

* atan2/asin (the double-precision definition) against numpy in float64,
  rounded to float32;
* the PCG hash against an independent integer implementation;
* the BVH: every node's box holds its children's boxes and its triangles,
  leaves hold at most 4 triangles, every triangle appears once, nodes are
  numbered depth first, the root's split has the least SAH cost of all bin
  splits, and coincident centroids fall back to the middle of the list;
* the BVH walk finds the same closest hit as a brute-force loop over all
  triangles; the watertight triangle test agrees with a float64
  Moller-Trumbore test away from edges;
* frames with closed forms: a diffuse ground of albedo a under a constant
  environment E gives a*E per sample (the bounce escapes; exact up to the FP32
  rounding of the barycentric blend of the albedo), a mirror ground E,
  an empty scene E (E a power of two per channel, which FP32 bilinear
  filtering returns exactly); resolve = rgba8(Reinhard(4 mean));
* cosine-weighted hemisphere directions: unit length, in the hemisphere,
  mean cos = 2/3.
"""
import numpy as np

from nprush_amd import camera, meshes


def _f32(x):
    return np.float32(x)


def test_atan2_asin_vs_float64(oracle):
    rng = np.random.default_rng(0)
    y = rng.normal(size=20000).astype(np.float32)
    x = rng.normal(size=20000).astype(np.float32)
    y[:8] = [0.0, 1.0, -1.0, 0.0, 1e-30, -3.0, 2.0, -0.5]
    x[:8] = [1.0, 0.0, 0.0, -1.0, -1e-30, -3.0, 2.0, 0.5]
    a, _ = oracle.atan2_asin_f32(y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    assert np.mean(a == ref) > 0.9999
    assert np.all(np.abs(a.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64)) <= 1)
    z = np.concatenate([rng.uniform(-1, 1, 20000), [-1.0, 1.0, 0.0, 0.5, -0.5]]).astype(np.float32)
    _, s = oracle.atan2_asin_f32(z, np.ones_like(z))
    ref = np.arcsin(z.astype(np.float64)).astype(np.float32)
    assert np.mean(s == ref) > 0.9999
    assert np.all(np.abs(s.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64)) <= 1)


def _pcg_py(v):
    m = 0xFFFFFFFF
    s = (v * 747796405 + 2891336453) & m
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & m
    return ((w >> 22) ^ w) & m


def test_pcg_hash(oracle):
    for v in [0, 1, 2, 12345, 0xFFFFFFFF, 0x80000000, 2654435761]:
        assert oracle.pt_pcg(v) == _pcg_py(v)


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-2, 2, (n, 1, 3))
    p = c + rng.normal(scale=0.2, size=(n, 3, 3))
    v = np.zeros((3 * n, 12), np.float32)
    v[:, :3] = p.reshape(-1, 3)
    nrm = rng.normal(size=(3 * n, 3))
    v[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    v[:, 6:9] = rng.uniform(0.1, 0.9, (3 * n, 3))
    v[:, 9] = 1.0
    return v, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def _boxes(nodes):
    f = nodes.view(np.float32)
    return f[:, 0:3].astype(np.float64), f[:, 4:7].astype(np.float64)


def _tri_box(v, t):
    p = v[t][:, :, :3].astype(np.float64)
    return p.min(axis=1) - 1e-4, p.max(axis=1) + 1e-4


def test_bvh_structure(oracle):
    v, idx = _soup(700, 1)
    nodes, tris = oracle.pt_bvh(v, idx)
    lo, hi = _boxes(nodes)
    leaf = (nodes[:, 7] & 0x80000000) != 0
    seen = []
    for k in range(len(nodes)):
        if leaf[k]:
            first, cnt = int(nodes[k, 3]), int(nodes[k, 7] & 0x7FFFFFFF)
            assert 1 <= cnt <= 4
            t = tris[first:first + cnt]
            tlo, thi = _tri_box(v, t)
            # float32 rounding of the double box: within one float ulp
            assert np.all(tlo >= lo[k] - 1e-6 * (1 + np.abs(lo[k]))) and np.all(thi <= hi[k] + 1e-6 * (1 + np.abs(hi[k])))
            seen.append(t)
        else:
            a, b = int(nodes[k, 3]), int(nodes[k, 7])
            assert a == k + 1 and b > a                 # depth first: left child follows its parent
            for c in (a, b):
                assert np.all(lo[c] >= lo[k]) and np.all(hi[c] <= hi[k])
    got = np.sort(np.concatenate(seen), axis=0)
    assert sorted(map(tuple, got)) == sorted(map(tuple, idx))
    assert len(got) == len(idx)
    # leaves are visited in triangle-list order
    firsts = [int(nodes[k, 3]) for k in range(len(nodes)) if leaf[k]]
    assert firsts == sorted(firsts)


def test_bvh_root_split_is_sah_minimum(oracle):
    v, idx = _soup(300, 2)
    nodes, tris = oracle.pt_bvh(v, idx)
    tlo, thi = _tri_box(v, idx)
    cen = (tlo + thi) * 0.5
    root_lo, root_hi = tlo.min(0), thi.max(0)

    def sa(lo, hi):
        d = hi - lo
        return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])

    psa = sa(root_lo, root_hi)
    best = np.inf
    for axis in range(3):
        cmin, cmax = cen[:, axis].min(), cen[:, axis].max()
        b = np.minimum(np.floor((cen[:, axis] - cmin) / (cmax - cmin) * 8), 7)
        for s in range(7):
            L = b <= s
            if L.all() or (~L).all():
                continue
            cost = (sa(tlo[L].min(0), thi[L].max(0)) * L.sum() + sa(tlo[~L].min(0), thi[~L].max(0)) * (~L).sum()) / psa
            best = min(best, cost)
    # the root's two subtrees: their triangle sets give the chosen partition
    def subtree_tris(k):
        out = []
        stack = [k]
        while stack:
            q = stack.pop()
            if nodes[q, 7] & 0x80000000:
                f, c = int(nodes[q, 3]), int(nodes[q, 7] & 0x7FFFFFFF)
                out += [tuple(t) for t in tris[f:f + c]]
            else:
                stack += [int(nodes[q, 3]), int(nodes[q, 7])]
        return out

    left = set(subtree_tris(int(nodes[0, 3])))
    L = np.array([tuple(t) in left for t in idx])
    chosen = (sa(tlo[L].min(0), thi[L].max(0)) * L.sum() + sa(tlo[~L].min(0), thi[~L].max(0)) * (~L).sum()) / psa
    assert abs(chosen - best) <= 1e-12 * best


def test_bvh_coincident_centroids_split_in_the_middle(oracle):
    v, idx = _soup(1, 3)
    v = np.tile(v, (10, 1))
    idx = np.arange(30, dtype=np.uint32).reshape(10, 3)
    nodes, tris = oracle.pt_bvh(v, idx)
    assert int(nodes[0, 3]) == 1
    left = nodes[1]
    # 10 -> 5 | 5 -> (2 | 3) ... the first leaf holds the first triangles in order
    assert np.array_equal(tris[:, 0], idx[:, 0])
    assert (left[7] & 0x80000000) == 0 or int(left[7] & 0x7FFFFFFF) == 5


def test_bvh_walk_equals_brute_force(oracle):
    v, idx = _soup(400, 4)
    nodes, tris = oracle.pt_bvh(v, idx)
    rng = np.random.default_rng(5)
    for _ in range(600):
        ro = rng.uniform(-4, 4, 3).astype(np.float32)
        rd = rng.normal(size=3)
        rd = (rd / np.linalg.norm(rd)).astype(np.float32)
        t_b, k = oracle.pt_closest_brute(ro, rd, v, idx)
        h = oracle.pt_closest_bvh(ro, rd, v, nodes, tris)
        assert h[0] == t_b
        if k >= 0:
            # same triangle: its barycentric normal is reproduced by the brute loop's triangle
            assert np.isfinite(h[1:4]).all()


def test_triangle_test_vs_moller_trumbore(oracle):
    rng = np.random.default_rng(6)
    agree = 0
    total = 0
    for _ in range(3000):
        p = rng.uniform(-1, 1, (3, 3))
        v = np.zeros((3, 12), np.float32)
        v[:, :3] = p
        v[:, 5] = 1.0
        v[:, 9] = 1.0
        ro = rng.uniform(-3, 3, 3)
        tgt = p.mean(0) + rng.normal(scale=0.5, size=3)
        rd = (tgt - ro) / np.linalg.norm(tgt - ro)
        t32, k = oracle.pt_closest_brute(ro.astype(np.float32), rd.astype(np.float32), v,
                                         np.arange(3, dtype=np.uint32))
        # float64 Moller-Trumbore on the float32 inputs
        P = v[:, :3].astype(np.float64)
        o = ro.astype(np.float32).astype(np.float64)
        d = rd.astype(np.float32).astype(np.float64)
        e1, e2 = P[1] - P[0], P[2] - P[0]
        h = np.cross(d, e2)
        a = e1 @ h
        if abs(a) < 1e-12:
            continue
        s = o - P[0]
        u = (s @ h) / a
        q = np.cross(s, e1)
        w = (d @ q) / a
        t = (e2 @ q) / a
        margin = min(u, w, 1 - u - w)
        if abs(margin) < 1e-4 or abs(t - 1e-4) < 1e-6:
            continue
        total += 1
        hit64 = margin > 0 and t >= 1e-4
        if hit64 == (t32 >= 0):
            agree += 1
            if hit64:
                assert abs(t32 - t) <= 1e-5 * max(1.0, abs(t))
    assert total > 2000 and agree == total


def _looking_down(h=4.0, W=32, H=24):
    vp, vi = camera.view_proj((0.0, 0.0, h), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), aspect=W / H)
    return vi


def test_diffuse_ground_constant_env_closed_form(oracle):
    W, H = 32, 24
    E = np.array([0.5, 0.25, 1.0], np.float32)   # powers of two: the bilinear filter returns them exactly
    a = np.array([0.8, 0.6, 0.4], np.float32)
    gv, gi = meshes.ground(50.0, 0.0, color=tuple(a) + (1.0,))
    nodes, tris = oracle.pt_bvh(gv, gi)
    env = meshes.constant_environment(E)
    acc = oracle.pt_trace(_looking_down(W=W, H=H), 0, W, H, gv, nodes, tris, env)
    # the albedo is the barycentric blend of three equal vertex colours: exact up
    # to the FP32 rounding of u + v + w
    assert np.allclose(acc[..., :3], (a * E)[None, None, :], rtol=4e-7, atol=0)
    assert np.all(acc[..., 3] == 1.0)
    # a second frame adds the same
    acc2 = oracle.pt_trace(_looking_down(W=W, H=H), 1, W, H, gv, nodes, tris, env, accum=acc)
    assert np.allclose(acc2[..., :3], (2 * a * E)[None, None, :], rtol=6e-7, atol=0)
    assert np.all(acc2[..., 3] == 2.0)
    out = oracle.pt_resolve(acc2)
    c = acc2[..., :3] / np.float32(2.0) * np.float32(4.0)
    expect = np.rint(np.clip(c / (c + np.float32(1.0)), 0, 1) * np.float32(255.0)) / np.float32(255.0)
    assert np.all(out[..., :3] == expect.astype(np.float32))
    assert np.all(out[..., 3] == 1.0)


def test_mirror_and_empty_scene_constant_env(oracle):
    W, H = 24, 16
    E = np.array([0.25, 0.5, 2.0], np.float32)   # powers of two: the bilinear filter returns them exactly
    env = meshes.constant_environment(E)
    gv, gi = meshes.ground(50.0, 0.0, color=(0.3, 0.3, 0.3, 0.0))       # alpha 0: mirror
    nodes, tris = oracle.pt_bvh(gv, gi)
    acc = oracle.pt_trace(_looking_down(W=W, H=H), 3, W, H, gv, nodes, tris, env)
    assert np.all(acc[..., :3] == E[None, None, :])
    nodes0, tris0 = oracle.pt_bvh(np.zeros((0, 12), np.float32), np.zeros((0, 3), np.uint32))
    assert nodes0.shape[0] == 1 and (nodes0[0, 7] & 0x80000000) and (nodes0[0, 7] & 0x7FFFFFFF) == 0
    acc0 = oracle.pt_trace(_looking_down(W=W, H=H), 0, W, H, np.zeros((0, 12), np.float32), nodes0, tris0, env)
    assert np.all(acc0[..., :3] == E[None, None, :])


def test_environment_background_and_sampler(oracle):
    """A texture that varies only in v (latitude) and is constant in u: the
    background of a camera looking straight up sees the top rows; looking at
    the horizon, v = 1/2 exactly between rows H/2 - 1 and H/2."""
    eh, ew = 8, 16
    env = np.ones((eh, ew, 4), np.float32)
    env[..., :3] = (np.arange(eh, dtype=np.float32) / eh)[:, None, None]
    W, H = 4, 3
    # horizon camera: the centre pixel's direction is horizontal (z = 0)
    vp, vi = camera.view_proj((0.0, 0.0, 0.0), target=(1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), aspect=W / H)
    out = oracle.env_background(vi, env, 5, 5)
    c = np.float32(((eh // 2 - 1) + (eh // 2)) / 2.0 / eh) * np.float32(4.0)
    mid = np.rint(np.clip(c / (c + np.float32(1.0)), 0, 1) * np.float32(255.0)) / np.float32(255.0)
    assert abs(out[2, 2, 0] - mid) <= 1.0 / 255.0
    # rows above the horizon sample smaller v (higher latitude) -> darker
    assert out[0, 2, 0] < out[2, 2, 0] < out[4, 2, 0]


def test_cosine_hemisphere_distribution(oracle):
    for n in ([0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [0.6, -0.48, 0.64]):
        n = np.asarray(n, np.float32)
        d = oracle.pt_hemisphere(n, 7, 40000).astype(np.float64)
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
        c = d @ n.astype(np.float64)
        assert np.all(c >= -1e-6)
        assert abs(c.mean() - 2.0 / 3.0) < 0.01                        # E[cos] of a cosine lobe
        t = d - c[:, None] * n.astype(np.float64)[None, :]
        assert np.linalg.norm(t.mean(0)) < 0.01


# ------------------------------------------------------------ mesh raster
def _rgba8(x):
    x = np.clip(np.asarray(x, np.float32), 0, 1)
    return (np.rint(x * np.float32(255.0)) / np.float32(255.0)).astype(np.float32)


def _reinhard(c):
    c = np.asarray(c, np.float32)
    return c / (c + np.float32(1.0))


def _down_camera(h, W, H):
    vp, vi = camera.view_proj((0.0, 0.0, h), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), aspect=W / H)
    view = camera.look_at((0.0, 0.0, h), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    return vp, view, vi


def test_raster_ground_closed_forms(oracle):
    """A ground quad facing a camera that looks straight down from height h:
    normal colour (0, 0, 1) -> (1/2, 1/2, 1); matcap mode with a constant matcap
    M and vertex colour c -> M c; depth 1 - 0.1 / h (every vertex has clip w = h);
    the reversed quad is culled (background, depth 1)."""
    W, H, h = 24, 16, 2.5
    vp, view, vi = _down_camera(h, W, H)
    c = np.array([0.6, 0.3, 0.9], np.float32)          # 4 M c away from the 8-bit rounding boundaries
    gv, gi = meshes.ground(40.0, 0.0, color=tuple(c) + (1.0,))
    nodes, tris = oracle.pt_bvh(gv, gi)
    M = np.array([0.5, 1.0, 0.25], np.float32)
    env = meshes.constant_environment(M)
    col, dep = oracle.mesh_raster(vp, view, vi, 0.0, W, H, gv, nodes, tris, env)
    assert np.all(col[..., :3] == _rgba8(_reinhard(np.float32([0.5, 0.5, 1.0]) * np.float32(4.0))))
    d = _rgba8(np.float32(1.0) - np.float32(0.1) / np.float32(h))
    assert np.all(dep[..., :3] == d) and np.all(dep[..., 3] == 1.0) and np.all(col[..., 3] == 1.0)
    col, dep = oracle.mesh_raster(vp, view, vi, 1.0, W, H, gv, nodes, tris, env)
    assert np.all(col[..., :3] == _rgba8(_reinhard(M * c * np.float32(4.0))))
    # reversed winding: back-facing -> culled
    nodes_r, tris_r = oracle.pt_bvh(gv, gi[:, ::-1].copy())
    col, dep = oracle.mesh_raster(vp, view, vi, 1.0, W, H, gv, nodes_r, tris_r, env)
    assert np.all(dep[..., :3] == 1.0)
    assert np.array_equal(col, oracle.env_background(vi, env, W, H))


def test_raster_nearer_surface_wins(oracle):
    W, H, h = 20, 20, 3.0
    vp, view, vi = _down_camera(h, W, H)
    lo = meshes.ground(40.0, 0.0, color=(1.0, 0.0, 0.0, 1.0))
    hi = meshes.ground(0.5, 1.0, color=(0.0, 1.0, 0.0, 1.0))     # a small square 1 unit above the ground
    v = np.concatenate([lo[0], hi[0]])
    i = np.concatenate([lo[1], hi[1] + len(lo[0])])
    nodes, tris = oracle.pt_bvh(v, i)
    env = meshes.constant_environment((1.0, 1.0, 1.0))
    col, dep = oracle.mesh_raster(vp, view, vi, 1.0, W, H, v, nodes, tris, env)
    centre = col[H // 2, W // 2, :3]
    corner = col[0, 0, :3]
    assert centre[1] > 0 and centre[0] == 0          # the raised square (green) covers the centre
    assert corner[0] > 0 and corner[1] == 0          # the ground (red) elsewhere
    assert dep[H // 2, W // 2, 0] == _rgba8(np.float32(1.0) - np.float32(0.1) / np.float32(h - 1.0))
    assert dep[0, 0, 0] == _rgba8(np.float32(1.0) - np.float32(0.1) / np.float32(h))
