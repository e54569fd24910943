// k_pt.hip -- progressive path tracer of the mesh scene (the training target
// generator of the reference's frame, path_trace.wgsl + path_trace_resolve.wgsl)
// and the equirectangular environment background (envmap.wgsl).
//
//   k_pt_trace   : one camera path per pixel per frame, 8 bounces: closest hit
//                  by a depth-first walk of the SAH BVH (the reference's order:
//                  right child pushed first, left child popped first, at most
//                  64 entries on the stack), watertight ray/triangle test, mirror
//                  (vertex alpha < 0.5) or diffuse bounce (cosine-weighted
//                  hemisphere from a PCG hash of pixel and frame), the
//                  environment sampled where a path escapes; the path's radiance
//                  and a sample count added to the pixel's accumulator;
//   k_pt_resolve : accumulator / count, x4, Reinhard, stored as rgba8 (k/255).
//
// Every float expression is written in the order the oracle evaluates it
// (oracle/oracle_pt.c, -ffp-contract=off on both sides): dot(a, b) =
// (ax bx + ay by) + az bz, M v = ((c0 x + c1 y) + c2 z) + c3 w, normalize(v) =
// v / sqrt(dot(v, v)), atan2/asin/sin/cos from nps_math.h.  A wave traces an 8x8
// pixel block (neighbouring primary rays walk the same nodes); the BVH walk's
// stack lives in LDS (scratch memory was 20-25 % slower).
#include "nps_common.h"
#include "nps_kernels.h"
#include "nps_pt.h"

namespace nps {

namespace {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize3(V3 v) {
    const float l = sqrtf(dot3(v, v));
    return v3(v.x / l, v.y / l, v.z / l);
}
__device__ __forceinline__ float comp(V3 v, uint32_t k) { return k == 0u ? v.x : (k == 1u ? v.y : v.z); }

// column-major 4x4 times (x, y, z, 1) -> homogeneous xyz / w
__device__ __forceinline__ V3 unproject(const float* m, float x, float y, float z) {
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = ((m[i] * x + m[4 + i] * y) + m[8 + i] * z) + m[12 + i] * 1.0f;
    return v3(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}

// The PCG hash of path_trace.wgsl:36-40 (u32 wrap-around arithmetic).
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float2 rand2(uint32_t& seed) {
    seed = pcg(seed);
    const uint32_t a = seed;
    seed = pcg(seed);
    const uint32_t b = seed;
    return make_float2((float)a / 4294967296.0f, (float)b / 4294967296.0f);
}

// Closest hit of a walk: distance, barycentric weights and the triangle's
// vertices (its attributes are blended once the walk is over; the blend of the
// closest hit is the shader's, which blends at every accepted hit).
struct PtHit {
    bool hit;
    float t, u, v, w;
    uint32_t i0, i1, i2;
};

__device__ __forceinline__ void load_vert(const float4* __restrict__ vt, uint32_t i, V3& p, V3& n, float4& c) {
    const float4 a = vt[3u * i], b = vt[3u * i + 1u], d = vt[3u * i + 2u];
    p = v3(a.x, a.y, a.z);
    n = v3(a.w, b.x, b.y);
    c = make_float4(b.z, b.w, d.x, d.y);
}
__device__ __forceinline__ V3 load_pos(const float4* __restrict__ vt, uint32_t i) {
    const float4 a = vt[3u * i];
    return v3(a.x, a.y, a.z);
}

// Watertight ray/triangle test (path_trace.wgsl:60-129): shear to the ray's
// dominant axis, scaled barycentrics, hit if all of U, V, W share a sign,
// 1e-4 <= t < t_max.  kCull: a triangle whose ccw normal (p1 - p0) x (p2 - p0)
// does not face the ray (dot >= 0) is skipped (the raster's back-face cull).
template <bool kCull>
__device__ __forceinline__ void intersect_tri(V3 ro, V3 rd, const float4* __restrict__ vt, uint32_t i0, uint32_t i1,
                                              uint32_t i2, float t_max, PtHit& best) {
    const V3 p0 = load_pos(vt, i0), p1 = load_pos(vt, i1), p2 = load_pos(vt, i2);
    if (kCull && !(dot3(cross3(sub(p1, p0), sub(p2, p0)), rd) < 0.0f)) return;
    const float ax = fabsf(rd.x), ay = fabsf(rd.y), az = fabsf(rd.z);
    uint32_t kz = 0u;
    if (ay > ax) kz = 1u;
    if (az > fmax_w(ax, ay)) kz = 2u;
    uint32_t kx = (kz + 1u) % 3u, ky = (kx + 1u) % 3u;
    if (comp(rd, kz) < 0.0f) {
        const uint32_t t = kx;
        kx = ky;
        ky = t;
    }
    const float Sz = 1.0f / comp(rd, kz);
    const float Sx = comp(rd, kx) * Sz, Sy = comp(rd, ky) * Sz;
    const V3 A = sub(p0, ro), B = sub(p1, ro), C = sub(p2, ro);
    const float Ax = comp(A, kx) - Sx * comp(A, kz), Ay = comp(A, ky) - Sy * comp(A, kz);
    const float Bx = comp(B, kx) - Sx * comp(B, kz), By = comp(B, ky) - Sy * comp(B, kz);
    const float Cx = comp(C, kx) - Sx * comp(C, kz), Cy = comp(C, ky) - Sy * comp(C, kz);
    const float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return;
    const float det = (U + V) + W;
    if (det == 0.0f) return;
    const float Az = Sz * comp(A, kz), Bz = Sz * comp(B, kz), Cz = Sz * comp(C, kz);
    const float T = (U * Az + V * Bz) + W * Cz;
    const float inv_det = 1.0f / det;
    const float t = T * inv_det;
    if (t < 1e-4f || t >= t_max) return;
    best.hit = true;
    best.t = t;
    best.u = U * inv_det;
    best.v = V * inv_det;
    best.w = W * inv_det;
    best.i0 = i0;
    best.i1 = i1;
    best.i2 = i2;
}

// Slab test (path_trace.wgsl:132-142) with WGSL min/max (a NaN operand loses).
__device__ __forceinline__ bool aabb_hit(V3 ro, V3 inv, const float4& mn, const float4& mx, float t_max) {
    const float t0x = (mn.x - ro.x) * inv.x, t0y = (mn.y - ro.y) * inv.y, t0z = (mn.z - ro.z) * inv.z;
    const float t1x = (mx.x - ro.x) * inv.x, t1y = (mx.y - ro.y) * inv.y, t1z = (mx.z - ro.z) * inv.z;
    const float tmin = fmax_w(fmax_w(fmin_w(t0x, t1x), fmin_w(t0y, t1y)), fmin_w(t0z, t1z));
    const float tmax = fmin_w(fmin_w(fmax_w(t0x, t1x), fmax_w(t0y, t1y)), fmax_w(t0z, t1z));
    return tmax >= 0.0f && tmax >= tmin && tmin < t_max;
}

// Closest hit (path_trace.wgsl:146-185).  The walk's stack (64 entries) is the
// lane's column of an LDS array: entry k at stack[64 k] (lanes side by side, so
// a wave's pushes and pops hit distinct banks).
template <bool kCull>
__device__ PtHit scene_hit(V3 ro, V3 rd, const float4* __restrict__ nodes, const uint32_t* __restrict__ tris,
                           const float4* __restrict__ vt, uint32_t nnodes, uint32_t* stack) {
    PtHit best;
    best.hit = false;
    best.t = 1e30f;
    best.u = best.v = best.w = 0.0f;
    best.i0 = best.i1 = best.i2 = 0u;
    if (nnodes == 0u) return best;
    const V3 inv = v3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    uint32_t top = 0u;
    stack[64u * top++] = 0u;
    while (top > 0u) {
        const uint32_t node = stack[64u * --top];
        const float4 mn = nodes[2u * node], mx = nodes[2u * node + 1u];
        if (!aabb_hit(ro, inv, mn, mx, best.t)) continue;
        const uint32_t d1 = __float_as_uint(mx.w), d0 = __float_as_uint(mn.w);
        if (d1 & kPtLeafFlag) {
            const uint32_t count = d1 & ~kPtLeafFlag;
            for (uint32_t k = 0u; k < count; ++k) {
                const uint32_t b = 3u * (d0 + k);
                intersect_tri<kCull>(ro, rd, vt, tris[b], tris[b + 1u], tris[b + 2u], best.t, best);
            }
        } else {
            if (top + 1u < kPtMaxStack) stack[64u * top++] = d1;
            if (top < kPtMaxStack) stack[64u * top++] = d0;
        }
    }
    return best;
}

// The hit's blended normal (normalised) and colour (path_trace.wgsl:126-127).
__device__ __forceinline__ void hit_attributes(const PtHit& h, const float4* __restrict__ vt, V3& n, float4& c) {
    V3 p0, p1, p2, n0, n1, n2;
    float4 c0, c1, c2;
    load_vert(vt, h.i0, p0, n0, c0);
    load_vert(vt, h.i1, p1, n1, c1);
    load_vert(vt, h.i2, p2, n2, c2);
    const float u = h.u, v = h.v, w = h.w;
    n = normalize3(add(add(mul(n0, u), mul(n1, v)), mul(n2, w)));
    c = make_float4((u * c0.x + v * c1.x) + w * c2.x, (u * c0.y + v * c1.y) + w * c2.y,
                    (u * c0.z + v * c1.z) + w * c2.z, (u * c0.w + v * c1.w) + w * c2.w);
}

// Equirectangular environment, Z up (path_trace.wgsl:190-196): bilinear filtering
// of the texels at x = u W - 1/2, y = v H - 1/2, repeat in u, clamp in v.
__device__ __forceinline__ V3 sample_env(const float4* __restrict__ env, int ew, int eh, V3 d) {
    const float u = atan2_det(d.y, d.x) / (2.0f * kPtPi) + 0.5f;
    const float v = 0.5f - asin_det(clamp_w(d.z, -1.0f, 1.0f)) / kPtPi;
    const float3 r = pt_env_bilinear_raw(env, ew, eh, u, v);
    return v3(r.x, r.y, r.z);
}

// Cosine-weighted hemisphere about n (path_trace.wgsl:200-210).
__device__ __forceinline__ V3 cosine_hemisphere(V3 n, uint32_t& seed) {
    const float2 r = rand2(seed);
    const float phi = (2.0f * kPtPi) * r.x;
    const float sr = sqrtf(r.y);
    const float x = cos_det(phi) * sr, y = sin_det(phi) * sr, z = sqrtf(fmax_w(0.0f, 1.0f - r.y));
    V3 up = v3(0.0f, 1.0f, 0.0f);
    if (fabsf(n.y) > 0.99f) up = v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize3(cross3(up, n));
    const V3 b = cross3(n, t);
    return normalize3(add(add(mul(t, x), mul(b, y)), mul(n, z)));
}

}  // namespace

// One thread per pixel, 16x16 pixels per workgroup, each wave an 8x8 block.
__global__ __launch_bounds__(64) void k_pt_trace(PtArgs a) {
    __shared__ uint32_t s_stack[kPtMaxStack * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t px = blockIdx.x * 8u + (lane & 7u);
    const uint32_t py = blockIdx.y * 8u + (lane >> 3);
    if (px >= a.W || py >= a.H) return;
    const uint32_t pix = py * a.W + px;
    uint32_t seed = pcg(pix ^ (a.frame * 2654435761u));
    const float2 r0 = rand2(seed);
    const float jx = r0.x - 0.5f, jy = r0.y - 0.5f;
    const float ux = (((float)px + 0.5f) + jx) / (float)a.W;
    const float uy = (((float)py + 0.5f) + jy) / (float)a.H;
    const float nx = ux * 2.0f - 1.0f, ny = 1.0f - uy * 2.0f;
    const V3 ro = unproject(a.ivp, nx, ny, 0.0f);
    const V3 far = unproject(a.ivp, nx, ny, 1.0f);
    V3 rd = normalize3(sub(far, ro));
    V3 rad = v3(0.f, 0.f, 0.f), thr = v3(1.f, 1.f, 1.f), o = ro;
    for (int bounce = 0; bounce < 8; ++bounce) {
        const PtHit h = scene_hit<false>(o, rd, a.nodes, a.tris, a.verts, a.nnodes, s_stack + lane);
        if (!h.hit) {
            rad = add(rad, mulv(thr, sample_env(a.env, a.ew, a.eh, rd)));
            break;
        }
        V3 hn;
        float4 hc;
        hit_attributes(h, a.verts, hn, hc);
        const V3 n = dot3(hn, rd) > 0.0f ? v3(-hn.x, -hn.y, -hn.z) : hn;
        o = add(add(o, mul(rd, h.t)), mul(n, 5e-4f));
        if (hc.w < 0.5f) {
            rd = sub(rd, mul(n, 2.0f * dot3(n, rd)));  // reflect(rd, n)
        } else {
            thr = mulv(thr, v3(hc.x, hc.y, hc.z));
            rd = cosine_hemisphere(n, seed);
        }
    }
    float4 acc = a.accum[pix];
    acc.x = acc.x + rad.x;
    acc.y = acc.y + rad.y;
    acc.z = acc.z + rad.z;
    acc.w = acc.w + 1.0f;
    a.accum[pix] = acc;
}

// accum / max(count, 1), x4, Reinhard, rgba8 store (path_trace_resolve.wgsl:21-30).
__global__ __launch_bounds__(256) void k_pt_resolve(const float4* __restrict__ accum, uint32_t npix,
                                                    float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float4 s = accum[i];
    const float n = fmax_w(s.w, 1.0f);
    out[i] = make_float4(pt_unorm8(pt_reinhard((s.x / n) * 4.0f)), pt_unorm8(pt_reinhard((s.y / n) * 4.0f)),
                         pt_unorm8(pt_reinhard((s.z / n) * 4.0f)), 1.0f);
}

// Environment background of the raster target (envmap.wgsl:38-60): per pixel the
// direction from the near to the far point under the pixel centre's NDC
// (x = (px + 1/2) / W * 2 - 1, y = 1 - (py + 1/2) / H * 2), the equirectangular
// texel, x4, Reinhard, rgba8.
__global__ __launch_bounds__(256) void k_env_background(EnvArgs a) {
    const uint32_t px = blockIdx.x * 16u + (threadIdx.x & 15u), py = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (px >= a.W || py >= a.H) return;
    const float nx = (((float)px + 0.5f) / (float)a.W) * 2.0f - 1.0f;
    const float ny = 1.0f - (((float)py + 0.5f) / (float)a.H) * 2.0f;
    const V3 nw = unproject(a.ivp, nx, ny, 0.0f), fw = unproject(a.ivp, nx, ny, 1.0f);
    const V3 c = sample_env(a.env, a.ew, a.eh, normalize3(sub(fw, nw)));
    a.out[(size_t)py * a.W + px] = make_float4(pt_unorm8(pt_reinhard(c.x * 4.0f)), pt_unorm8(pt_reinhard(c.y * 4.0f)),
                                               pt_unorm8(pt_reinhard(c.z * 4.0f)), 1.0f);
}

// clip-space w of a point (row 3 of the column-major VP)
__device__ __forceinline__ float clip_w(const float* m, V3 p) {
    return ((m[3] * p.x + m[7] * p.y) + m[11] * p.z) + m[15] * 1.0f;
}
// view-space direction of a normal (view * vec4(n, 0)).xyz
__device__ __forceinline__ V3 view_dir(const float* m, V3 n) {
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = ((m[i] * n.x + m[4 + i] * n.y) + m[8 + i] * n.z) + m[12 + i] * 0.0f;
    return v3(r[0], r[1], r[2]);
}
__device__ __forceinline__ float view_depth(float w) {
    return clamp_w(1.0f - 0.1f / fmax_w(w, 0.1f), 0.0f, 1.0f);  // mesh.wgsl:42-43
}

// Per pixel: the envmap background, then the closest front-facing triangle
// along the pixel centre's ray shaded as mesh.wgsl frag (:55-74): normals
// colour (normalize(n) + 1) / 2, or the matcap at the view normal's xy (v
// flipped) times the vertex colour; Reinhard(4 c); depth the blended vertex
// depths 1 - 0.1 / max(w, 0.1).  No triangle: background colour, depth 1.
__global__ __launch_bounds__(256) void k_mesh_raster(RasterArgs a) {
    __shared__ uint32_t s_stack[kPtMaxStack * 64 * 4];
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint32_t px = blockIdx.x * 16u + (wid & 1u) * 8u + (lane & 7u);
    const uint32_t py = blockIdx.y * 16u + (wid >> 1) * 8u + (lane >> 3);
    if (px >= a.W || py >= a.H) return;
    const float nx = (((float)px + 0.5f) / (float)a.W) * 2.0f - 1.0f;
    const float ny = 1.0f - (((float)py + 0.5f) / (float)a.H) * 2.0f;
    const V3 nw = unproject(a.ivp, nx, ny, 0.0f), fw = unproject(a.ivp, nx, ny, 1.0f);
    const V3 rd = normalize3(sub(fw, nw));
    const size_t pix = (size_t)py * a.W + px;
    const PtHit h = scene_hit<true>(nw, rd, a.nodes, a.tris, a.verts, a.nnodes, s_stack + wid * 64u * kPtMaxStack + lane);
    if (!h.hit) {
        const V3 c = sample_env(a.env, a.ew, a.eh, rd);
        a.color[pix] = make_float4(pt_unorm8(pt_reinhard(c.x * 4.0f)), pt_unorm8(pt_reinhard(c.y * 4.0f)),
                                   pt_unorm8(pt_reinhard(c.z * 4.0f)), 1.0f);
        a.depth[pix] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        return;
    }
    V3 p[3], n[3];
    float4 c[3];
    load_vert(a.verts, h.i0, p[0], n[0], c[0]);
    load_vert(a.verts, h.i1, p[1], n[1], c[1]);
    load_vert(a.verts, h.i2, p[2], n[2], c[2]);
    const float u = h.u, v = h.v, w = h.w;
    const float d = (u * view_depth(clip_w(a.vp, p[0])) + v * view_depth(clip_w(a.vp, p[1]))) +
                    w * view_depth(clip_w(a.vp, p[2]));
    const V3 nn = add(add(mul(n[0], u), mul(n[1], v)), mul(n[2], w));
    V3 fc;
    if (a.shading < 0.5f) {
        const V3 q = normalize3(nn);
        fc = v3((q.x + 1.0f) * 0.5f, (q.y + 1.0f) * 0.5f, (q.z + 1.0f) * 0.5f);
    } else {
        const V3 vn = normalize3(
            add(add(mul(view_dir(a.view, n[0]), u), mul(view_dir(a.view, n[1]), v)), mul(view_dir(a.view, n[2]), w)));
        const float mu = vn.x * 0.5f + 0.5f, mv = vn.y * 0.5f + 0.5f;
        const float3 m = pt_env_bilinear_raw<false>(a.env, a.ew, a.eh, mu, 1.0f - mv);
        fc = v3(m.x * ((u * c[0].x + v * c[1].x) + w * c[2].x), m.y * ((u * c[0].y + v * c[1].y) + w * c[2].y),
                m.z * ((u * c[0].z + v * c[1].z) + w * c[2].z));
    }
    a.color[pix] = make_float4(pt_unorm8(pt_reinhard(fc.x * 4.0f)), pt_unorm8(pt_reinhard(fc.y * 4.0f)),
                               pt_unorm8(pt_reinhard(fc.z * 4.0f)), 1.0f);
    const float dq = pt_unorm8(d);
    a.depth[pix] = make_float4(dq, dq, dq, 1.0f);
}

void launch_mesh_raster(hipStream_t st, const RasterArgs& a) {
    if (a.W == 0 || a.H == 0) return;
    hipLaunchKernelGGL(k_mesh_raster, dim3((a.W + 15) / 16, (a.H + 15) / 16), dim3(256), 0, st, a);
}

void launch_env_background(hipStream_t st, const float* ivp, const float4* env, int ew, int eh, uint32_t W,
                           uint32_t H, float4* out) {
    if (W == 0 || H == 0) return;
    EnvArgs a;
    for (int k = 0; k < 16; ++k) a.ivp[k] = ivp[k];
    a.env = env;
    a.ew = ew;
    a.eh = eh;
    a.W = W;
    a.H = H;
    a.out = out;
    hipLaunchKernelGGL(k_env_background, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, st, a);
}

void launch_pt_trace(hipStream_t st, const PtArgs& a) {
    if (a.W == 0 || a.H == 0) return;
    hipLaunchKernelGGL(k_pt_trace, dim3((a.W + 7) / 8, (a.H + 7) / 8), dim3(64), 0, st, a);
}

void launch_pt_resolve(hipStream_t st, const float4* accum, uint32_t npix, float4* out) {
    if (npix == 0) return;
    hipLaunchKernelGGL(k_pt_resolve, dim3((npix + 255) / 256), dim3(256), 0, st, accum, npix, out);
}

}  // namespace nps
