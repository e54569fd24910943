// k_bezier.hip -- the Bezier stroke optimizer on gfx950: per-view curve
// preprocess, fused per-tile forward composite + loss seeds + reverse pass,
// Adam step and densification.  Semantics of bezier_backward.wgsl,
// bezier_step.wgsl and bezier_adc.wgsl, bit for bit.
//
// The reference bins every curve against every 16x16 tile in an LDS bitmask and
// walks each tile's first 1024 ascending ids per pixel, evaluating 16 polyline
// segments per (pixel, curve) in the forward and again in the reverse.  Here:
//   k_bz_preprocess : projects each curve once per view: the 17 polyline points
//                     (the bezier_at values every pair would recompute), the
//                     clip w's, the backprojection rows of the 4 control points,
//                     the polyline box, and the exact range of tiles whose
//                     rectangle passes the reference's float overlap test;
//                     the capped, id-ordered tile lists then come from the same
//                     cut / key / radix sort / range kernels as the splats;
//   k_bz_composite  : one 256-thread workgroup per tile, one thread per pixel,
//                     each wave an 8x8 block.  Records are staged in LDS 64 at a
//                     time.  An entry whose polyline box is farther from the
//                     wave's block than width + softness (with a margin that
//                     covers every float rounding of the distance) has alpha
//                     exactly 0 on all 64 pixels and is not evaluated.  Inside
//                     the segment loop the division is skipped where the clamp
//                     decides it (num <= 0, num >= len2) and the square root
//                     where the squared distance cannot win the strict minimum.
//                     The reverse recomputes alphas from transmittance
//                     checkpoints every 8 entries (the reference keeps 1024
//                     alphas per thread) and reduces the 18 fixed-point
//                     gradients and the loss contribution of each pair over the
//                     wave with integer transposes before one LDS atomic per
//                     value; one global atomic per (curve, parameter) per tile.
//   k_bz_step       : Adam, clamps, prune, off-screen kill, one thread per curve.
//   k_bz_adc        : one 1024-thread workgroup: dead list, kills, candidate
//                     lists, pixel scans and resets as block scans; the two
//                     split/clone walks and each spawn's bookkeeping stay
//                     sequential, in the reference's order.
#include "nps_bezier.h"
#include "nps_common.h"
#include "nps_kernels.h"
#include "nps_wave.h"

namespace nps {

// ------------------------------------------------------------------ helpers
struct V2 {
    float x, y;
};

// bezier_backward.wgsl:56-62
__host__ __device__ __forceinline__ V2 bz_at(V2 p0, V2 p1, V2 p2, V2 p3, float t) {
    float omt = 1.0f - t;
    float c0 = omt * omt * omt, c1 = 3.0f * omt * omt * t, c2 = 3.0f * omt * t * t, c3 = t * t * t;
    V2 r;
    r.x = ((p0.x * c0 + p1.x * c1) + p2.x * c2) + p3.x * c3;
    r.y = ((p0.y * c0 + p1.y * c1) + p2.y * c2) + p3.y * c3;
    return r;
}
// :64-67
__device__ __forceinline__ void bz_bern(float t, float B[4]) {
    float omt = 1.0f - t;
    B[0] = omt * omt * omt;
    B[1] = 3.0f * omt * omt * t;
    B[2] = 3.0f * omt * t * t;
    B[3] = t * t * t;
}
// WGSL smoothstep(e0, e1, x)
__device__ __forceinline__ float bz_smoothstep(float e0, float e1, float x) {
    float t = clamp_w((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
// float -> int64 toward zero, saturating, NaN -> 0
__device__ __forceinline__ uint64_t bz_f2l(float x) {
    if (x != x) return 0ull;
    if (x >= 9223372036854775808.0f) return 0x7FFFFFFFFFFFFFFFull;
    if (x <= -9223372036854775808.0f) return 0x8000000000000000ull;
    return (uint64_t)(int64_t)x;
}
__device__ __forceinline__ uint32_t bz_gi(float g, float fp) { return (uint32_t)f2i_hw(g * fp); }

// ---------------------------------------------------------------- preprocess
__global__ __launch_bounds__(256) void k_bz_preprocess(const float4* __restrict__ curves, uint32_t n, ViewParams v,
                                                       BzFwd* __restrict__ fwd, BzBwd* __restrict__ bwd,
                                                       uint32_t* __restrict__ counts, uint4* __restrict__ rects) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = curves[5 * (size_t)i + k];
    const float4 col = curves[5 * (size_t)i + 4];
    counts[i] = 0u;
    if (col.w < kBzKill) return;  // BW:128
    const float aspect = v.aspect;
    F4 k4[4];
    V2 p[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        k4[k] = mat_vec(v.vp, c[k].x, c[k].y, c[k].z, 1.0f);
        p[k].x = k4[k].x / k4[k].w * aspect;  // project_center (BW:69-72)
        p[k].y = k4[k].y / k4[k].w;
    }
    if (k4[0].w <= 0.0f || k4[1].w <= 0.0f || k4[2].w <= 0.0f || k4[3].w <= 0.0f) return;  // BW:139
    const float width = fmax_w(c[0].w, 0.001f), soft = fmax_w(c[1].w, 0.001f);
    const float oc = width + soft;
    const float mnx = fmin_w(fmin_w(p[0].x, p[1].x), fmin_w(p[2].x, p[3].x)) - oc;
    const float mny = fmin_w(fmin_w(p[0].y, p[1].y), fmin_w(p[2].y, p[3].y)) - oc;
    const float mxx = fmax_w(fmax_w(p[0].x, p[1].x), fmax_w(p[2].x, p[3].x)) + oc;
    const float mxy = fmax_w(fmax_w(p[0].y, p[1].y), fmax_w(p[2].y, p[3].y)) + oc;
    // tiles whose rectangle passes BW:150-151, as index ranges (the rectangle
    // edges are monotone in the tile index)
    const uint32_t W = v.W, H = v.H, T = (uint32_t)kTile;
    const uint4 tr = tile_rect(mnx, mxx, mny, mxy, W, H, T, v.tiles_x, v.tiles_y, aspect);
    const uint32_t tx0 = tr.x, tx_end = tr.y, ty0 = tr.z, ty_end = tr.w;
    if (tx0 >= tx_end || ty0 >= ty_end) return;
    counts[i] = (tx_end - tx0) * (ty_end - ty0);
    rects[i] = make_uint4(tx0, tx_end, ty0, ty_end);

    BzFwd f;
    f.x[0] = p[0].x;
    f.y[0] = p[0].y;
#pragma unroll
    for (int k = 1; k <= kBzSeg; ++k) {
        V2 q = bz_at(p[0], p[1], p[2], p[3], (float)k / (float)kBzSeg);
        f.x[k] = q.x;
        f.y[k] = q.y;
    }
    float bx0 = f.x[0], bx1 = f.x[0], by0 = f.y[0], by1 = f.y[0];
#pragma unroll
    for (int k = 1; k <= kBzSeg; ++k) {
        bx0 = fmin_w(bx0, f.x[k]);
        bx1 = fmax_w(bx1, f.x[k]);
        by0 = fmin_w(by0, f.y[k]);
        by1 = fmax_w(by1, f.y[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) f.w[k] = k4[k].w;
    f.width = width;
    f.soft = soft;
    f.op = col.w;
    f.r = col.x;
    f.g = col.y;
    f.b = col.z;
    f.cx0 = bx0;
    f.cy0 = by0;
    f.cx1 = bx1;
    f.cy1 = by1;
    // every pixel farther than rcull from the box has a computed distance above
    // width + softness >= the outer radius at any pressure, so alpha = 0: the
    // relative and absolute margins exceed the rounding of the distance by
    // orders of magnitude
    const float M = fmax_w(fmax_w(fabsf(bx0), fabsf(bx1)), fmax_w(fabsf(by0), fabsf(by1)));
    f.rcull = (width + soft) * 1.0001f + 1e-5f * (4.0f + M);
    f.pad[0] = f.pad[1] = f.pad[2] = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        float gx0 = f.x[4 * g], gx1 = f.x[4 * g], gy0 = f.y[4 * g], gy1 = f.y[4 * g];
#pragma unroll
        for (int j = 1; j <= 4; ++j) {
            gx0 = fmin_w(gx0, f.x[4 * g + j]);
            gx1 = fmax_w(gx1, f.x[4 * g + j]);
            gy0 = fmin_w(gy0, f.y[4 * g + j]);
            gy1 = fmax_w(gy1, f.y[4 * g + j]);
        }
        f.gb[4 * g + 0] = gx0;
        f.gb[4 * g + 1] = gy0;
        f.gb[4 * g + 2] = gx1;
        f.gb[4 * g + 3] = gy1;
    }
    fwd[i] = f;

    BzBwd b;
    V2 q0 = bz_at(p[0], p[1], p[2], p[3], 0.0f);
    b.q0x = q0.x;
    b.q0y = q0.y;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float w = k4[k].w, w2 = w * w;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float v0 = v.vp[4 * ax + 0], v1 = v.vp[4 * ax + 1], v3 = v.vp[4 * ax + 3];
            b.J0[3 * k + ax] = aspect * (v0 * w - k4[k].x * v3) / w2;  // BW:74-88
            b.J1[3 * k + ax] = (v1 * w - k4[k].y * v3) / w2;
        }
    }
    b.pad[0] = b.pad[1] = 0.0f;
    bwd[i] = b;
}

void launch_bz_preprocess(hipStream_t st, const float* curves, uint32_t n, const ViewParams& v, BzFwd* fwd,
                          BzBwd* bwd, uint32_t* counts, uint4* rects) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_bz_preprocess, dim3((n + 255) / 256), dim3(256), 0, st, (const float4*)curves, n, v, fwd,
                       bwd, counts, rects);
}

// ------------------------------------------------------------- composite
struct BzEval {
    float a, dval;
    uint32_t k;
};

// The wave's 8x8 pixel block in screen space and the squared distance beyond
// which every alpha of the entry is exactly 0 (rcull plus the block's margin).
struct BzBlock {
    float xa, xb, ya, yb;
};

// Forward of one (pixel, curve) pair, BW:196-256.  The minimum over the 16
// segments is taken with the reference's strict `<` on the distances.  Exact
// shortcuts, each leaving alpha (and, where alpha > 0, the segment, its
// parameter and the depth) unchanged:
//  - a group of 4 segments whose box is farther than the cull radius from the
//    wave's block is skipped: all their distances exceed width + softness, so
//    none is the minimum of a pixel with alpha > 0, and a pixel whose minimum
//    lies there has alpha 0 either way (the test is conservative, so it uses
//    the native min/max);
//  - the square root is only taken where d^2 < (d^2 of the current minimum),
//    since otherwise sqrt(d^2) >= the current minimum (sqrt is monotone and
//    correctly rounded);
//  - u = clamp(num / len2, 0, 1) is 0 for num <= 0 and 1 for num >= len2
//    (finite len2) without the division;
//  - a minimum distance >= fl(width + softness) >= the outer radius at any
//    pressure gives alpha 0 without the taper and depth terms.
__device__ __forceinline__ BzEval bz_eval(const BzFwd& f, float px, float py, const BzBlock& bk, float thr2) {
    float min_d = 1e9f, min_d2 = __builtin_inff(), min_u = 0.0f;
    uint32_t min_k = 1u;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
    const float gx = __builtin_fmaxf(__builtin_fmaxf(f.gb[4 * g + 0] - bk.xb, bk.xa - f.gb[4 * g + 2]), 0.0f);
    const float gy = __builtin_fmaxf(__builtin_fmaxf(f.gb[4 * g + 1] - bk.yb, bk.ya - f.gb[4 * g + 3]), 0.0f);
    // the operands are the same on every lane (LDS broadcast, the wave's
    // block): a scalar branch
    if (__builtin_amdgcn_readfirstlane((int)(gx * gx + gy * gy > thr2))) continue;
    float prevx = f.x[4 * g], prevy = f.y[4 * g];
#pragma unroll
    for (int k = 4 * g + 1; k <= 4 * g + 4; ++k) {
        const float cx = f.x[k], cy = f.y[k];
        const float sx = cx - prevx, sy = cy - prevy;
        const float len2 = fmax_w(sx * sx + sy * sy, 1e-8f);
        const float qx = px - prevx, qy = py - prevy;
        const float num = qx * sx + qy * sy;
        float u;
        if (num <= 0.0f) u = 0.0f;
        else if (num >= len2 && len2 < __builtin_inff()) u = 1.0f;
        else u = clamp_w(num / len2, 0.0f, 1.0f);
        const float ox = prevx + sx * u, oy = prevy + sy * u;
        const float dx = px - ox, dy = py - oy;
        const float d2 = dx * dx + dy * dy;
        if (d2 < min_d2) {
            const float d = sqrtf(d2);
            if (d < min_d) {
                min_d = d;
                min_d2 = d2;
                min_k = (uint32_t)k;
                min_u = u;
            }
        }
        prevx = cx;
        prevy = cy;
    }
    }
    BzEval e;
    if (min_d >= f.width + f.soft) {
        e.a = 0.0f;
        e.dval = 0.0f;
        e.k = min_k;
        return e;
    }
    const float t = ((float)(min_k - 1u) + min_u) / (float)kBzSeg;
    const float dt = t - 0.5f;
    const float pressure = 1.0f - 4.0f * dt * dt;
    const float lw = f.width * pressure, ls = f.soft * pressure, lo = f.op * pressure;
    const float a_geom = 1.0f - bz_smoothstep(lw - ls, lw + ls, min_d);
    e.a = clamp_w(a_geom * lo, 0.0f, 0.999f);
    float B[4];
    bz_bern(t, B);
    const float raw_w = ((B[0] * f.w[0] + B[1] * f.w[1]) + B[2] * f.w[2]) + B[3] * f.w[3];
    const float lin = fmax_w(raw_w, 0.1f);
    e.dval = clamp_w(1.0f - 0.1f / lin, 0.0f, 1.0f);
    e.k = min_k;
    return e;
}

// Loss seeds of one pixel, BW:259-320, and its fine-layer flow direction (:449-480).
struct BzSeeds {
    float dC[3];
    float dT;
    float dD;      // dD_total
    float is_bg;   // step(0.995, tgt_depth)
    float fx, fy;  // flow direction
    bool have_flow;
};

__device__ __forceinline__ float bz_luma(const float4* t, int W, int H, int x, int y) {
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    float4 c = t[(size_t)y * W + x];
    return (c.x * 0.2126f + c.y * 0.7152f) + c.z * 0.0722f;
}
__device__ __forceinline__ float bz_avg(const float4* t, int W, int H, int x, int y) {
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    float4 c = t[(size_t)y * W + x];
    return (c.x * 0.333f + c.y * 0.333f) + c.z * 0.333f;
}

// Reverse-pass body of one active (pixel, curve) pair, BW:330-562: the 18
// gradients as the reference's i32 conversions (direction and main parts
// converted separately, as its two atomicAdds are) and the 16-bit pieces of the
// pair's 2^-32 loss contribution.
__device__ __forceinline__ void bz_pair_grads(const BzFwd& f, const BzBwd& bb, float px, float py, uint32_t k,
                                              float a, float T_prev, float d_val, bool is_fine, float bg_penalty,
                                              BzSeeds& s, uint32_t (&g16)[16], uint32_t (&g8)[8]) {
    const float width = f.width, soft = f.soft, opacity = f.op;
    const float cr = f.r, cg = f.g, cb = f.b;
    float dColor[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dColor[c] = s.dC[c] * (T_prev * a);
    const float da = (s.dT * (-T_prev) + ((s.dC[0] * (T_prev * cr) + s.dC[1] * (T_prev * cg)) + s.dC[2] * (T_prev * cb))) +
                     s.dD * (T_prev * d_val);
    s.dT = (s.dT * (1.0f - a) + ((s.dC[0] * (a * cr) + s.dC[1] * (a * cg)) + s.dC[2] * (a * cb))) + s.dD * (a * d_val);
    const float t_prev = (float)(k - 1u) / (float)kBzSeg, t_curr = (float)k / (float)kBzSeg;
    const float ppx = (k == 1u) ? bb.q0x : f.x[k - 1], ppy = (k == 1u) ? bb.q0y : f.y[k - 1];
    const float cpx = f.x[k], cpy = f.y[k];
    const float segx = cpx - ppx, segy = cpy - ppy;
    const float len2 = fmax_w(segx * segx + segy * segy, 1e-8f);
    const float uc = clamp_w(((px - ppx) * segx + (py - ppy) * segy) / len2, 0.0f, 1.0f);
    const float ox = ppx + segx * uc, oy = ppy + segy * uc;
    const float dvx = px - ox, dvy = py - oy;
    const float d = fmax_w(sqrtf(dvx * dvx + dvy * dvy), 1e-6f);
    const float t_pix = ((float)(k - 1u) + uc) / (float)kBzSeg;
    const float dtp = t_pix - 0.5f;
    const float pressure = 1.0f - 4.0f * dtp * dtp;
    float Bp[4];
    bz_bern(t_pix, Bp);
    const float lwid = width * pressure, lsoft = soft * pressure, lop = opacity * pressure;
    const float inner = lwid - lsoft, outer = lwid + lsoft;
    const float denom = fmax_w(outer - inner, 1e-6f);
    const float x_in = clamp_w((d - inner) / denom, 0.0f, 1.0f);
    const float ss_d = 6.0f * x_in * (1.0f - x_in) / denom;
    const bool in_band = (d > inner) && (d < outer);
    const float a_geom = 1.0f - bz_smoothstep(inner, outer, d);
    const float d_op = da * a_geom * pressure + bg_penalty * s.is_bg;
    const float da_eff = da * lop;
    const float dDd = in_band ? -da_eff * ss_d : 0.0f;
    const float dWidth = in_band ? da_eff * ss_d * pressure / denom : 0.0f;
    const float dmi = clamp_w(d - inner, 0.0f, denom);
    const float dx_ds = pressure * (denom - 2.0f * dmi) / fmax_w(denom * denom, 1e-12f);
    float dSoft = in_band ? -da_eff * ss_d * dx_ds : 0.0f;
    const float raw_w = ((Bp[0] * f.w[0] + Bp[1] * f.w[1]) + Bp[2] * f.w[2]) + Bp[3] * f.w[3];
    const float lin = fmax_w(raw_w, 0.1f);
    const float dd_dw = 0.1f / (lin * lin);
    const float dProjx = -dDd * dvx / d, dProjy = -dDd * dvy / d;
    const float dPrevx = dProjx * (1.0f - uc), dPrevy = dProjy * (1.0f - uc);
    const float dCurrx = dProjx * uc, dCurry = dProjy * uc;
    float Bpr[4], Bcu[4];
    bz_bern(t_prev, Bpr);
    bz_bern(t_curr, Bcu);
    if (is_fine) dSoft = dSoft + 5.0f * 2.0f * soft;
    // direction regulariser (fine layer, BW:446-521)
    uint32_t gdir[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) gdir[q] = 0u;
    if (is_fine && len2 > 1e-10f && s.have_flow) {
        const float il = sqrtf(len2);
        const float tgx = segx / il, tgy = segy / il;
        const float tgd = tgx * s.fx + tgy * s.fy;
        const float dld = 1.5f * 2.0f * tgd;
        const float dtvx = s.fx * dld, dtvy = s.fy * dld;
        const float inv_len = 1.0f / sqrtf(len2);
        const float dtt = dtvx * tgx + dtvy * tgy;
        const float dsx = (dtvx - tgx * dtt) * inv_len, dsy = (dtvy - tgy * dtt) * inv_len;
        const float dpx = -dsx, dpy = -dsy;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float gx = dpx * Bpr[c] + dsx * Bcu[c], gy = dpy * Bpr[c] + dsy * Bcu[c];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) gdir[3 * c + ax] = bz_gi(gx * bb.J0[3 * c + ax] + gy * bb.J1[3 * c + ax], 10000.0f);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float gx = dPrevx * Bpr[c] + dCurrx * Bcu[c], gy = dPrevy * Bpr[c] + dCurry * Bcu[c];
        const float dZ = s.dD * (T_prev * a) * dd_dw * Bp[c];
        const float g3x = gx * bb.J0[3 * c + 0] + gy * bb.J1[3 * c + 0];
        const float g3y = gx * bb.J0[3 * c + 1] + gy * bb.J1[3 * c + 1];
        const float g3z = gx * bb.J0[3 * c + 2] + gy * bb.J1[3 * c + 2];
        g16[3 * c + 0] = gdir[3 * c + 0] + bz_gi(g3x, 10000.0f);
        g16[3 * c + 1] = gdir[3 * c + 1] + bz_gi(g3y, 10000.0f);
        g16[3 * c + 2] = gdir[3 * c + 2] + bz_gi(g3z + dZ, 10000.0f);
    }
    g16[12] = bz_gi(dColor[0], 100000.0f);
    g16[13] = bz_gi(dColor[1], 100000.0f);
    g16[14] = bz_gi(dColor[2], 100000.0f);
    g16[15] = bz_gi(d_op, 100000.0f);
    g8[0] = bz_gi(dWidth, 10000.0f);
    g8[1] = bz_gi(dSoft, 10000.0f);
    const float lc = ((s.dC[0] * s.dC[0] + s.dC[1] * s.dC[1]) + s.dC[2] * s.dC[2]) * (T_prev * a);
    const uint64_t l = bz_f2l(lc * 4294967296.0f);
    g8[2] = (uint32_t)(l & 0xFFFFu);
    g8[3] = (uint32_t)((l >> 16) & 0xFFFFu);
    g8[4] = (uint32_t)((l >> 32) & 0xFFFFu);
    g8[5] = (uint32_t)(l >> 48);
    g8[6] = 0u;
    g8[7] = 0u;
}

__device__ __forceinline__ void bz_stage(const BzFwd* __restrict__ fwd, const BzBwd* __restrict__ bwd,
                                         const uint64_t* __restrict__ keys, uint32_t base, int nb, BzFwd* s_f,
                                         BzBwd* s_b, uint32_t* s_ids, bool with_bwd) {
    const int t = threadIdx.x;
    constexpr int FQ = (int)(sizeof(BzFwd) / 16), BQ = (int)(sizeof(BzBwd) / 16);
    for (int e = t; e < nb * FQ; e += 256) {
        const int j = e / FQ, part = e - j * FQ;
        const uint32_t id = (uint32_t)keys[base + j];
        reinterpret_cast<float4*>(&s_f[j])[part] = reinterpret_cast<const float4*>(&fwd[id])[part];
        if (part == 0) s_ids[j] = id;
    }
    if (with_bwd) {
        for (int e = t; e < nb * BQ; e += 256) {
            const int j = e / BQ, part = e - j * BQ;
            const uint32_t id = (uint32_t)keys[base + j];
            reinterpret_cast<float4*>(&s_b[j])[part] = reinterpret_cast<const float4*>(&bwd[id])[part];
        }
    }
}

// Forward half: per tile the front-to-back composite (BW:188-257), images,
// loss, transmittance checkpoints, the pair buffers and activity masks of the
// reverse, and the per-pixel loss seeds (BW:259-320), which the reverse kernel
// reads back (seeds, masks and pair values are written only with `backward`).
__global__ __launch_bounds__(256) void k_bz_forward(ViewParams v, BzCompositeArgs A) {
    __shared__ BzFwd s_f[kBatch];
    __shared__ uint32_t s_ids[kBatch];
    __shared__ double s_red[2][4];
    __shared__ uint32_t s_cnt[4];

    const uint32_t tile = xcd_remap(blockIdx.x, v.num_tiles);
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t start = A.ranges[2 * tile], end = A.ranges[2 * tile + 1];
    const int count = (int)min(end - start, A.cap);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t bx = tx * kTile + (wid & 1) * 8, by = ty * kTile + (wid >> 1) * 8;
    const uint32_t px = bx + (lane & 7), py = by + (lane >> 3);
    const bool valid = px < v.W && py < v.H;
    const float pxf = px_to_x(px, v.W, v.aspect);
    const float pyf = py_to_y(py, v.H);
    // the wave's 8x8 block in screen space (x increases, y decreases with the pixel index)
    const float xa = px_to_x(bx, v.W, v.aspect), xb = px_to_x(bx + 7, v.W, v.aspect);
    const float ya = py_to_y(by + 7, v.H), yb = py_to_y(by, v.H);
    const float blk_margin = 1e-5f * (((fabsf(xa) + fabsf(xb)) + fabsf(ya)) + fabsf(yb));
    const BzBlock blk = {xa, xb, ya, yb};
    float* ckpt = A.ckpt + (size_t)tile * A.nchunk_max * kTilePx;
    const bool is_fine = A.prm.max_width > 0.0f;
    const bool color_mode = A.prm.mode > 0.5f;

    auto thr2_of = [&](const BzFwd& f) {
        const float thr = f.rcull + blk_margin;
        return thr * thr;
    };
    auto culled = [&](const BzFwd& f, float thr2) {
        const float gx = __builtin_fmaxf(__builtin_fmaxf(f.cx0 - xb, xa - f.cx1), 0.0f);
        const float gy = __builtin_fmaxf(__builtin_fmaxf(f.cy0 - yb, ya - f.cy1), 0.0f);
        return __builtin_amdgcn_readfirstlane((int)(gx * gx + gy * gy > thr2)) != 0;
    };

    // ---------------- forward (BW:188-257) ----------------
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, D = 0.0f;
    uint32_t actword = 0u, nzword = 0u;
    uint32_t n_feval = 0u;
    for (int b0 = 0; b0 < count; b0 += kBatch) {
        const int nb = min(kBatch, count - b0);
        __syncthreads();
        bz_stage(A.fwd, A.bwd, A.keys, start + (uint32_t)b0, nb, s_f, nullptr, s_ids, false);
        __syncthreads();
        for (int j = 0; j < nb; ++j) {
            const int e = b0 + j;
            if ((e & (kChunk - 1)) == 0) ckpt[(size_t)(e / kChunk) * kTilePx + tid] = T;
            const BzFwd& f = s_f[j];
            float a = 0.0f, dval = 0.0f;
            const float thr2 = thr2_of(f);
            if (!culled(f, thr2)) {
                n_feval += 1u;
                BzEval ev = bz_eval(f, pxf, pyf, blk, thr2);
                a = ev.a;
                dval = ev.dval;
                // the reverse reloads the pairs of entries where some lane has alpha > 0
                // (all-zero entries leave the transmittance unchanged)
                if (A.backward && __ballot(a != 0.0f) != 0ull) {
                    nzword |= 1u << (e & 31);
                    const size_t q = (size_t)(start + (uint32_t)e) * kTilePx + tid;
                    A.pa[q] = a;
                    A.pd[q] = dval;
                    A.pk[q] = (uint8_t)ev.k;
                }
            }
            if (__ballot(valid && !(a < 0.001f)) != 0ull) actword |= 1u << (e & 31);
            if ((e & 31) == 31 || e == count - 1) {
                if (lane == 0 && A.backward) {
                    const size_t mo = ((size_t)tile * 4 + wid) * (kBzCap / 32) + (e >> 5);
                    A.act[mo] = actword;
                    A.nzm[mo] = nzword;
                }
                actword = 0u;
                nzword = 0u;
            }
            const float Ta = T * a;
            C0 = C0 + Ta * f.r;
            C1 = C1 + Ta * f.g;
            C2 = C2 + Ta * f.b;
            D = D + Ta * dval;
            T = T * (1.0f - a);
        }
    }

    // ---------------- background and loss seeds (BW:259-320) ----------------
    const int W = (int)v.W, H = (int)v.H;
    const size_t pix = valid ? (size_t)py * v.W + px : 0;
    float4 tg = make_float4(0.f, 0.f, 0.f, 0.f);
    float td = 1.0f, bgc0 = 0.0f, bgc1 = 0.0f, bgc2 = 0.0f, bgd = 1.0f;
    if (valid) {
        tg = A.tgt[pix];
        td = A.tgt_depth[pix];
        if (color_mode) {
            float4 bgs = A.bg[pix];
            bgc0 = bgs.x;
            bgc1 = bgs.y;
            bgc2 = bgs.z;
            bgd = A.bg_depth[pix];
        }
    }
    const float Tn = T;
    C0 = C0 + Tn * bgc0;
    C1 = C1 + Tn * bgc1;
    C2 = C2 + Tn * bgc2;
    D = D + Tn * bgd;
    double l2 = 0.0, l1 = 0.0;
    if (valid) {
        A.c_img[3 * pix + 0] = C0;
        A.c_img[3 * pix + 1] = C1;
        A.c_img[3 * pix + 2] = C2;
        A.d_img[pix] = D;
        A.t_img[pix] = Tn;
        double e0 = (double)C0 - (double)tg.x, e1 = (double)C1 - (double)tg.y, e2 = (double)C2 - (double)tg.z;
        l2 = (e0 * e0 + e1 * e1) + e2 * e2;
        l1 = (fabs(e0) + fabs(e1)) + fabs(e2);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        l2 += __shfl_xor(l2, o, 64);
        l1 += __shfl_xor(l1, o, 64);
    }
    if (lane == 0) {
        s_red[0][wid] = l2;
        s_red[1][wid] = l1;
    }
    if (A.backward) {
        const float L0 = 0.2126f, L1 = 0.7152f, L2 = 0.0722f;
        const float dr0 = 2.0f * (C0 - tg.x), dr1 = 2.0f * (C1 - tg.y), dr2 = 2.0f * (C2 - tg.z);
        const bool fg = td < 0.99f;
        BzSeeds s;
        s.dD = (fg && color_mode) ? 0.5f * 2.0f * (D - td) : 0.0f;
        const float h0 = dr0 * 0.5f, h1 = dr1 * 0.5f, h2 = dr2 * 0.5f;
        const float luma_err = (h0 * L0 + h1 * L1) + h2 * L2;
        const int ix = (int)px, iy = (int)py;
        float contrast_w = 1.0f;
        s.fx = s.fy = 0.0f;
        s.have_flow = false;
        if (valid) {
            const float lr = bz_luma(A.tgt, W, H, ix + 1, iy), ll = bz_luma(A.tgt, W, H, ix - 1, iy);
            const float lu = bz_luma(A.tgt, W, H, ix, iy + 1), ld = bz_luma(A.tgt, W, H, ix, iy - 1);
            const float contrast = sqrtf((lr - ll) * (lr - ll) + (lu - ld) * (lu - ld));
            contrast_w = 1.0f + clamp_w(contrast * 8.0f, 0.0f, 3.0f);
            if (is_fine) {
                const float c_r = bz_avg(A.tgt, W, H, ix + 2, iy), c_l = bz_avg(A.tgt, W, H, ix - 2, iy);
                const float c_u = bz_avg(A.tgt, W, H, ix, iy + 2), c_d = bz_avg(A.tgt, W, H, ix, iy - 2);
                const float gcx = (c_r - c_l) * 0.25f * v.aspect, gcy = -(c_u - c_d) * 0.25f;
                const float n_r = bz_avg(A.normal, W, H, ix + 2, iy), n_l = bz_avg(A.normal, W, H, ix - 2, iy);
                const float n_u = bz_avg(A.normal, W, H, ix, iy + 2), n_d = bz_avg(A.normal, W, H, ix, iy - 2);
                const float gnx = (n_r - n_l) * 0.25f * v.aspect, gny = -(n_u - n_d) * 0.25f;
                const float frx = gcx + gnx * 2.0f, fry = gcy + gny * 2.0f;
                const float fl = sqrtf(frx * frx + fry * fry);
                if (fl > 1e-4f) {
                    s.fx = frx / fl;
                    s.fy = fry / fl;
                    s.have_flow = true;
                }
            }
        }
        const float dr_l = (dr0 * L0 + dr1 * L1) + dr2 * L2;
        s.dC[0] = (luma_err * L0 * 6.0f + (dr0 - dr_l * L0)) * contrast_w;
        s.dC[1] = (luma_err * L1 * 6.0f + (dr1 - dr_l * L1)) * contrast_w;
        s.dC[2] = (luma_err * L2 * 6.0f + (dr2 - dr_l * L2)) * contrast_w;
        s.dT = ((s.dC[0] * bgc0 + s.dC[1] * bgc1) + s.dC[2] * bgc2) + s.dD * bgd;
        const float coverage = 1.0f - Tn;
        const float d_cov = A.prm.mode < 0.5f ? 2.0f * 2.0f * (coverage - tg.x) : 0.0f;
        s.dT = s.dT + -d_cov;
        s.is_bg = td >= 0.995f ? 1.0f : 0.0f;
        if (valid) {
            const float residual = ((h0 * h0 + h1 * h1) + h2 * h2) * Tn;
            A.pixel_loss[pix] = (int32_t)((uint32_t)A.pixel_loss[pix] + (uint32_t)f2i_hw(residual * 10000.0f));
        }

        const size_t so = (size_t)tile * kTilePx + tid, ss = (size_t)v.num_tiles * kTilePx;
        A.seeds[so + 0 * ss] = s.dC[0];
        A.seeds[so + 1 * ss] = s.dC[1];
        A.seeds[so + 2 * ss] = s.dC[2];
        A.seeds[so + 3 * ss] = s.dT;
        A.seeds[so + 4 * ss] = s.dD;
        A.seeds[so + 5 * ss] = s.is_bg;
        A.seeds[so + 6 * ss] = s.fx;
        A.seeds[so + 7 * ss] = s.fy;
        A.seeds[so + 8 * ss] = s.have_flow ? 1.0f : 0.0f;
    }
    if (lane == 0) s_cnt[wid] = n_feval;
    __syncthreads();
    if (tid == 0) {
        double a2 = (s_red[0][0] + s_red[0][1]) + (s_red[0][2] + s_red[0][3]);
        double a1 = (s_red[1][0] + s_red[1][1]) + (s_red[1][2] + s_red[1][3]);
        A.tile_loss[2 * tile] = a2;
        A.tile_loss[2 * tile + 1] = a1;
        uint32_t vw = min(v.W - tx * kTile, (uint32_t)kTile), vh = min(v.H - ty * kTile, (uint32_t)kTile);
        // the tile's counters in its slot of A.part (summed by launch_sum_parts)
        unsigned long long* pc = A.part + 5 * (size_t)tile;
        pc[0] = (unsigned long long)count * (unsigned long long)(vw * vh);
        pc[1] = (unsigned long long)s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    }
}

// Reverse half (BW:325-563) from the forward's seeds, checkpoints, masks and
// pair buffers: gradients and loss contributions of every active pair.
__global__ __launch_bounds__(256) void k_bz_reverse(ViewParams v, BzCompositeArgs A) {
    __shared__ BzFwd s_f[kBatch];
    __shared__ BzBwd s_b[kBatch];
    __shared__ uint32_t s_ids[kBatch];
    __shared__ uint32_t s_grad[kBatch][kBzParams];
    __shared__ uint32_t s_lossp[kBatch][4];
    __shared__ uint32_t s_act[4][kBzCap / 32];
    __shared__ uint32_t s_nz[4][kBzCap / 32];
    __shared__ unsigned long long s_cnt[3][4];

    const uint32_t tile = xcd_remap(blockIdx.x, v.num_tiles);
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t start = A.ranges[2 * tile], end = A.ranges[2 * tile + 1];
    const int count = (int)min(end - start, A.cap);
    if (count == 0) {
        if (threadIdx.x < 3) A.part[5 * (size_t)tile + 2 + threadIdx.x] = 0ull;
        return;
    }
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t bx = tx * kTile + (wid & 1) * 8, by = ty * kTile + (wid >> 1) * 8;
    const uint32_t px = bx + (lane & 7), py = by + (lane >> 3);
    const bool valid = px < v.W && py < v.H;
    const float pxf = px_to_x(px, v.W, v.aspect);
    const float pyf = py_to_y(py, v.H);
    float* ckpt = A.ckpt + (size_t)tile * A.nchunk_max * kTilePx;
    const bool is_fine = A.prm.max_width > 0.0f;
    const int nwords = (count + 31) >> 5;
    for (int e = tid; e < 4 * (kBzCap / 32); e += 256) {
        const int w = e / (kBzCap / 32), k = e - w * (kBzCap / 32);
        if (k < nwords) {
            const size_t mo = ((size_t)tile * 4 + w) * (kBzCap / 32) + k;
            s_act[w][k] = A.act[mo];
            s_nz[w][k] = A.nzm[mo];
        }
    }
    BzSeeds s;
    {
        const size_t so = (size_t)tile * kTilePx + tid, ss = (size_t)v.num_tiles * kTilePx;
        s.dC[0] = A.seeds[so + 0 * ss];
        s.dC[1] = A.seeds[so + 1 * ss];
        s.dC[2] = A.seeds[so + 2 * ss];
        s.dT = A.seeds[so + 3 * ss];
        s.dD = A.seeds[so + 4 * ss];
        s.is_bg = A.seeds[so + 5 * ss];
        s.fx = A.seeds[so + 6 * ss];
        s.fy = A.seeds[so + 7 * ss];
        s.have_flow = A.seeds[so + 8 * ss] != 0.0f;
    }
    unsigned long long n_active = 0, n_went = 0, n_reval = 0;
        // ---------------- reverse pass (BW:325-563) ----------------
        const int nbatch = (count + kBatch - 1) / kBatch;
        for (int bi = nbatch - 1; bi >= 0; --bi) {
            const int b0 = bi * kBatch;
            const int nb = min(kBatch, count - b0);
            __syncthreads();
            bz_stage(A.fwd, A.bwd, A.keys, start + (uint32_t)b0, nb, s_f, s_b, s_ids, true);
            for (int e = tid; e < kBatch * kBzParams; e += 256) (&s_grad[0][0])[e] = 0u;
            for (int e = tid; e < kBatch * 4; e += 256) (&s_lossp[0][0])[e] = 0u;
            __syncthreads();
            const int nch = (nb + kChunk - 1) / kChunk;
            for (int c = nch - 1; c >= 0; --c) {
                const int e0 = b0 + c * kChunk;
                const uint32_t bits = (s_act[wid][e0 >> 5] >> (e0 & 31)) & ((1u << kChunk) - 1u);
                if (bits == 0u) continue;
                const int hi = 31 - __clz((int)bits);
                const uint32_t nzb = (s_nz[wid][e0 >> 5] >> (e0 & 31)) & ((1u << kChunk) - 1u);
                float Tc = ckpt[(size_t)(e0 / kChunk) * kTilePx + tid];
                float ra[kChunk], rT[kChunk], rd[kChunk];
                uint32_t rk[kChunk];
                // alphas, depths and segments stored by the forward; the
                // transmittance chain is recomputed from the checkpoint
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                    ra[j] = 0.0f;
                    rd[j] = 0.0f;
                    rk[j] = 1u;
                    if (j <= hi && ((nzb >> j) & 1u)) {
                        const size_t q = (size_t)(start + (uint32_t)(e0 + j)) * kTilePx + tid;
                        ra[j] = A.pa[q];
                        rd[j] = A.pd[q];
                        rk[j] = A.pk[q];
                        n_reval += 1ull;
                    }
                }
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                    rT[j] = Tc;
                    if (j <= hi) Tc = Tc * (1.0f - ra[j]);
                }
#pragma unroll 1
                for (int j = kChunk - 1; j >= 0; --j) {
                    if ((bits >> j) & 1u) {
                        const int jj = c * kChunk + j;
                        const bool active = valid && !(ra[j] < 0.001f);
                        n_active += (unsigned long long)__popcll(__ballot(active));
                        n_went += 1ull;
                        uint32_t g16[16], g8[8];
                        if (active) {
                            bz_pair_grads(s_f[jj], s_b[jj], pxf, pyf, rk[j], ra[j], rT[j], rd[j], is_fine,
                                          A.prm.bg_penalty, s, g16, g8);
                        } else {
#pragma unroll
                            for (int q = 0; q < 16; ++q) g16[q] = 0u;
#pragma unroll
                            for (int q = 0; q < 8; ++q) g8[q] = 0u;
                        }
                        const uint32_t t16 = wave_reduce16(g16, lane);
                        const uint32_t t8 = wave_reduce8(g8, lane);
                        if ((lane & 3) == 0 && t16 != 0u) {
                            const int idx = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                                            ((lane >> 2) & 1);
                            atomicAdd(&s_grad[jj][idx], t16);
                        }
                        if ((lane & 7) == 0 && t8 != 0u) {
                            const int idx = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
                            if (idx < 2) atomicAdd(&s_grad[jj][16 + idx], t8);
                            else if (idx < 6) atomicAdd(&s_lossp[jj][idx - 2], t8);
                        }
                    }
                }
            }
            __syncthreads();
            for (int e = tid; e < nb * kBzParams; e += 256) {
                const uint32_t val = (&s_grad[0][0])[e];
                const int j = e / kBzParams, q = e - j * kBzParams;
                if (val != 0u) atomicAdd((unsigned int*)&A.grads[kBzParams * (size_t)s_ids[j] + q], val);
            }
            for (int j = tid; j < nb; j += 256) {
                const unsigned long long l = (unsigned long long)s_lossp[j][0] +
                                             ((unsigned long long)s_lossp[j][1] << 16) +
                                             ((unsigned long long)s_lossp[j][2] << 32) +
                                             ((unsigned long long)s_lossp[j][3] << 48);
                if (l != 0ull) atomicAdd(&A.loss_acc[s_ids[j]], l);
            }
        }
    if (lane == 0) {
        s_cnt[0][wid] = n_active;
        s_cnt[1][wid] = n_went;
        s_cnt[2][wid] = n_reval;
    }
    __syncthreads();
    if (tid < 3)
        A.part[5 * (size_t)tile + 2 + tid] = (s_cnt[tid][0] + s_cnt[tid][1]) + (s_cnt[tid][2] + s_cnt[tid][3]);
}

void launch_bz_composite(hipStream_t st, const ViewParams& v, const BzCompositeArgs& a) {
    if (v.num_tiles == 0) return;
    hipLaunchKernelGGL(k_bz_forward, dim3(v.num_tiles), dim3(256), 0, st, v, a);
    if (a.backward) hipLaunchKernelGGL(k_bz_reverse, dim3(v.num_tiles), dim3(256), 0, st, v, a);
}

// ------------------------------------------------------------------- step
struct Mat16 {
    float m[16];
};

// bezier_step.wgsl:49-189 for one curve.
__global__ __launch_bounds__(256) void k_bz_step(float* __restrict__ curves, int32_t* __restrict__ grads,
                                                 float* __restrict__ m, float* __restrict__ v,
                                                 float* __restrict__ grad_accum, uint32_t n, Mat16 vp, BzParams prm,
                                                 float bc1, float bc2, float pixel_norm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float b[20];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        float4 q = reinterpret_cast<const float4*>(curves)[5 * (size_t)i + k];
        b[4 * k + 0] = q.x;
        b[4 * k + 1] = q.y;
        b[4 * k + 2] = q.z;
        b[4 * k + 3] = q.w;
    }
    int32_t* g = grads + kBzParams * (size_t)i;
    if (b[19] < kBzKill) {
#pragma unroll
        for (int lp = 0; lp < kBzParams; ++lp) g[lp] = 0;
        return;
    }
    constexpr int slot[18] = {0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 16, 17, 18, 19, 3, 7};
    const float width_hi = prm.max_width > 0.0f ? prm.max_width : 0.1f;
    float pn2 = 0.0f;
#pragma unroll
    for (int lp = 0; lp < kBzParams; ++lp) {
        const size_t pi = kBzParams * (size_t)i + lp;
        const int32_t raw = g[lp];
        g[lp] = 0;
        const float fps = (lp >= 12 && lp < 16) ? 100000.0f : 10000.0f;
        const float lr = lp < 12 ? 0.005f : (lp < 15 ? 0.01f : (lp == 15 ? 0.005f : 0.002f));
        const float mu = lp < 12 ? 0.005f : (lp < 15 ? 0.01f : (lp == 15 ? 0.005f : 0.003f));
        const float lo = lp < 12 ? -1e9f : (lp < 16 ? 0.0f : 0.001f);
        const float hi = lp < 12 ? 1e9f : (lp < 15 ? 1.0f : (lp == 15 ? 0.99f : (lp == 16 ? width_hi : 0.03f)));
        const float grad = (float)raw / fps * pixel_norm;
        const float beta1 = 0.9f, beta2 = 0.999f;
        const float mm = beta1 * m[pi] + (1.0f - beta1) * grad;
        const float vv = beta2 * v[pi] + (1.0f - beta2) * grad * grad;
        m[pi] = mm;
        v[pi] = vv;
        const float m_hat = mm / bc1, v_hat = vv / bc2;
        const float raw_update = lr * m_hat / (sqrtf(v_hat) + 1e-8f);
        const float upd = clamp_w(raw_update, -mu, mu);
        b[slot[lp]] = clamp_w(b[slot[lp]] - upd, lo, hi);
        if (lp <= 11) pn2 = pn2 + grad * grad;
    }
    grad_accum[i] = grad_accum[i] + sqrtf(pn2);
    const float at = prm.prune_alpha > 0.0f ? prm.prune_alpha : 0.001f;
    const float wt = prm.prune_width > 0.0f ? prm.prune_width : 0.001f;
    if (b[19] < at || b[3] <= wt) b[19] = 0.0f;
    if (b[19] > 0.0f && prm.no_kill < 0.5f) {
        const float mg = 1.2f;
        bool L = true, R = true, Bo = true, Ab = true, Be = true;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            F4 k = mat_vec(vp.m, b[4 * c + 0], b[4 * c + 1], b[4 * c + 2], 1.0f);
            L = L && (k.x < -mg * k.w);
            R = R && (k.x > mg * k.w);
            Bo = Bo && (k.y < -mg * k.w);
            Ab = Ab && (k.y > mg * k.w);
            Be = Be && (k.w < 0.0f);
        }
        if (L || R || Bo || Ab || Be) b[19] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
        reinterpret_cast<float4*>(curves)[5 * (size_t)i + k] =
            make_float4(b[4 * k + 0], b[4 * k + 1], b[4 * k + 2], b[4 * k + 3]);
}

void launch_bz_step(hipStream_t st, float* curves, int32_t* grads, float* m, float* v, float* grad_accum, uint32_t n,
                    const float vp[16], const BzParams& prm, float t, float pixel_count) {
    if (n == 0) return;
    Mat16 M;
    for (int k = 0; k < 16; ++k) M.m[k] = vp[k];
    const float bc1 = 1.0f - pow_det(0.9f, t), bc2 = 1.0f - pow_det(0.999f, t);
    const float pixel_norm = 1.0f / fmax_w(pixel_count, 1.0f);
    hipLaunchKernelGGL(k_bz_step, dim3((n + 255) / 256), dim3(256), 0, st, curves, grads, m, v, grad_accum, n, M, prm,
                       bc1, bc2, pixel_norm);
}

// -------------------------------------------------------------------- ADC
constexpr int kAdcThreads = 1024;
constexpr float kBzAdcPeriod = 50.0f, kBzTauPos = 0.0002f, kBzTauLoss = 0.001f, kBzSplitLen = 0.25f;

__device__ __forceinline__ float bz_hash(float x) { return fract_det(sin_det(x) * 43758.5453f); }

// Exclusive block-wide prefix of a 0/1 flag over the 1024 threads; *total gets the count.
__device__ __forceinline__ uint32_t bz_block_scan(bool f, uint32_t* s_w, uint32_t* total) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const unsigned long long bal = __ballot(f);
    const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) s_w[wid] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < kAdcThreads / 64; ++w) {
        const uint32_t c = s_w[w];
        if (w < wid) off += c;
        tot += c;
    }
    *total = tot;
    return off + below;
}

// BW ADC:134-181: split (de Casteljau at 1/2) or jittered clone of curve i into
// slot j, Adam state of both cleared.  One thread.
__device__ void bz_split_or_clone(float* curves, float* m, float* v, uint32_t i, uint32_t j, float adam_t) {
    float* b = curves + 20 * (size_t)i;
    float o[20], nb[20];
#pragma unroll
    for (int k = 0; k < 20; ++k) o[k] = b[k];
#pragma unroll
    for (int k = 0; k < 20; ++k) nb[k] = o[k];
    float mid[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) mid[c] = (((o[c] + 3.0f * o[4 + c]) + 3.0f * o[8 + c]) + o[12 + c]) * 0.125f;
    const float d0x = mid[0] - o[0], d0y = mid[1] - o[1], d0z = mid[2] - o[2];
    const float d1x = o[12] - mid[0], d1y = o[13] - mid[1], d1z = o[14] - mid[2];
    const float len_approx = sqrtf((d0x * d0x + d0y * d0y) + d0z * d0z) + sqrtf((d1x * d1x + d1y * d1y) + d1z * d1z);
    if (len_approx > kBzSplitLen) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float q0 = (o[c] + o[4 + c]) * 0.5f, q1 = (o[4 + c] + o[8 + c]) * 0.5f,
                        q2 = (o[8 + c] + o[12 + c]) * 0.5f;
            const float r0 = (q0 + q1) * 0.5f, r1 = (q1 + q2) * 0.5f;
            const float s = (r0 + r1) * 0.5f;
            b[4 + c] = q0;
            b[8 + c] = r0;
            b[12 + c] = s;
            nb[0 + c] = s;
            nb[4 + c] = r1;
            nb[8 + c] = q2;
            nb[12 + c] = o[12 + c];
        }
        b[11] = 0.0f;
        b[15] = 0.0f;
    } else {
        const float seed = (float)i * 3.14159f + adam_t;
        const float jx = (bz_hash(seed * 12.9898f) - 0.5f) * 0.001f;
        const float jy = (bz_hash(seed * 78.233f) - 0.5f) * 0.001f;
        const float jz = (bz_hash(seed * 43.123f) - 0.5f) * 0.001f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            nb[4 * k + 0] = o[4 * k + 0] + jx;
            nb[4 * k + 1] = o[4 * k + 1] + jy;
            nb[4 * k + 2] = o[4 * k + 2] + jz;
        }
    }
    nb[11] = 0.0f;
    nb[15] = 0.0f;
    float* dst = curves + 20 * (size_t)j;
#pragma unroll
    for (int k = 0; k < 20; ++k) dst[k] = nb[k];
    for (int p = 0; p < kBzParams; ++p) {
        m[kBzParams * (size_t)i + p] = 0.0f;
        v[kBzParams * (size_t)i + p] = 0.0f;
        m[kBzParams * (size_t)j + p] = 0.0f;
        v[kBzParams * (size_t)j + p] = 0.0f;
    }
}

__global__ __launch_bounds__(kAdcThreads) void k_bz_adc(float* curves, float* m, float* v, float* grad_accum,
                                                        unsigned long long* loss_acc, int32_t* pixel_loss, uint32_t n,
                                                        uint32_t W, uint32_t H, Mat16 vp, Mat16 vpi, BzParams prm,
                                                        float adam_t, BzAdcBuffers B) {
    __shared__ uint32_t s_w[kAdcThreads / 64];
    __shared__ uint32_t s_dc;
    __shared__ int32_t s_bv[kAdcThreads / 64];
    __shared__ uint32_t s_bk[kAdcThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    uint32_t tot;
    // pass 0: dead slots, ascending (ADC:78-83)
    uint32_t dc = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += kAdcThreads) {
        const uint32_t i = i0 + tid;
        const bool f = i < n && curves[20 * (size_t)i + 19] < kBzKill;
        const uint32_t pos = bz_block_scan(f, s_w, &tot);
        if (f) B.dead_list[dc + pos] = i;
        dc += tot;
    }
    const uint32_t d0 = dc;
    // pass 1: kills of stuck / high-loss and off-screen curves, appended ascending (ADC:93-119)
    for (uint32_t i0 = 0; i0 < n; i0 += kAdcThreads) {
        const uint32_t i = i0 + tid;
        bool kill = false;
        if (i < n) {
            const float* b = curves + 20 * (size_t)i;
            if (!(b[19] < kBzKill)) {
                const float gn = grad_accum[i] / kBzAdcPeriod;
                const float ln = (float)((double)(int64_t)loss_acc[i] * 0x1p-32) / kBzAdcPeriod;
                if (prm.no_kill < 0.5f && gn <= kBzTauPos && ln > kBzTauLoss) {
                    kill = true;
                } else {
                    bool all_off = true;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        F4 k = mat_vec(vp.m, b[4 * c + 0], b[4 * c + 1], b[4 * c + 2], 1.0f);
                        bool off;
                        if (k.w < 0.0f) {
                            off = true;
                        } else {
                            const float w = fmax_w(k.w, 1e-5f);
                            off = fabsf(k.x / w) > 1.2f || fabsf(k.y / w) > 1.2f;
                        }
                        all_off = all_off && off;
                    }
                    kill = all_off;
                }
            }
        }
        const uint32_t pos = bz_block_scan(kill, s_w, &tot);
        if (kill) {
            B.dead_list[dc + pos] = i;
            curves[20 * (size_t)i + 19] = 0.0f;
        }
        dc += tot;
    }
    const uint32_t d1 = dc;
    // candidates of both split/clone passes: grad_norm above TAU_POS, ascending
    uint32_t nc = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += kAdcThreads) {
        const uint32_t i = i0 + tid;
        const bool f = i < n && !(grad_accum[i] / kBzAdcPeriod <= kBzTauPos);
        const uint32_t pos = bz_block_scan(f, s_w, &tot);
        if (f) B.cand[nc + pos] = i;
        nc += tot;
    }
    __syncthreads();
    int64_t st_split1 = 0, st_split3 = 0, st_spawn = 0;
    // pass 1.5 (ADC:122-182): in index order, each live candidate takes the
    // dead slot at the top of the stack
    if (tid == 0) {
        for (uint32_t c = 0; c < nc; ++c) {
            const uint32_t i = B.cand[c];
            if (curves[20 * (size_t)i + 19] < kBzKill) continue;
            if (dc == 0) break;
            dc -= 1u;
            bz_split_or_clone(curves, m, v, i, B.dead_list[dc], adam_t);
            st_split1 += 1;
        }
        s_dc = dc;
    }
    __syncthreads();
    dc = s_dc;
    // pass 2 (ADC:187-243): spawns at the highest-loss pixels
    const uint32_t npx = W * H;
    const uint32_t max_spawns = min(dc, 512u);
    const uint32_t stride = max(1u, npx / 4096u);
    for (uint32_t sp = 0; sp < max_spawns; ++sp) {
        dc = s_dc;
        if (dc == 0) break;
        const uint32_t st0 = (sp * 997u) % stride;
        const uint32_t K = st0 < npx ? (npx - st0 + stride - 1u) / stride : 0u;
        int32_t bv = 0;
        uint32_t bk = 0xFFFFFFFFu;
        for (uint32_t k = tid; k < K; k += kAdcThreads) {
            const int32_t val = pixel_loss[st0 + k * stride];
            if (val > bv) {
                bv = val;
                bk = k;
            }
        }
        // first index of the maximum (the scan's strict `>`)
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const int32_t ov = __shfl_xor(bv, o, 64);
            const uint32_t ok = __shfl_xor(bk, o, 64);
            if (ov > bv || (ov == bv && ok < bk)) {
                bv = ov;
                bk = ok;
            }
        }
        if (lane == 0) {
            s_bv[wid] = bv;
            s_bk[wid] = bk;
        }
        __syncthreads();
        if (tid == 0) {
            int32_t best_v = 0;
            uint32_t best_k = 0xFFFFFFFFu;
            for (int w = 0; w < kAdcThreads / 64; ++w) {
                if (s_bv[w] > best_v || (s_bv[w] == best_v && s_bk[w] < best_k)) {
                    best_v = s_bv[w];
                    best_k = s_bk[w];
                }
            }
            uint32_t spawn_px;
            if (best_v <= 0) {
                const float seed_fb = (float)sp * 1234.5678f + adam_t * 0.1f;
                const uint32_t rx = (uint32_t)(bz_hash(seed_fb * 12.9898f) * (float)W);
                const uint32_t ry = (uint32_t)(bz_hash(seed_fb * 78.233f) * (float)H);
                spawn_px = ry * W + rx;
            } else {
                spawn_px = st0 + best_k * stride;
                pixel_loss[spawn_px] = 0;
            }
            dc -= 1u;
            const uint32_t slot = B.dead_list[dc];
            // pixel_to_world (ADC:52-65) with this view's size, spawn depth 0.5
            const uint32_t pxx = spawn_px % W, pyy = spawn_px / W;
            const float ux = ((float)pxx + 0.5f) / (float)W, uy = ((float)pyy + 0.5f) / (float)H;
            const float nx = ux * 2.0f - 1.0f, ny = 1.0f - uy * 2.0f;
            const float w = 0.1f / fmax_w(1.0f - 0.5f, 1e-5f);
            F4 wc = mat_vec(vpi.m, nx * w, ny * w, w, w);
            const float cxw = wc.x / wc.w, cyw = wc.y / wc.w, czw = wc.z / wc.w;
            const float seed = (float)spawn_px * 1.61803f + (float)sp * 2.71828f;
            const float angle = bz_hash(seed * 127.1f) * 6.28318f;
            const float tgx = cos_det(angle) * 0.025f, tgz = sin_det(angle) * 0.025f;
            float* nb = curves + 20 * (size_t)slot;
            const float ctr[3] = {cxw, cyw, czw}, tg[3] = {tgx, 0.0f, tgz};
            for (int c = 0; c < 3; ++c) {
                nb[0 + c] = ctr[c] - tg[c];
                nb[4 + c] = ctr[c] - tg[c] * 0.33f;
                nb[8 + c] = ctr[c] + tg[c] * 0.33f;
                nb[12 + c] = ctr[c] + tg[c];
            }
            nb[3] = 0.015f;
            nb[7] = 0.005f;
            nb[11] = 0.0f;
            nb[15] = 0.0f;
            nb[16] = nb[17] = nb[18] = nb[19] = 0.5f;
            for (int p = 0; p < kBzParams; ++p) {
                m[kBzParams * (size_t)slot + p] = 0.0f;
                v[kBzParams * (size_t)slot + p] = 0.0f;
            }
            st_spawn += 1;
            s_dc = dc;
        }
        __syncthreads();
    }
    // pass 3 (ADC:246-305)
    if (tid == 0) {
        dc = s_dc;
        for (uint32_t c = 0; c < nc; ++c) {
            if (dc == 0) break;
            const uint32_t i = B.cand[c];
            if (curves[20 * (size_t)i + 19] < kBzKill) continue;
            dc -= 1u;
            bz_split_or_clone(curves, m, v, i, B.dead_list[dc], adam_t);
            st_split3 += 1;
        }
        s_dc = dc;
        B.stats[0] = (int64_t)d0;
        B.stats[1] = (int64_t)(d1 - d0);
        B.stats[2] = st_split1;
        B.stats[3] = st_spawn;
        B.stats[4] = st_split3;
        B.stats[5] = (int64_t)dc;
    }
    __syncthreads();
    // resets (ADC:307-314)
    for (uint32_t k = tid; k < npx; k += kAdcThreads) pixel_loss[k] = 0;
    for (uint32_t i = tid; i < n; i += kAdcThreads) {
        grad_accum[i] = 0.0f;
        loss_acc[i] = 0ull;
    }
}

void launch_bz_adc(hipStream_t st, float* curves, float* m, float* v, float* grad_accum, unsigned long long* loss_acc,
                   int32_t* pixel_loss, uint32_t n, uint32_t W, uint32_t H, const float vp[16], const float vp_inv[16],
                   const BzParams& prm, float adam_t, const BzAdcBuffers& b) {
    Mat16 A, Bi;
    for (int k = 0; k < 16; ++k) {
        A.m[k] = vp[k];
        Bi.m[k] = vp_inv[k];
    }
    hipLaunchKernelGGL(k_bz_adc, dim3(1), dim3(kAdcThreads), 0, st, curves, m, v, grad_accum, loss_acc, pixel_loss, n,
                       W, H, A, Bi, prm, adam_t, b);
}

// ------------------------------------------------------------ display raster
// bezier_forward.wgsl per curve (:47-112): projection, the near cull (any control
// point at w <= 0, :74-81), the quad (hull of the control points -+ width +
// softness, clamped to [-4, 4], :88-96) and its tiles, and the polyline points
// the fragment shader recomputes per pixel (:133-136).  A curve with opacity
// below 0.0009 is dropped: its alpha, a_geom (<= 1) * brush (<= 1 up to a few
// ulps) * opacity * pressure (<= 1), stays below 0.001 on every pixel, where
// the fragment is discarded.
__global__ __launch_bounds__(256) void k_bz_draw_preprocess(const float4* __restrict__ curves, uint32_t n,
                                                            ViewParams v, BzDraw* __restrict__ rec,
                                                            uint32_t* __restrict__ counts,
                                                            uint4* __restrict__ rects) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = curves[5 * (size_t)i + k];
    const float4 col = curves[5 * (size_t)i + 4];
    counts[i] = 0u;
    if (col.w < 0.0009f) return;
    const float aspect = v.aspect;
    F4 k4[4];
    V2 p[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        k4[k] = mat_vec(v.vp, c[k].x, c[k].y, c[k].z, 1.0f);
        p[k].x = k4[k].x / k4[k].w * aspect;  // project_to_screen (:36-40)
        p[k].y = k4[k].y / k4[k].w;
    }
    if (k4[0].w <= 0.0f || k4[1].w <= 0.0f || k4[2].w <= 0.0f || k4[3].w <= 0.0f) return;
    BzDraw f;
    f.width = fmax_w(c[0].w, 0.0001f);
    f.soft = fmax_w(c[1].w, 0.0001f);
    const float pad = f.width + f.soft;
    f.qx0 = fmax_w(fmin_w(fmin_w(p[0].x, p[1].x), fmin_w(p[2].x, p[3].x)) - pad, -4.0f);
    f.qy0 = fmax_w(fmin_w(fmin_w(p[0].y, p[1].y), fmin_w(p[2].y, p[3].y)) - pad, -4.0f);
    f.qx1 = fmin_w(fmax_w(fmax_w(p[0].x, p[1].x), fmax_w(p[2].x, p[3].x)) + pad, 4.0f);
    f.qy1 = fmin_w(fmax_w(fmax_w(p[0].y, p[1].y), fmax_w(p[2].y, p[3].y)) + pad, 4.0f);
    // every tile holding a pixel centre inside the quad
    const uint4 tr = tile_rect(f.qx0, f.qx1, f.qy0, f.qy1, v.W, v.H, (uint32_t)kTile, v.tiles_x, v.tiles_y, aspect);
    if (tr.x >= tr.y || tr.z >= tr.w) return;
    counts[i] = (tr.y - tr.x) * (tr.w - tr.z);
    rects[i] = tr;
    f.x[0] = p[0].x;
    f.y[0] = p[0].y;
#pragma unroll
    for (int k = 1; k <= kBzSeg; ++k) {
        V2 q = bz_at(p[0], p[1], p[2], p[3], (float)k / (float)kBzSeg);
        f.x[k] = q.x;
        f.y[k] = q.y;
    }
    float bx0 = f.x[0], bx1 = f.x[0], by0 = f.y[0], by1 = f.y[0];
#pragma unroll
    for (int k = 1; k <= kBzSeg; ++k) {
        bx0 = fmin_w(bx0, f.x[k]);
        bx1 = fmax_w(bx1, f.x[k]);
        by0 = fmin_w(by0, f.y[k]);
        by1 = fmax_w(by1, f.y[k]);
    }
    f.op = col.w;
    f.r = col.x;
    f.g = col.y;
    f.b = col.z;
    f.cx0 = bx0;
    f.cy0 = by0;
    f.cx1 = bx1;
    f.cy1 = by1;
    // as in training: beyond rcull from the polyline box every computed distance
    // exceeds width + softness, where alpha is 0
    const float M = fmax_w(fmax_w(fabsf(bx0), fabsf(bx1)), fmax_w(fabsf(by0), fabsf(by1)));
    f.rcull = (f.width + f.soft) * 1.0001f + 1e-5f * (4.0f + M);
    f.pad[0] = f.pad[1] = f.pad[2] = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        float gx0 = f.x[4 * g], gx1 = f.x[4 * g], gy0 = f.y[4 * g], gy1 = f.y[4 * g];
#pragma unroll
        for (int j = 1; j <= 4; ++j) {
            gx0 = fmin_w(gx0, f.x[4 * g + j]);
            gx1 = fmax_w(gx1, f.x[4 * g + j]);
            gy0 = fmin_w(gy0, f.y[4 * g + j]);
            gy1 = fmax_w(gy1, f.y[4 * g + j]);
        }
        f.gb[4 * g + 0] = gx0;
        f.gb[4 * g + 1] = gy0;
        f.gb[4 * g + 2] = gx1;
        f.gb[4 * g + 3] = gy1;
    }
    rec[i] = f;
}

// The brush sampler (linear filtering, repeat in u, clamp-to-edge in v, one mip
// level; GpuBezierForwardPipelineManager.ts:34-41) as FP32 bilinear filtering
// of the texel values at x = u*bw - 0.5, y = v*bh - 0.5.
__device__ __forceinline__ float bz_brush(const float* __restrict__ tex, int bw, int bh, float u, float v) {
    const float x = u * (float)bw - 0.5f, y = v * (float)bh - 0.5f;
    const float x0f = floorf(x), y0f = floorf(y);
    const float fx = x - x0f, fy = y - y0f;
    const int ix = (int)x0f, iy = (int)y0f;
    const int x0 = ((ix % bw) + bw) % bw, x1 = (((ix + 1) % bw) + bw) % bw;
    const int y0 = min(max(iy, 0), bh - 1), y1 = min(max(iy + 1, 0), bh - 1);
    const float t00 = tex[(size_t)y0 * bw + x0], t10 = tex[(size_t)y0 * bw + x1];
    const float t01 = tex[(size_t)y1 * bw + x0], t11 = tex[(size_t)y1 * bw + x1];
    const float top = t00 * (1.0f - fx) + t10 * fx;
    const float bot = t01 * (1.0f - fx) + t11 * fx;
    return top * (1.0f - fy) + bot * fy;
}

// fs_main (:115-182) for one pixel: alpha, 0 where the pixel is outside the quad.
// The closest segment is found as in training (bz_eval: skipped segment groups,
// square roots only for a possible new minimum, divisions only where the clamp
// does not decide u), then its signed cross distance is recomputed from the
// same operands.  A minimum >= fl(width + softness) gives a_geom = 0 and alpha 0.
__device__ __forceinline__ float bz_draw_alpha(const BzDraw& f, float px, float py, const BzBlock& bk, float thr2,
                                               const float* __restrict__ brush, int bw, int bh) {
    const bool cov = (px >= f.qx0) & (px <= f.qx1) & (py >= f.qy0) & (py <= f.qy1);
    float min_d = 1e9f, min_d2 = __builtin_inff(), min_u = 0.0f;
    uint32_t min_k = 1u;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float gx = __builtin_fmaxf(__builtin_fmaxf(f.gb[4 * g + 0] - bk.xb, bk.xa - f.gb[4 * g + 2]), 0.0f);
        const float gy = __builtin_fmaxf(__builtin_fmaxf(f.gb[4 * g + 1] - bk.yb, bk.ya - f.gb[4 * g + 3]), 0.0f);
        if (__builtin_amdgcn_readfirstlane((int)(gx * gx + gy * gy > thr2))) continue;
        float prevx = f.x[4 * g], prevy = f.y[4 * g];
#pragma unroll
        for (int k = 4 * g + 1; k <= 4 * g + 4; ++k) {
            const float cx = f.x[k], cy = f.y[k];
            const float sx = cx - prevx, sy = cy - prevy;
            const float len2 = fmax_w(sx * sx + sy * sy, 1e-8f);
            const float qx = px - prevx, qy = py - prevy;
            const float num = qx * sx + qy * sy;
            float u;
            if (num <= 0.0f) u = 0.0f;
            else if (num >= len2 && len2 < __builtin_inff()) u = 1.0f;
            else u = clamp_w(num / len2, 0.0f, 1.0f);
            const float ox = prevx + sx * u, oy = prevy + sy * u;
            const float dx = px - ox, dy = py - oy;
            const float d2 = dx * dx + dy * dy;
            if (d2 < min_d2) {
                const float d = sqrtf(d2);
                if (d < min_d) {
                    min_d = d;
                    min_d2 = d2;
                    min_k = (uint32_t)k;
                    min_u = u;
                }
            }
            prevx = cx;
            prevy = cy;
        }
    }
    if (!cov || min_d >= f.width + f.soft) return 0.0f;
    const float prevx = f.x[min_k - 1u], prevy = f.y[min_k - 1u];
    const float sx = f.x[min_k] - prevx, sy = f.y[min_k] - prevy;
    const float len2 = fmax_w(sx * sx + sy * sy, 1e-8f);
    const float dx = px - (prevx + sx * min_u), dy = py - (prevy + sy * min_u);
    const float seg_len = sqrtf(len2);
    const float dirx = sx / seg_len, diry = sy / seg_len;
    const float cross = dx * (-diry) + dy * dirx;
    const float t = ((float)(min_k - 1u) + min_u) / (float)kBzSeg;
    const float dt = t - 0.5f;
    const float pressure = 1.0f - 4.0f * dt * dt;
    const float lw = f.width * pressure, ls = f.soft * pressure, lo = f.op * pressure;
    const float a_geom = 1.0f - bz_smoothstep(lw - ls, lw + ls, min_d);
    const float bv = clamp_w(cross / fmax_w(lw + ls, 1e-6f) * 0.5f + 0.5f, 0.0f, 1.0f);
    const float br = bz_brush(brush, bw, bh, t, bv);
    return clamp_w(a_geom * br * lo, 0.0f, 0.999f);
}

__device__ __forceinline__ float bz_q8(float x) { return rintf(clamp_w(x, 0.0f, 1.0f) * 255.0f) / 255.0f; }

constexpr int kDrawBatch = 32;

// The instanced draw (:47-182) with its blend state (premultiplied output,
// one / one-minus-src-alpha, GpuBezierForwardPipelineManager.ts:79-92) as a
// tiled pass: one 256-thread workgroup per 16x16 tile, each wave an 8x8 block,
// the tile's curves walked from the highest id down (instance ii draws curve
// N-1-ii), records staged in LDS 32 at a time.  An entry whose polyline box is
// farther than the cull radius from the wave's block, or whose quad misses the
// block, draws nothing there and is skipped by the whole wave.
__global__ __launch_bounds__(256) void k_bz_draw(ViewParams v, BzDrawArgs A) {
    __shared__ BzDraw s_f[kDrawBatch];
    const uint32_t tile = xcd_remap(blockIdx.x, v.num_tiles);
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t start = A.ranges[2 * tile], end = A.ranges[2 * tile + 1];
    const int count = (int)(end - start);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t bx = tx * kTile + (wid & 1) * 8, by = ty * kTile + (wid >> 1) * 8;
    const uint32_t px = bx + (lane & 7), py = by + (lane >> 3);
    const bool valid = px < v.W && py < v.H;
    const float pxf = px_to_x(px, v.W, v.aspect);
    const float pyf = py_to_y(py, v.H);
    const float xa = px_to_x(bx, v.W, v.aspect), xb = px_to_x(bx + 7, v.W, v.aspect);
    const float ya = py_to_y(by + 7, v.H), yb = py_to_y(by, v.H);
    const float blk_margin = 1e-5f * (((fabsf(xa) + fabsf(xb)) + fabsf(ya)) + fabsf(yb));
    const BzBlock blk = {xa, xb, ya, yb};
    const size_t pix = valid ? (size_t)py * v.W + px : 0;
    float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!A.clear && valid) c = A.rgba[pix];
    constexpr int FQ = (int)(sizeof(BzDraw) / 16);
    for (int b0 = count > 0 ? ((count - 1) / kDrawBatch) * kDrawBatch : -1; b0 >= 0; b0 -= kDrawBatch) {
        const int nb = min(kDrawBatch, count - b0);
        __syncthreads();
        for (int e = tid; e < nb * FQ; e += 256) {
            const int j = e / FQ, part = e - j * FQ;
            const uint32_t id = (uint32_t)A.keys[start + (uint32_t)(b0 + j)];
            reinterpret_cast<float4*>(&s_f[j])[part] = reinterpret_cast<const float4*>(&A.rec[id])[part];
        }
        __syncthreads();
        for (int j = nb - 1; j >= 0; --j) {
            const BzDraw& f = s_f[j];
            const float thr = f.rcull + blk_margin;
            const float thr2 = thr * thr;
            const float gx = __builtin_fmaxf(__builtin_fmaxf(f.cx0 - xb, xa - f.cx1), 0.0f);
            const float gy = __builtin_fmaxf(__builtin_fmaxf(f.cy0 - yb, ya - f.cy1), 0.0f);
            const bool miss = (gx * gx + gy * gy > thr2) | (f.qx0 > xb) | (f.qx1 < xa) | (f.qy0 > yb) | (f.qy1 < ya);
            if (__builtin_amdgcn_readfirstlane((int)miss)) continue;
            const float a = bz_draw_alpha(f, pxf, pyf, blk, thr2, A.brush, A.bw, A.bh);
            if (valid && !(a < 0.001f)) {
                const float one_m = 1.0f - a;
                c.x = f.r * a + c.x * one_m;
                c.y = f.g * a + c.y * one_m;
                c.z = f.b * a + c.z * one_m;
                c.w = a + c.w * one_m;
                if (A.quantize) c = make_float4(bz_q8(c.x), bz_q8(c.y), bz_q8(c.z), bz_q8(c.w));
            }
        }
    }
    if (valid) A.rgba[pix] = c;
}

void launch_bz_draw_preprocess(hipStream_t st, const float* curves, uint32_t n, const ViewParams& v, BzDraw* rec,
                               uint32_t* counts, uint4* rects) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_bz_draw_preprocess, dim3((n + 255) / 256), dim3(256), 0, st, (const float4*)curves, n, v,
                       rec, counts, rects);
}

void launch_bz_draw(hipStream_t st, const ViewParams& v, const BzDrawArgs& a) {
    if (v.num_tiles == 0) return;
    hipLaunchKernelGGL(k_bz_draw, dim3(v.num_tiles), dim3(256), 0, st, v, a);
}

}  // namespace nps
