/*
 * oracle_bezier.c -- CPU restatement of the reference's Bezier stroke optimizer
 * (src/lib/gpu/bezier_backward.wgsl, bezier_step.wgsl, bezier_adc.wgsl, driven by
 * GpuBezierOptimizerManager.ts).  TEST INFRASTRUCTURE ONLY.
 *
 * FP32, expressions in the WGSL's order, no contraction, transcendentals from
 * ref_math.h (the same definitions the splat oracle uses).  Defined where the
 * reference is not:
 *   - adc.loss_accum is updated non-atomically by every pixel (a race,
 *     bezier_backward.wgsl:564-566): here each (pixel, curve) contribution is
 *     converted to 2^-32 fixed point (truncation) and summed in int64, so the
 *     sum is order-independent; ADC reads it as (float)(sum * 2^-32);
 *   - adam.t is read by every workgroup while one thread increments it
 *     (bezier_step.wgsl:53-58): t is the host step counter;
 *   - pixel_loss is sized to the view (the reference fixes 512^2,
 *     constants.ts:10) and the ADC reads it with the view's dimensions (the
 *     reference compiles the ADC for 256 x 256, GpuBezierOptimizerManager.ts:146);
 *   - WGSL cos() in the spawn code = sin(x + pi/2) with the deterministic sin.
 * Curve layout (20 floats): p0.xyz width | p1.xyz softness | p2.xyz pad | p3.xyz pad | rgba.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ref_math.h"

#define BZ_NSEG 16u
#define BZ_NEAR 0.1f
#define BZ_CAP 1024
#define BZ_KILL 0.005f
#define BZ_PRUNE_ALPHA 0.001f
#define BZ_PRUNE_WIDTH 0.001f
#define BZ_OFFSCREEN 1.2f
#define BZ_FP_POS 10000.0f
#define BZ_FP_COL 100000.0f
#define BZ_ADC_PERIOD 50.0f
#define BZ_TAU_POS 0.0002f
#define BZ_TAU_LOSS 0.001f
#define BZ_SPLIT_LEN 0.25f
#define BZ_MAX_SPAWNS 512u
#define BZ_SPAWN_TLEN 0.025f
#define BZ_SPAWN_WIDTH 0.015f
#define BZ_SPAWN_SOFT 0.005f
#define BZ_LOSS_FP 4294967296.0f /* 2^32 */

typedef struct {
    float vp[16];
    float mode, max_width, prune_alpha, prune_width, bg_penalty;
    float vp_inv[16];
} bz_uniforms;

typedef struct { float x, y; } v2;
typedef struct { float x, y, z; } v3;

static inline v2 v2s(v2 a, float s) { v2 r = { a.x * s, a.y * s }; return r; }
static inline v2 v2a(v2 a, v2 b) { v2 r = { a.x + b.x, a.y + b.y }; return r; }
static inline v2 v2m(v2 a, v2 b) { v2 r = { a.x - b.x, a.y - b.y }; return r; }
static inline float v2d(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }

/* bezier_backward.wgsl:60-66 */
static v2 bezier_at(v2 p0, v2 p1, v2 p2, v2 p3, float t) {
    float omt = 1.0f - t;
    v2 a = v2s(p0, omt * omt * omt);
    v2 b = v2s(p1, 3.0f * omt * omt * t);
    v2 c = v2s(p2, 3.0f * omt * t * t);
    v2 d = v2s(p3, t * t * t);
    return v2a(v2a(v2a(a, b), c), d);
}
/* :68-71 */
static void bernstein(float t, float B[4]) {
    float omt = 1.0f - t;
    B[0] = omt * omt * omt;
    B[1] = 3.0f * omt * omt * t;
    B[2] = 3.0f * omt * t * t;
    B[3] = t * t * t;
}
/* :73-76 -> (x, y, w) */
static v3 project_center(const float* vp, const float* p, float aspect) {
    float cx = ((vp[0] * p[0] + vp[4] * p[1]) + vp[8] * p[2]) + vp[12];
    float cy = ((vp[1] * p[0] + vp[5] * p[1]) + vp[9] * p[2]) + vp[13];
    float cw = ((vp[3] * p[0] + vp[7] * p[1]) + vp[11] * p[2]) + vp[15];
    v3 r = { cx / cw * aspect, cy / cw, cw };
    return r;
}
/* :78-92 */
static v3 backproject(const float* vp, const float* p, float aspect, v2 g) {
    float cx = ((vp[0] * p[0] + vp[4] * p[1]) + vp[8] * p[2]) + vp[12];
    float cy = ((vp[1] * p[0] + vp[5] * p[1]) + vp[9] * p[2]) + vp[13];
    float w = ((vp[3] * p[0] + vp[7] * p[1]) + vp[11] * p[2]) + vp[15];
    float w2 = w * w;
    float o[3];
    for (int ax = 0; ax < 3; ++ax) {
        float v0 = vp[4 * ax + 0], v1 = vp[4 * ax + 1], v3_ = vp[4 * ax + 3];
        float dsx = aspect * (v0 * w - cx * v3_) / w2;
        float dsy = (v1 * w - cy * v3_) / w2;
        o[ax] = g.x * dsx + g.y * dsy;
    }
    v3 r = { o[0], o[1], o[2] };
    return r;
}
static inline float clampf_(float x, float lo, float hi) {
    float a = x > lo ? x : lo; /* max then min, NaN-propagating like the splat oracle's WGSL clamp */
    if (x != x) a = lo;
    return a < hi ? a : hi;
}
static inline float maxf_(float a, float b) { return (a != a) ? b : (b != b) ? a : (a > b ? a : b); }
static inline float minf_(float a, float b) { return (a != a) ? b : (b != b) ? a : (a < b ? a : b); }
/* WGSL smoothstep(e0, e1, x) */
static inline float smoothstep_(float e0, float e1, float x) {
    float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static inline v2 pixel_to_p(uint32_t px, uint32_t py, uint32_t W, uint32_t H, float aspect) {
    float ux = ((float)px + 0.5f) / (float)W, uy = ((float)py + 0.5f) / (float)H;
    v2 p = { ux * 2.0f - 1.0f, uy * 2.0f - 1.0f };
    p.y = -p.y;
    p.x = p.x * aspect;
    return p;
}
static inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
/* atomic wrap-around add of a fixed-point gradient (the reference's atomicAdd) */
static inline void gadd(int32_t* p, int32_t x) {
    uint32_t* q = (uint32_t*)p;
    uint32_t d = (uint32_t)x;
#pragma omp atomic
    *q += d;
}
/* float -> int64 toward zero, saturating, NaN -> 0 (the 2^-32 loss accumulator) */
static inline int64_t f2l(float x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0f) return INT64_MAX;
    if (x <= -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)x;
}

/* Tile lists (bezier_backward.wgsl:112-177): first BZ_CAP ascending ids of the
 * curves whose hull box (+ width + softness) overlaps the tile rectangle. */
int ora_bz_bin(const float* bz, int n, const bz_uniforms* u, int W, int H, const int32_t* tiles, int ntiles,
               uint32_t* lists, int32_t* counts, int32_t* full) {
    const float aspect = (float)W / (float)H;
    const uint32_t ntx = ((uint32_t)W + 15u) / 16u;
    /* per-curve hull box (the same values for every tile, computed once) */
    float* box = (float*)malloc(sizeof(float) * 4 * (size_t)(n > 0 ? n : 1));
    uint8_t* live = (uint8_t*)malloc((size_t)(n > 0 ? n : 1));
    if (!box || !live) { free(box); free(live); return -1; }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const float* b = bz + 20 * (size_t)i;
        live[i] = 0;
        if (b[19] < BZ_KILL) continue;
        float width = maxf_(b[3], 0.001f), soft = maxf_(b[7], 0.001f);
        v3 q0 = project_center(u->vp, b + 0, aspect), q1 = project_center(u->vp, b + 4, aspect);
        v3 q2 = project_center(u->vp, b + 8, aspect), q3 = project_center(u->vp, b + 12, aspect);
        if (q0.z <= 0.0f || q1.z <= 0.0f || q2.z <= 0.0f || q3.z <= 0.0f) continue;
        float oc = width + soft;
        box[4 * (size_t)i + 0] = minf_(minf_(q0.x, q1.x), minf_(q2.x, q3.x)) - oc;
        box[4 * (size_t)i + 1] = minf_(minf_(q0.y, q1.y), minf_(q2.y, q3.y)) - oc;
        box[4 * (size_t)i + 2] = maxf_(maxf_(q0.x, q1.x), maxf_(q2.x, q3.x)) + oc;
        box[4 * (size_t)i + 3] = maxf_(maxf_(q0.y, q1.y), maxf_(q2.y, q3.y)) + oc;
        live[i] = 1;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < ntiles; ++k) {
        uint32_t t = (uint32_t)tiles[k], tx = t % ntx, ty = t / ntx;
        uint32_t x0 = tx * 16u, y0 = ty * 16u;
        uint32_t x1 = x0 + 16u < (uint32_t)W ? x0 + 16u : (uint32_t)W, y1 = y0 + 16u < (uint32_t)H ? y0 + 16u : (uint32_t)H;
        v2 p00 = pixel_to_p(x0, y0, W, H, aspect), p11 = pixel_to_p(x1, y1, W, H, aspect);
        float tminx = p00.x, tminy = p11.y, tmaxx = p11.x, tmaxy = p00.y;
        int c = 0, f = 0;
        for (int i = 0; i < n; ++i) {
            if (!live[i]) continue;
            const float* bb = box + 4 * (size_t)i;
            float mnx = bb[0], mny = bb[1], mxx = bb[2], mxy = bb[3];
            if (mnx > tmaxx || mxx < tminx || mny > tmaxy || mxy < tminy) continue;
            if (c < BZ_CAP) lists[(size_t)k * BZ_CAP + c++] = (uint32_t)i;
            f++;
        }
        counts[k] = c;
        if (full) full[k] = f;
    }
    free(box);
    free(live);
    return 0;
}

/* Per-(pixel, curve) evaluation of the forward (:196-246). */
typedef struct {
    float a, d_val;
    uint32_t k;
} bz_eval;

static bz_eval bz_pair(const float* b, const float* vp, float aspect, v2 p) {
    bz_eval e;
    float width = maxf_(b[3], 0.001f), soft = maxf_(b[7], 0.001f), opacity = b[19];
    v3 q0 = project_center(vp, b + 0, aspect), q1 = project_center(vp, b + 4, aspect);
    v3 q2 = project_center(vp, b + 8, aspect), q3 = project_center(vp, b + 12, aspect);
    v2 p0 = { q0.x, q0.y }, p1 = { q1.x, q1.y }, p2 = { q2.x, q2.y }, p3 = { q3.x, q3.y };
    float min_d = 1e9f, min_u = 0.0f;
    uint32_t min_k = 1u;
    v2 prev = p0;
    for (uint32_t k = 1u; k <= BZ_NSEG; ++k) {
        v2 curr = bezier_at(p0, p1, p2, p3, (float)k / (float)BZ_NSEG);
        v2 seg = v2m(curr, prev);
        float len2 = maxf_(v2d(seg, seg), 1e-8f);
        float uu = clampf_(v2d(v2m(p, prev), seg) / len2, 0.0f, 1.0f);
        v2 proj = v2a(prev, v2s(seg, uu));
        v2 dv = v2m(p, proj);
        float d = sqrtf(v2d(dv, dv));
        if (d < min_d) { min_d = d; min_k = k; min_u = uu; }
        prev = curr;
    }
    float t = ((float)(min_k - 1u) + min_u) / (float)BZ_NSEG;
    float dt = t - 0.5f;
    float pressure = 1.0f - 4.0f * dt * dt;
    float lw = width * pressure, ls = soft * pressure, lo = opacity * pressure;
    float a_geom = 1.0f - smoothstep_(lw - ls, lw + ls, min_d);
    e.a = clampf_(a_geom * lo, 0.0f, 0.999f);
    float B[4];
    bernstein(t, B);
    float raw_w = ((B[0] * q0.z + B[1] * q1.z) + B[2] * q2.z) + B[3] * q3.z;
    float lin = maxf_(raw_w, BZ_NEAR);
    e.d_val = clampf_(1.0f - BZ_NEAR / lin, 0.0f, 1.0f);
    e.k = min_k;
    return e;
}

static inline float tex_luma(const float* t, int W, int H, int x, int y) {
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    const float* c = t + 4 * ((size_t)y * W + x);
    return (c[0] * 0.2126f + c[1] * 0.7152f) + c[2] * 0.0722f;
}
static inline float tex_avg(const float* t, int W, int H, int x, int y) {
    x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
    y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    const float* c = t + 4 * ((size_t)y * W + x);
    return (c[0] * 0.333f + c[1] * 0.333f) + c[2] * 0.333f;
}

/*
 * bezier_backward.wgsl::main over the given tiles (lists from ora_bz_bin).
 * Textures are H*W*4 floats: tgt (rgb; in edge mode .r is the edge target),
 * tgt_depth (.r), bg (rgb), bg_depth (.r), normal (rgb).  Outputs: grads
 * (i32[18N], added), loss_acc (int64[N], 2^-32 fixed point, added),
 * pixel_loss (i32[W*H], added), c_out (H*W*3), d_out (H*W), t_out (H*W, final
 * transmittance), loss[2] = sum of |C - tgt|^2 over colour channels and sum |C - tgt|.
 * Any output pointer may be NULL.
 */
int ora_bz_backward_f32(const float* bz, int n, const bz_uniforms* u, const float* tgt, const float* tgt_depth,
                        const float* bg, const float* bg_depth, const float* normal, int W, int H,
                        const int32_t* tiles, int ntiles, const uint32_t* lists, const int32_t* counts,
                        int32_t* grads, int64_t* loss_acc, int32_t* pixel_loss, float* c_out, float* d_out,
                        float* t_out, double* loss) {
    const float aspect = (float)W / (float)H;
    const uint32_t ntx = ((uint32_t)W + 15u) / 16u;
    const float lw3[3] = { 0.2126f, 0.7152f, 0.0722f };
    double l2 = 0.0, l1 = 0.0;
    /* Tiles run in parallel: every accumulation below is an integer wrap-around
     * sum (order independent) or a per-pixel write; the two loss sums are
     * reported only. */
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : l2, l1)
    for (int kk = 0; kk < ntiles; ++kk) {
        float as[BZ_CAP], depths[BZ_CAP], Ts[BZ_CAP + 1];
        uint32_t segs[BZ_CAP];
        uint32_t tile = (uint32_t)tiles[kk], tx = tile % ntx, ty = tile / ntx;
        const uint32_t* list = lists + (size_t)kk * BZ_CAP;
        const int cnt = counts[kk];
        for (uint32_t ly = 0; ly < 16u; ++ly)
            for (uint32_t lx = 0; lx < 16u; ++lx) {
                uint32_t px = tx * 16u + lx, py = ty * 16u + ly;
                if (px >= (uint32_t)W || py >= (uint32_t)H) continue;
                const size_t pix = (size_t)py * W + px;
                v2 p = pixel_to_p(px, py, W, H, aspect);
                const float* tg = tgt + 4 * pix;
                const float td = tgt_depth[4 * pix];
                Ts[0] = 1.0f;
                float C[3] = { 0.0f, 0.0f, 0.0f }, D = 0.0f;
                for (int e = 0; e < cnt; ++e) {
                    const float* b = bz + 20 * (size_t)list[e];
                    bz_eval ev = bz_pair(b, u->vp, aspect, p);
                    as[e] = ev.a;
                    segs[e] = ev.k;
                    depths[e] = ev.d_val;
                    for (int c = 0; c < 3; ++c) C[c] = C[c] + (Ts[e] * ev.a) * b[16 + c];
                    D = D + (Ts[e] * ev.a) * ev.d_val;
                    Ts[e + 1] = Ts[e] * (1.0f - ev.a);
                }
                const int color_mode = u->mode > 0.5f;
                float bgc[3] = { 0.0f, 0.0f, 0.0f }, bgd = 1.0f;
                if (color_mode) {
                    for (int c = 0; c < 3; ++c) bgc[c] = bg[4 * pix + c];
                    bgd = bg_depth[4 * pix];
                }
                const float Tn = Ts[cnt];
                for (int c = 0; c < 3; ++c) C[c] = C[c] + Tn * bgc[c];
                D = D + Tn * bgd;
                if (c_out) for (int c = 0; c < 3; ++c) c_out[3 * pix + c] = C[c];
                if (d_out) d_out[pix] = D;
                if (t_out) t_out[pix] = Tn;
                float dC_raw[3];
                for (int c = 0; c < 3; ++c) {
                    dC_raw[c] = 2.0f * (C[c] - tg[c]);
                    double er = (double)C[c] - (double)tg[c];
                    l2 += er * er;
                    l1 += fabs(er);
                }
                const int fg = td < 0.99f;
                const float dD_depth = (fg && color_mode) ? 0.5f * 2.0f * (D - td) : 0.0f;
                float h[3] = { dC_raw[0] * 0.5f, dC_raw[1] * 0.5f, dC_raw[2] * 0.5f };
                float luma_err = (h[0] * lw3[0] + h[1] * lw3[1]) + h[2] * lw3[2];
                float lr = tex_luma(tgt, W, H, (int)px + 1, (int)py), ll = tex_luma(tgt, W, H, (int)px - 1, (int)py);
                float lu = tex_luma(tgt, W, H, (int)px, (int)py + 1), ld = tex_luma(tgt, W, H, (int)px, (int)py - 1);
                float contrast = sqrtf((lr - ll) * (lr - ll) + (lu - ld) * (lu - ld));
                float cw = 1.0f + clampf_(contrast * 8.0f, 0.0f, 3.0f);
                float dr_l = (dC_raw[0] * lw3[0] + dC_raw[1] * lw3[1]) + dC_raw[2] * lw3[2];
                float dC[3];
                for (int c = 0; c < 3; ++c) {
                    float luma_c = luma_err * lw3[c] * 6.0f;
                    float chroma = dC_raw[c] - dr_l * lw3[c];
                    dC[c] = (luma_c + chroma) * cw;
                }
                const float dD_total = dD_depth;
                float dT = ((dC[0] * bgc[0] + dC[1] * bgc[1]) + dC[2] * bgc[2]) + dD_total * bgd;
                const float coverage = 1.0f - Tn;
                const float d_cov = (u->mode < 0.5f) ? 2.0f * 2.0f * (coverage - tg[0]) : 0.0f;
                dT = dT + -d_cov;
                const float is_bg = td >= 0.995f ? 1.0f : 0.0f;
                /* direction flow field of the pixel (fine layer only) */
                const int is_fine = u->max_width > 0.0f;
                v2 flow_dir = { 0.0f, 0.0f };
                int have_flow = 0;
                if (is_fine) {
                    int ix = (int)px, iy = (int)py;
                    float cr = tex_avg(tgt, W, H, ix + 2, iy), cl = tex_avg(tgt, W, H, ix - 2, iy);
                    float cu = tex_avg(tgt, W, H, ix, iy + 2), cd = tex_avg(tgt, W, H, ix, iy - 2);
                    v2 gc = { (cr - cl) * 0.25f * aspect, -(cu - cd) * 0.25f };
                    float nr = tex_avg(normal, W, H, ix + 2, iy), nl = tex_avg(normal, W, H, ix - 2, iy);
                    float nu = tex_avg(normal, W, H, ix, iy + 2), nd = tex_avg(normal, W, H, ix, iy - 2);
                    v2 gn = { (nr - nl) * 0.25f * aspect, -(nu - nd) * 0.25f };
                    v2 fr = v2a(gc, v2s(gn, 2.0f));
                    float fl = sqrtf(v2d(fr, fr));
                    if (fl > 1e-4f) {
                        flow_dir.x = fr.x / fl;
                        flow_dir.y = fr.y / fl;
                        have_flow = 1;
                    }
                }
                for (int j = 0; j < cnt; ++j) {
                    int idx = cnt - 1 - j;
                    float a = as[idx];
                    if (a < 0.001f) continue;
                    uint32_t i = list[idx];
                    const float* b = bz + 20 * (size_t)i;
                    float width = maxf_(b[3], 0.001f), soft = maxf_(b[7], 0.001f), opacity = b[19];
                    v3 q0 = project_center(u->vp, b + 0, aspect), q1 = project_center(u->vp, b + 4, aspect);
                    v3 q2 = project_center(u->vp, b + 8, aspect), q3 = project_center(u->vp, b + 12, aspect);
                    v2 p0 = { q0.x, q0.y }, p1 = { q1.x, q1.y }, p2 = { q2.x, q2.y }, p3 = { q3.x, q3.y };
                    float T_prev = Ts[idx], d_val = depths[idx];
                    float dColor[3];
                    for (int c = 0; c < 3; ++c) dColor[c] = dC[c] * (T_prev * a);
                    float da = (dT * (-T_prev) + ((dC[0] * (T_prev * b[16]) + dC[1] * (T_prev * b[17])) +
                                                  dC[2] * (T_prev * b[18]))) + dD_total * (T_prev * d_val);
                    dT = (dT * (1.0f - a) + ((dC[0] * (a * b[16]) + dC[1] * (a * b[17])) + dC[2] * (a * b[18]))) +
                         dD_total * (a * d_val);
                    uint32_t k = segs[idx];
                    float t_prev = (float)(k - 1u) / (float)BZ_NSEG, t_curr = (float)k / (float)BZ_NSEG;
                    v2 prev_pt = bezier_at(p0, p1, p2, p3, t_prev), curr_pt = bezier_at(p0, p1, p2, p3, t_curr);
                    v2 seg = v2m(curr_pt, prev_pt);
                    float len2 = maxf_(v2d(seg, seg), 1e-8f);
                    float uc = clampf_(v2d(v2m(p, prev_pt), seg) / len2, 0.0f, 1.0f);
                    v2 proj = v2a(prev_pt, v2s(seg, uc));
                    v2 dvec = v2m(p, proj);
                    float d = maxf_(sqrtf(v2d(dvec, dvec)), 1e-6f);
                    float t_pix = ((float)(k - 1u) + uc) / (float)BZ_NSEG;
                    float dtp = t_pix - 0.5f;
                    float pressure = 1.0f - 4.0f * dtp * dtp;
                    float Bp[4];
                    bernstein(t_pix, Bp);
                    float lwid = width * pressure, lsoft = soft * pressure, lop = opacity * pressure;
                    float inner = lwid - lsoft, outer = lwid + lsoft;
                    float denom = maxf_(outer - inner, 1e-6f);
                    float x_in = clampf_((d - inner) / denom, 0.0f, 1.0f);
                    float ss_d = 6.0f * x_in * (1.0f - x_in) / denom;
                    int in_band = (d > inner) && (d < outer);
                    float a_geom = 1.0f - smoothstep_(inner, outer, d);
                    float d_op = da * a_geom * pressure + u->bg_penalty * is_bg;
                    float da_eff = da * lop;
                    float dDd = in_band ? -da_eff * ss_d : 0.0f;
                    float dWidth = in_band ? da_eff * ss_d * pressure / denom : 0.0f;
                    float dmi = clampf_(d - inner, 0.0f, denom);
                    float dx_ds = pressure * (denom - 2.0f * dmi) / maxf_(denom * denom, 1e-12f);
                    float dSoft = in_band ? -da_eff * ss_d * dx_ds : 0.0f;
                    float w0 = q0.z, w1 = q1.z, w2 = q2.z, w3 = q3.z;
                    float raw_w = ((Bp[0] * w0 + Bp[1] * w1) + Bp[2] * w2) + Bp[3] * w3;
                    float lin = maxf_(raw_w, BZ_NEAR);
                    float dd_dw = BZ_NEAR / (lin * lin);
                    float dZ[4];
                    for (int c = 0; c < 4; ++c) dZ[c] = dD_total * (T_prev * a) * dd_dw * Bp[c];
                    v2 dProj;
                    dProj.x = -dDd * dvec.x / d;
                    dProj.y = -dDd * dvec.y / d;
                    v2 dPrev = v2s(dProj, 1.0f - uc), dCurr = v2s(dProj, uc);
                    float Bpr[4], Bcu[4];
                    bernstein(t_prev, Bpr);
                    bernstein(t_curr, Bcu);
                    int32_t* g = grads + 18 * (size_t)i;
                    if (is_fine) dSoft = dSoft + 5.0f * 2.0f * soft;
                    if (is_fine && len2 > 1e-10f && have_flow) {
                        float il = sqrtf(len2);
                        v2 tang = { seg.x / il, seg.y / il };
                        float tgd = v2d(tang, flow_dir);
                        float dld = 1.5f * 2.0f * tgd;
                        v2 dtv = v2s(flow_dir, dld);
                        float inv_len = 1.0f / sqrtf(len2);
                        float dtt = v2d(dtv, tang);
                        v2 dseg = { (dtv.x - tang.x * dtt) * inv_len, (dtv.y - tang.y * dtt) * inv_len };
                        v2 dpd = { -dseg.x, -dseg.y };
                        const float* cp[4] = { b + 0, b + 4, b + 8, b + 12 };
                        for (int c = 0; c < 4; ++c) {
                            v2 g2 = v2a(v2s(dpd, Bpr[c]), v2s(dseg, Bcu[c]));
                            v3 g3 = backproject(u->vp, cp[c], aspect, g2);
                            gadd(g + 3 * c + 0, f2i(g3.x * BZ_FP_POS));
                            gadd(g + 3 * c + 1, f2i(g3.y * BZ_FP_POS));
                            gadd(g + 3 * c + 2, f2i(g3.z * BZ_FP_POS));
                        }
                    }
                    const float* cp[4] = { b + 0, b + 4, b + 8, b + 12 };
                    for (int c = 0; c < 4; ++c) {
                        v2 g2 = v2a(v2s(dPrev, Bpr[c]), v2s(dCurr, Bcu[c]));
                        v3 g3 = backproject(u->vp, cp[c], aspect, g2);
                        gadd(g + 3 * c + 0, f2i(g3.x * BZ_FP_POS));
                        gadd(g + 3 * c + 1, f2i(g3.y * BZ_FP_POS));
                        gadd(g + 3 * c + 2, f2i((g3.z + dZ[c]) * BZ_FP_POS));
                    }
                    for (int c = 0; c < 3; ++c) gadd(g + 12 + c, f2i(dColor[c] * BZ_FP_COL));
                    gadd(g + 15, f2i(d_op * BZ_FP_COL));
                    gadd(g + 16, f2i(dWidth * BZ_FP_POS));
                    gadd(g + 17, f2i(dSoft * BZ_FP_POS));
                    float lc = ((dC[0] * dC[0] + dC[1] * dC[1]) + dC[2] * dC[2]) * (T_prev * a);
                    if (loss_acc) {
                        uint64_t* la = (uint64_t*)(loss_acc + i);
                        uint64_t dl = (uint64_t)f2l(lc * BZ_LOSS_FP);
#pragma omp atomic
                        *la += dl;
                    }
                }
                float residual = ((h[0] * h[0] + h[1] * h[1]) + h[2] * h[2]) * Tn;
                if (pixel_loss) pixel_loss[pix] = wrap_add(pixel_loss[pix], f2i(residual * 10000.0f));
            }
    }
    if (loss) { loss[0] = l2; loss[1] = l1; }
    return 0;
}

/*
 * bezier_step.wgsl::main for every curve: Adam on the 18 fixed-point gradients
 * (drained to 0), clamps, grad_accum, prune and off-screen kill.  t is the step
 * number this update uses (adam.t + 1 in the reference); pixel_count = W*H of
 * the view the gradients came from (GpuBezierOptimizerManager.ts:375-380).
 */
int ora_bz_step_f32(float* bz, int n, int32_t* grads, float* m, float* v, float* grad_accum, const bz_uniforms* u,
                    float t, float pixel_count, float no_kill) {
    static const float lr_t[18] = { 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f,
                                    0.005f, 0.005f, 0.005f, 0.01f, 0.01f, 0.01f, 0.005f, 0.002f, 0.002f };
    static const float mu_t[18] = { 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.005f,
                                    0.005f, 0.005f, 0.005f, 0.01f, 0.01f, 0.01f, 0.005f, 0.003f, 0.003f };
    static const float fps_t[18] = { 10000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f,
                                     10000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f,
                                     100000.0f, 100000.0f, 100000.0f, 100000.0f, 10000.0f, 10000.0f };
    /* parameter lp -> float slot in the 20-float curve */
    static const int slot[18] = { 0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 16, 17, 18, 19, 3, 7 };
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - rm_powf(beta1, t), bc2 = 1.0f - rm_powf(beta2, t);
    const float width_hi = u->max_width > 0.0f ? u->max_width : 0.1f;
    for (int i = 0; i < n; ++i) {
        float* b = bz + 20 * (size_t)i;
        int32_t* g = grads + 18 * (size_t)i;
        if (b[19] < BZ_KILL) {
            for (int lp = 0; lp < 18; ++lp) g[lp] = 0;
            continue;
        }
        float pn2 = 0.0f;
        for (int lp = 0; lp < 18; ++lp) {
            size_t pi = 18 * (size_t)i + lp;
            int32_t raw = g[lp];
            g[lp] = 0;
            float pixel_norm = 1.0f / maxf_(pixel_count, 1.0f);
            float grad = (float)raw / fps_t[lp] * pixel_norm;
            float mm = beta1 * m[pi] + (1.0f - beta1) * grad;
            float vv = beta2 * v[pi] + (1.0f - beta2) * grad * grad;
            m[pi] = mm;
            v[pi] = vv;
            float m_hat = mm / bc1, v_hat = vv / bc2;
            float raw_update = lr_t[lp] * m_hat / (sqrtf(v_hat) + eps);
            float upd = clampf_(raw_update, -mu_t[lp], mu_t[lp]);
            float lo = lp < 12 ? -1e9f : (lp < 16 ? 0.0f : 0.001f);
            float hi = lp < 12 ? 1e9f : (lp < 15 ? 1.0f : (lp == 15 ? 0.99f : (lp == 16 ? width_hi : 0.03f)));
            b[slot[lp]] = clampf_(b[slot[lp]] - upd, lo, hi);
            if (lp <= 11) pn2 = pn2 + grad * grad;
        }
        grad_accum[i] = grad_accum[i] + sqrtf(pn2);
        float at = u->prune_alpha > 0.0f ? u->prune_alpha : BZ_PRUNE_ALPHA;
        float wt = u->prune_width > 0.0f ? u->prune_width : BZ_PRUNE_WIDTH;
        if (b[19] < at || b[3] <= wt) b[19] = 0.0f;
        if (b[19] > 0.0f && no_kill < 0.5f) {
            float cx[4], cy[4], cw[4];
            for (int c = 0; c < 4; ++c) {
                const float* P = b + 4 * c;
                const float* M = u->vp;
                cx[c] = ((M[0] * P[0] + M[4] * P[1]) + M[8] * P[2]) + M[12];
                cy[c] = ((M[1] * P[0] + M[5] * P[1]) + M[9] * P[2]) + M[13];
                cw[c] = ((M[3] * P[0] + M[7] * P[1]) + M[11] * P[2]) + M[15];
            }
            const float mg = BZ_OFFSCREEN;
            int L = 1, R = 1, Bo = 1, A = 1, Be = 1;
            for (int c = 0; c < 4; ++c) {
                L = L && (cx[c] < -mg * cw[c]);
                R = R && (cx[c] > mg * cw[c]);
                Bo = Bo && (cy[c] < -mg * cw[c]);
                A = A && (cy[c] > mg * cw[c]);
                Be = Be && (cw[c] < 0.0f);
            }
            if (L || R || Bo || A || Be) b[19] = 0.0f;
        }
    }
    return 0;
}

/* ------------------------------------------------------------ densification */
static inline float len3(const float* a, const float* b) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrtf((dx * dx + dy * dy) + dz * dz);
}
static inline float hash_fract(float x) { return rm_fractf(rm_sinf(x) * 43758.5453f); }

static void clip_of(const float* M, const float* P, float c[4]) {
    c[0] = ((M[0] * P[0] + M[4] * P[1]) + M[8] * P[2]) + M[12];
    c[1] = ((M[1] * P[0] + M[5] * P[1]) + M[9] * P[2]) + M[13];
    c[2] = ((M[2] * P[0] + M[6] * P[1]) + M[10] * P[2]) + M[14];
    c[3] = ((M[3] * P[0] + M[7] * P[1]) + M[11] * P[2]) + M[15];
}
/* bezier_adc.wgsl:67-72 */
static int check_offscreen(const float c[4]) {
    if (c[3] < 0.0f) return 1;
    float w = maxf_(c[3], 1e-5f);
    float nx = c[0] / w, ny = c[1] / w;
    return fabsf(nx) > BZ_OFFSCREEN || fabsf(ny) > BZ_OFFSCREEN;
}

/* Split (de Casteljau at t = 1/2) or jittered clone of curve i into slot j
 * (bezier_adc.wgsl:134-181); zeroes the Adam state of both. */
static void split_or_clone(float* bz, float* m, float* v, uint32_t i, uint32_t j, float adam_t) {
    float* b = bz + 20 * (size_t)i;
    float nb[20];
    memcpy(nb, b, sizeof nb);
    const float *p0 = b, *p1 = b + 4, *p2 = b + 8, *p3 = b + 12;
    float mid[3];
    for (int c = 0; c < 3; ++c) mid[c] = (((p0[c] + 3.0f * p1[c]) + 3.0f * p2[c]) + p3[c]) * 0.125f;
    float len_approx = len3(mid, p0) + len3(p3, mid);
    if (len_approx > BZ_SPLIT_LEN) {
        float q0[3], q1[3], q2[3], r0[3], r1[3], s[3];
        for (int c = 0; c < 3; ++c) {
            q0[c] = (p0[c] + p1[c]) * 0.5f;
            q1[c] = (p1[c] + p2[c]) * 0.5f;
            q2[c] = (p2[c] + p3[c]) * 0.5f;
        }
        for (int c = 0; c < 3; ++c) {
            r0[c] = (q0[c] + q1[c]) * 0.5f;
            r1[c] = (q1[c] + q2[c]) * 0.5f;
        }
        for (int c = 0; c < 3; ++c) s[c] = (r0[c] + r1[c]) * 0.5f;
        float p3c[3] = { p3[0], p3[1], p3[2] };
        for (int c = 0; c < 3; ++c) {
            /* first half stays in slot i (p0 and both w's kept, p2.w/p3.w = 0) */
            b[4 + c] = q0[c];
            b[8 + c] = r0[c];
            b[12 + c] = s[c];
            nb[0 + c] = s[c];
            nb[4 + c] = r1[c];
            nb[8 + c] = q2[c];
            nb[12 + c] = p3c[c];
        }
        b[11] = 0.0f;
        b[15] = 0.0f;
        nb[11] = 0.0f;
        nb[15] = 0.0f;
    } else {
        float seed = (float)i * 3.14159f + adam_t;
        float jx = (hash_fract(seed * 12.9898f) - 0.5f) * 0.001f;
        float jy = (hash_fract(seed * 78.233f) - 0.5f) * 0.001f;
        float jz = (hash_fract(seed * 43.123f) - 0.5f) * 0.001f;
        for (int k = 0; k < 4; ++k) {
            nb[4 * k + 0] = b[4 * k + 0] + jx;
            nb[4 * k + 1] = b[4 * k + 1] + jy;
            nb[4 * k + 2] = b[4 * k + 2] + jz;
        }
        nb[11] = 0.0f;
        nb[15] = 0.0f;
    }
    memcpy(bz + 20 * (size_t)j, nb, sizeof nb);
    for (int p = 0; p < 18; ++p) {
        m[18 * (size_t)i + p] = 0.0f;
        v[18 * (size_t)i + p] = 0.0f;
        m[18 * (size_t)j + p] = 0.0f;
        v[18 * (size_t)j + p] = 0.0f;
    }
}

/* bezier_adc.wgsl:52-65 with the view's W x H in place of the compiled 256 x 256 */
static void pixel_to_world(const float* vp_inv, uint32_t px_idx, int W, int H, float spawn_depth, float out[3]) {
    uint32_t px = px_idx % (uint32_t)W, py = px_idx / (uint32_t)W;
    float ux = ((float)px + 0.5f) / (float)W, uy = ((float)py + 0.5f) / (float)H;
    float nx = ux * 2.0f - 1.0f, ny = 1.0f - uy * 2.0f;
    float w = 0.1f / maxf_(1.0f - spawn_depth, 1e-5f);
    float P[3] = { nx * w, ny * w, w };
    float c[4];
    /* vp_inv * (x, y, w, w): the last column multiplies w, not 1 */
    const float* M = vp_inv;
    c[0] = ((M[0] * P[0] + M[4] * P[1]) + M[8] * P[2]) + M[12] * w;
    c[1] = ((M[1] * P[0] + M[5] * P[1]) + M[9] * P[2]) + M[13] * w;
    c[2] = ((M[2] * P[0] + M[6] * P[1]) + M[10] * P[2]) + M[14] * w;
    c[3] = ((M[3] * P[0] + M[7] * P[1]) + M[11] * P[2]) + M[15] * w;
    out[0] = c[0] / c[3];
    out[1] = c[1] / c[3];
    out[2] = c[2] / c[3];
}

/*
 * bezier_adc.wgsl::main.  loss_acc in 2^-32 fixed point (read as float);
 * pixel_loss is W*H.  stats[6] = {dead slots found, killed, split/clone pass 1.5,
 * spawned, split/clone pass 3, dead slots left}.
 */
int ora_bz_adc_f32(float* bz, int n, float* m, float* v, float* grad_accum, int64_t* loss_acc, int32_t* pixel_loss,
                   int W, int H, const bz_uniforms* u, float adam_t, float no_kill, int64_t* stats) {
    uint32_t* dead = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    if (!dead) return -1;
    uint32_t dc = 0;
    int64_t st[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < n; ++i)
        if (bz[20 * (size_t)i + 19] < BZ_KILL) dead[dc++] = (uint32_t)i;
    st[0] = dc;
    /* pass 1: kill stuck / high-loss and off-screen curves */
    for (int i = 0; i < n; ++i) {
        float* b = bz + 20 * (size_t)i;
        if (b[19] < BZ_KILL) continue;
        float gn = grad_accum[i] / BZ_ADC_PERIOD;
        float ln = (float)((double)loss_acc[i] * 0x1p-32) / BZ_ADC_PERIOD;
        if (no_kill < 0.5f && gn <= BZ_TAU_POS && ln > BZ_TAU_LOSS) {
            b[19] = 0.0f;
            dead[dc++] = (uint32_t)i;
            st[1]++;
            continue;
        }
        float c0[4], c1[4], c2[4], c3[4];
        clip_of(u->vp, b + 0, c0);
        clip_of(u->vp, b + 4, c1);
        clip_of(u->vp, b + 8, c2);
        clip_of(u->vp, b + 12, c3);
        if (check_offscreen(c0) && check_offscreen(c1) && check_offscreen(c2) && check_offscreen(c3)) {
            b[19] = 0.0f;
            dead[dc++] = (uint32_t)i;
            st[1]++;
        }
    }
    /* pass 1.5: split or clone high-gradient curves into dead slots */
    for (int i = 0; i < n; ++i) {
        if (bz[20 * (size_t)i + 19] < BZ_KILL) continue;
        if (grad_accum[i] / BZ_ADC_PERIOD <= BZ_TAU_POS) continue;
        if (dc == 0) continue;
        dc--;
        split_or_clone(bz, m, v, (uint32_t)i, dead[dc], adam_t);
        st[2]++;
    }
    /* pass 2: spawn at the highest-loss pixels */
    const uint32_t npx = (uint32_t)W * (uint32_t)H;
    uint32_t max_spawns = dc < BZ_MAX_SPAWNS ? dc : BZ_MAX_SPAWNS;
    uint32_t stride = npx / 4096u;
    if (stride < 1u) stride = 1u;
    for (uint32_t sp = 0; sp < max_spawns; ++sp) {
        if (dc == 0) break;
        uint32_t best_px = 0;
        int32_t best_val = 0;
        uint32_t start = (sp * 997u) % stride;
        for (uint32_t px = start; px < npx; px += stride) {
            int32_t val = pixel_loss[px];
            if (val > best_val) { best_val = val; best_px = px; }
        }
        uint32_t spawn_px = best_px;
        if (best_val <= 0) {
            float seed_fb = (float)sp * 1234.5678f + adam_t * 0.1f;
            uint32_t rx = (uint32_t)(hash_fract(seed_fb * 12.9898f) * (float)W);
            uint32_t ry = (uint32_t)(hash_fract(seed_fb * 78.233f) * (float)H);
            spawn_px = ry * (uint32_t)W + rx;
        } else {
            pixel_loss[best_px] = 0;
        }
        dc--;
        uint32_t slot_i = dead[dc];
        float center[3];
        pixel_to_world(u->vp_inv, spawn_px, W, H, 0.5f, center);
        float seed = (float)spawn_px * 1.61803f + (float)sp * 2.71828f;
        float angle = hash_fract(seed * 127.1f) * 6.28318f;
        float tx = rm_cosf(angle) * BZ_SPAWN_TLEN, tz = rm_sinf(angle) * BZ_SPAWN_TLEN;
        float tg[3] = { tx, 0.0f, tz };
        float* nb = bz + 20 * (size_t)slot_i;
        for (int c = 0; c < 3; ++c) {
            nb[0 + c] = center[c] - tg[c];
            nb[4 + c] = center[c] - tg[c] * 0.33f;
            nb[8 + c] = center[c] + tg[c] * 0.33f;
            nb[12 + c] = center[c] + tg[c];
        }
        nb[3] = BZ_SPAWN_WIDTH;
        nb[7] = BZ_SPAWN_SOFT;
        nb[11] = 0.0f;
        nb[15] = 0.0f;
        nb[16] = nb[17] = nb[18] = nb[19] = 0.5f;
        for (int p = 0; p < 18; ++p) {
            m[18 * (size_t)slot_i + p] = 0.0f;
            v[18 * (size_t)slot_i + p] = 0.0f;
        }
        st[3]++;
    }
    /* pass 3: clone / split high-gradient live curves into what is left */
    for (int i = 0; i < n; ++i) {
        if (dc == 0) break;
        if (bz[20 * (size_t)i + 19] < BZ_KILL) continue;
        if (grad_accum[i] / BZ_ADC_PERIOD <= BZ_TAU_POS) continue;
        dc--;
        split_or_clone(bz, m, v, (uint32_t)i, dead[dc], adam_t);
        st[4]++;
    }
    for (uint32_t px = 0; px < npx; ++px) pixel_loss[px] = 0;
    for (int i = 0; i < n; ++i) {
        grad_accum[i] = 0.0f;
        loss_acc[i] = 0;
    }
    st[5] = dc;
    if (stats) memcpy(stats, st, sizeof st);
    free(dead);
    return 0;
}

/*
 * Display raster of the curves: bezier_forward.wgsl:47-182 with the blend state of
 * GpuBezierForwardPipelineManager.ts:79-96, brute force per pixel.
 *   - instances drawn from curve N-1 down to 0 (bezier_forward.wgsl:61);
 *   - a curve with any control point at clip w <= 0 draws nothing (:74-81);
 *   - its quad covers the pixel centres inside the closed box hull(P0..P3) -+ pad,
 *     pad = max(width, 1e-4) + max(softness, 1e-4), clamped to [-4, 4] (:88-96);
 *     the fragment position is the pixel centre (raster edge rules and varying
 *     interpolation are not modelled);
 *   - closest of the 16 polyline segments with the strict <, and the signed cross
 *     distance on that segment (:133-152);
 *   - pressure taper, smoothstep edge, brush .r at (t, cross / (w_l + s_l) / 2 + 1/2)
 *     (:154-173), a = clamp(a_geom * brush * opacity, 0, 0.999), discard a < 0.001;
 *   - premultiplied output (rgb a, a) blended with (one, one-minus-src-alpha):
 *     rgb = rgb*a + dst*(1-a), alpha = a + dst_alpha*(1-a).
 * The sampler (linear filtering, repeat in u, clamp-to-edge in v, one mip level,
 * GpuBezierForwardPipelineManager.ts:34-41) is restated as FP32 bilinear filtering
 * of the texel values at x = u*bw - 0.5, y = v*bh - 0.5; hardware filter weights
 * are implementation-defined.  brush: bh rows of bw texel values (r / 255).
 * rgba: H*W*4 in/out, cleared to 0 when clear != 0; quantize != 0 rounds every
 * channel to k/255 after each blend (rgba8unorm target).
 */
float ora_bz_brush_sample(const float* tex, int bw, int bh, float u, float v) {
    float x = u * (float)bw - 0.5f, y = v * (float)bh - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int ix = (int)x0f, iy = (int)y0f;
    int x0 = ((ix % bw) + bw) % bw, x1 = (((ix + 1) % bw) + bw) % bw;
    int y0 = iy < 0 ? 0 : (iy > bh - 1 ? bh - 1 : iy);
    int y1 = iy + 1 < 0 ? 0 : (iy + 1 > bh - 1 ? bh - 1 : iy + 1);
    float t00 = tex[(size_t)y0 * bw + x0], t10 = tex[(size_t)y0 * bw + x1];
    float t01 = tex[(size_t)y1 * bw + x0], t11 = tex[(size_t)y1 * bw + x1];
    float top = t00 * (1.0f - fx) + t10 * fx;
    float bot = t01 * (1.0f - fx) + t11 * fx;
    return top * (1.0f - fy) + bot * fy;
}

int ora_bz_render_f32(const float* bz, int n, const float* vp, int W, int H, const float* brush, int bw, int bh,
                      float* rgba, int clear, int quantize) {
    if (bw <= 0 || bh <= 0) return -1;
    float aspect = (float)W / (float)H;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py) {
        for (int px = 0; px < W; ++px) {
            float* c = rgba + 4 * ((size_t)py * W + px);
            if (clear) c[0] = c[1] = c[2] = c[3] = 0.0f;
            v2 p = pixel_to_p((uint32_t)px, (uint32_t)py, (uint32_t)W, (uint32_t)H, aspect);
            for (int ii = 0; ii < n; ++ii) {
                const float* b = bz + 20 * (size_t)(n - 1 - ii);
                v3 q0 = project_center(vp, b + 0, aspect), q1 = project_center(vp, b + 4, aspect);
                v3 q2 = project_center(vp, b + 8, aspect), q3 = project_center(vp, b + 12, aspect);
                if (q0.z <= 0.0f || q1.z <= 0.0f || q2.z <= 0.0f || q3.z <= 0.0f) continue;
                v2 p0 = { q0.x, q0.y }, p1 = { q1.x, q1.y }, p2 = { q2.x, q2.y }, p3 = { q3.x, q3.y };
                float width = maxf_(b[3], 0.0001f), soft = maxf_(b[7], 0.0001f);
                float pad = width + soft;
                float mnx = maxf_(minf_(minf_(p0.x, p1.x), minf_(p2.x, p3.x)) - pad, -4.0f);
                float mny = maxf_(minf_(minf_(p0.y, p1.y), minf_(p2.y, p3.y)) - pad, -4.0f);
                float mxx = minf_(maxf_(maxf_(p0.x, p1.x), maxf_(p2.x, p3.x)) + pad, 4.0f);
                float mxy = minf_(maxf_(maxf_(p0.y, p1.y), maxf_(p2.y, p3.y)) + pad, 4.0f);
                if (!(p.x >= mnx && p.x <= mxx && p.y >= mny && p.y <= mxy)) continue;
                float min_d = 1e9f, min_u = 0.0f, min_cross = 0.0f;
                uint32_t min_k = 1u;
                v2 prev = p0;
                for (uint32_t k = 1u; k <= BZ_NSEG; ++k) {
                    v2 curr = bezier_at(p0, p1, p2, p3, (float)k / (float)BZ_NSEG);
                    v2 seg = v2m(curr, prev);
                    float len2 = maxf_(v2d(seg, seg), 1e-8f);
                    float uu = clampf_(v2d(v2m(p, prev), seg) / len2, 0.0f, 1.0f);
                    v2 proj = v2a(prev, v2s(seg, uu));
                    v2 dv = v2m(p, proj);
                    float d = sqrtf(v2d(dv, dv));
                    if (d < min_d) {
                        min_d = d;
                        min_k = k;
                        min_u = uu;
                        float seg_len = sqrtf(len2);
                        v2 dir = { seg.x / seg_len, seg.y / seg_len };
                        min_cross = dv.x * (-dir.y) + dv.y * dir.x;
                    }
                    prev = curr;
                }
                float t = ((float)(min_k - 1u) + min_u) / (float)BZ_NSEG;
                float dt = t - 0.5f;
                float pressure = 1.0f - 4.0f * dt * dt;
                float lw = width * pressure, ls = soft * pressure, lo = b[19] * pressure;
                float a_geom = 1.0f - smoothstep_(lw - ls, lw + ls, min_d);
                float bv = clampf_(min_cross / maxf_(lw + ls, 1e-6f) * 0.5f + 0.5f, 0.0f, 1.0f);
                float br = ora_bz_brush_sample(brush, bw, bh, t, bv);
                float a = clampf_(a_geom * br * lo, 0.0f, 0.999f);
                if (a < 0.001f) continue;
                float one_m = 1.0f - a;
                c[0] = b[16] * a + c[0] * one_m;
                c[1] = b[17] * a + c[1] * one_m;
                c[2] = b[18] * a + c[2] * one_m;
                c[3] = a + c[3] * one_m;
                if (quantize)
                    for (int k = 0; k < 4; ++k) c[k] = rintf(clampf_(c[k], 0.0f, 1.0f) * 255.0f) / 255.0f;
            }
        }
    }
    return 0;
}
