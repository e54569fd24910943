/*
 * nps_oracle.c -- CPU oracle for the splat rasterizer and optimizer loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (nprush_amd) never links it.
 *
 * It restates, in plain C, the reference's WGSL kernels:
 *   ora_bin           tile binning block of splat_backward.wgsl:141-211
 *   ora_backward_f32  splat_backward.wgsl:136-457 (binning + forward + loss +
 *                     reverse pass + i32 fixed-point atomics), FP32 REF semantics
 *   ora_backward_f64  the same algebra in FP64 (no fixed point), for error
 *                     statements and finite differences
 *   ora_step_f32      splat_step.wgsl:32-106 (fused Adam, clamp, renorm, kill)
 *   ora_adc_f32       splat_adc.wgsl:27-87 (serial clone/split into dead slots)
 *   ora_render_f32    splat_forward.wgsl:42-156 + the blend state of
 *                     GpuSplatForwardPipelineManager.ts:54-82,114-137
 * Transcendentals in FP32 mode come from ref_math.h (see there).
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "ref_math.h"

/* ---------------- FP32 instantiation (bit-level REF semantics) ---------------- */
#define R float
#define SFX(name) name##_f32
#define REXP2 rm_exp2f
#define REXP rm_expf
#define RLOG2SQRT rm_log2_sqrtf
#define RLN2 RM_LN2F
#define GRAD_T uint32_t
/* atomicAdd(&grads[k], i32(g * FP)) with wrapping two's-complement sums */
#define GRAD_ADD(p, g, fp)                                   \
    do {                                                     \
        uint32_t v_ = (uint32_t)rm_f2i_sat((g) * (fp));      \
        _Pragma("omp atomic") *(p) += v_;                    \
    } while (0)
#include "oracle_core.inc"
#undef R
#undef SFX
#undef REXP2
#undef REXP
#undef RLOG2SQRT
#undef RLN2
#undef GRAD_T
#undef GRAD_ADD

/* ---------------- FP64 twin ---------------- */
#define R double
#define SFX(name) name##_f64
#define REXP2 exp2
#define REXP exp
#define RLOG2SQRT(x) (0.5 * log2(x))
#define RLN2 0.69314718055994530942
#define GRAD_T double
#define GRAD_ADD(p, g, fp) do { (void)(fp); *(p) += (g); } while (0)
#include "oracle_core.inc"
#undef R
#undef SFX
#undef REXP2
#undef REXP
#undef RLOG2SQRT
#undef RLN2
#undef GRAD_T
#undef GRAD_ADD

#define TILE 16u

/* Tile rectangle in aspect-corrected NDC (BW:148-153). */
static void tile_rect(uint32_t tx, uint32_t ty, uint32_t W, uint32_t H, float aspect,
                      float* tmin_x, float* tmin_y, float* tmax_x, float* tmax_y) {
    uint32_t x0 = tx * TILE, y0 = ty * TILE;
    uint32_t x1 = x0 + TILE < W ? x0 + TILE : W;
    uint32_t y1 = y0 + TILE < H ? y0 + TILE : H;
    v2_f32 p00 = pixel_to_p_f32(x0, y0, W, H, aspect);
    v2_f32 p11 = pixel_to_p_f32(x1, y1, W, H, aspect);
    *tmin_x = p00.x; *tmin_y = p11.y;
    *tmax_x = p11.x; *tmax_y = p00.y;
}

/*
 * Binning of one tile (BW:141-211): every splat is tested; the first `cap`
 * ascending ids are kept.  Returns the uncapped count in *full.
 */
static uint32_t bin_tile(const float* splats, uint32_t n, const float* vp, uint32_t W, uint32_t H,
                         uint32_t tx, uint32_t ty, uint32_t cap, uint32_t* ids, uint32_t* full) {
    float aspect = (float)W / (float)H;
    float tmin_x, tmin_y, tmax_x, tmax_y;
    tile_rect(tx, ty, W, H, aspect, &tmin_x, &tmin_y, &tmax_x, &tmax_y);
    uint32_t count = 0, all = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float* s = splats + 16 * (size_t)i;
        if (s[7] < 0.005f) continue;
        v4_f32 clip = mat_vec_f32(vp, s[0], s[1], s[2], 1.0f);
        float px = clip.x / clip.w * aspect, py = clip.y / clip.w, pz = clip.w;
        if (pz < 0.0f) continue;
        float sx = fmax__f32(s[3], 0.0001f);
        float sy = fmax__f32(s[12], 0.0001f);
        float sz = fmax__f32(s[15], 0.0001f);
        float safe_sb = fmax__f32(s[14], 0.0001f);
        float Rr = rm_powf(15.0f / safe_sb, 1.0f / s[13]);
        v3_f32 ex = { 1, 0, 0 }, ey = { 0, 1, 0 }, ez = { 0, 0, 1 };
        v3_f32 ax_w = quat_rotate_f32(s + 8, ex);
        v3_f32 ay_w = quat_rotate_f32(s + 8, ey);
        v3_f32 az_w = quat_rotate_f32(s + 8, ez);
        float w = clip.w;
        v2_f32 ax_s = project_axis_f32(vp, ax_w, clip.x, clip.y, w, aspect);
        v2_f32 ay_s = project_axis_f32(vp, ay_w, clip.x, clip.y, w, aspect);
        v2_f32 az_s = project_axis_f32(vp, az_w, clip.x, clip.y, w, aspect);
        float m0x = ax_s.x * sx, m0y = ax_s.y * sx;
        float m1x = ay_s.x * sy, m1y = ay_s.y * sy;
        float m2x = az_s.x * sz, m2y = az_s.y * sz;
        float cov00 = (m0x * m0x + m1x * m1x) + m2x * m2x;
        float cov11 = (m0y * m0y + m1y * m1y) + m2y * m2y;
        float filter_std = 0.3f * (2.0f / (float)H);
        float filter2 = filter_std * filter_std;
        cov00 = cov00 + filter2;
        cov11 = cov11 + filter2;
        float max_r = Rr * sqrtf(fmax__f32(fmax__f32(cov00, cov11), 1e-9f));
        float sminx = px - max_r, sminy = py - max_r;
        float smaxx = px + max_r, smaxy = py + max_r;
        if (!(sminx > tmax_x || smaxx < tmin_x || sminy > tmax_y || smaxy < tmin_y)) {
            if (count < cap) ids[count++] = i;
            all++;
        }
    }
    if (full) *full = all;
    return count;
}

static void tile_xy(uint32_t t, uint32_t W, uint32_t* tx, uint32_t* ty) {
    uint32_t ntx = (W + TILE - 1) / TILE;
    *tx = t % ntx; *ty = t / ntx;
}

/* Binning for a set of tiles (tile id = ty*ceil(W/16)+tx).  ids_out is ntiles*cap. */
int ora_bin(const float* splats, int n, const float* vp, int W, int H, const int32_t* tiles, int ntiles,
            int cap, uint32_t* ids_out, int32_t* count_out, int32_t* full_out) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < ntiles; ++k) {
        uint32_t tx, ty, full;
        tile_xy((uint32_t)tiles[k], (uint32_t)W, &tx, &ty);
        uint32_t c = bin_tile(splats, (uint32_t)n, vp, (uint32_t)W, (uint32_t)H, tx, ty, (uint32_t)cap,
                              ids_out + (size_t)k * cap, &full);
        count_out[k] = (int32_t)c;
        if (full_out) full_out[k] = (int32_t)full;
    }
    return 0;
}

/*
 * FP32 REF backward over a set of tiles (all tiles when tiles == NULL).
 * grads: int32[16n], accumulated with wrap-around.  c_out: H*W*3, d_out: H*W
 * (may be NULL).  loss[0] = sum L2 (as the reference's gradient defines it),
 * loss[1] = sum L1 of colour.  Both summed in double.
 */
int ora_backward_f32(const float* splats, int n, const float* vp, const float* vp_inv,
                     const float* tgt_rgba, const float* tgt_depth, int W, int H, int cap,
                     const int32_t* tiles, int ntiles, int32_t* grads, float* c_out, float* d_out,
                     double* loss, int32_t* full_counts) {
    uint32_t ntx = ((uint32_t)W + TILE - 1) / TILE, nty = ((uint32_t)H + TILE - 1) / TILE;
    int total = tiles ? ntiles : (int)(ntx * nty);
    double l2 = 0.0, l1 = 0.0;
    #pragma omp parallel reduction(+ : l2, l1)
    {
        uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
        float* alphas = (float*)malloc(sizeof(float) * (size_t)cap);
        float* Ts = (float*)malloc(sizeof(float) * ((size_t)cap + 1));
        #pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < total; ++k) {
            uint32_t t = tiles ? (uint32_t)tiles[k] : (uint32_t)k;
            uint32_t tx = t % ntx, ty = t / ntx, full;
            uint32_t count = bin_tile(splats, (uint32_t)n, vp, (uint32_t)W, (uint32_t)H, tx, ty,
                                      (uint32_t)cap, ids, &full);
            if (full_counts) full_counts[k] = (int32_t)full;
            for (uint32_t ly = 0; ly < TILE; ++ly)
                for (uint32_t lx = 0; lx < TILE; ++lx) {
                    uint32_t px = tx * TILE + lx, py = ty * TILE + ly;
                    if (px >= (uint32_t)W || py >= (uint32_t)H) continue;
                    size_t pix = (size_t)py * (size_t)W + px;
                    const float* tg = tgt_rgba + 4 * pix;
                    pixel_backward_f32(splats, vp, vp_inv, ids, count, px, py, (uint32_t)W, (uint32_t)H,
                                       tg[0], tg[1], tg[2], tgt_depth[pix], alphas, Ts,
                                       (uint32_t*)grads, c_out ? c_out + 3 * pix : NULL,
                                       d_out ? d_out + pix : NULL, &l2, &l1, 0);
                }
        }
        free(ids); free(alphas); free(Ts);
    }
    (void)nty;
    if (loss) { loss[0] = l2; loss[1] = l1; }
    return 0;
}

/*
 * FP64 twin.  The tile lists come from ora_bin (lists: ntiles*cap, counts).
 * grads: double[16n] (unscaled, no truncation) or NULL for forward + loss only.
 * flags: ORA_NO_REG (drop the shape regulariser), ORA_NO_PULL (drop the 3D pull).
 */
int ora_backward_f64(const double* splats, int n, const double* vp, const double* vp_inv,
                     const float* tgt_rgba, const float* tgt_depth, int W, int H,
                     const int32_t* tiles, int ntiles, const uint32_t* lists, const int32_t* counts, int cap,
                     double* grads, double* c_out, double* d_out, double* loss, int flags) {
    uint32_t ntx = ((uint32_t)W + TILE - 1) / TILE;
    int maxc = 0;
    for (int k = 0; k < ntiles; ++k) if (counts[k] > maxc) maxc = counts[k];
    double l2 = 0.0, l1 = 0.0;
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    double* tgrads = NULL;
    if (grads) tgrads = (double*)calloc((size_t)nthreads * 16 * (size_t)n, sizeof(double));
    #pragma omp parallel reduction(+ : l2, l1)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* alphas = (double*)malloc(sizeof(double) * ((size_t)maxc + 1));
        double* Ts = (double*)malloc(sizeof(double) * ((size_t)maxc + 2));
        double* g = grads ? tgrads + (size_t)tid * 16 * (size_t)n : NULL;
        #pragma omp for schedule(static)
        for (int k = 0; k < ntiles; ++k) {
            uint32_t t = (uint32_t)tiles[k];
            uint32_t tx = t % ntx, ty = t / ntx;
            const uint32_t* ids = lists + (size_t)k * cap;
            for (uint32_t ly = 0; ly < TILE; ++ly)
                for (uint32_t lx = 0; lx < TILE; ++lx) {
                    uint32_t px = tx * TILE + lx, py = ty * TILE + ly;
                    if (px >= (uint32_t)W || py >= (uint32_t)H) continue;
                    size_t pix = (size_t)py * (size_t)W + px;
                    const float* tg = tgt_rgba + 4 * pix;
                    pixel_backward_f64(splats, vp, vp_inv, ids, (uint32_t)counts[k], px, py, (uint32_t)W,
                                       (uint32_t)H, tg[0], tg[1], tg[2], tgt_depth[pix], alphas, Ts, g,
                                       c_out ? c_out + 3 * pix : NULL, d_out ? d_out + pix : NULL,
                                       &l2, &l1, flags);
                }
        }
        free(alphas); free(Ts);
    }
    if (grads) {
        for (size_t j = 0; j < 16 * (size_t)n; ++j) {
            double acc = 0.0;
            for (int t = 0; t < nthreads; ++t) acc += tgrads[(size_t)t * 16 * (size_t)n + j];
            grads[j] = acc;
        }
        free(tgrads);
    }
    if (loss) { loss[0] = l2; loss[1] = l1; }
    return 0;
}

/* splat_step.wgsl:48-52,85-86 tables */
static const float LR_T[16] = { 0.0005f, 0.0005f, 0.0005f, 0.01f, 0.02f, 0.02f, 0.02f, 0.01f,
                                0.005f, 0.005f, 0.005f, 0.005f, 0.01f, 0.01f, 0.01f, 0.01f };
static const float MU_T[16] = { 0.005f, 0.005f, 0.005f, 0.005f, 0.001f, 0.001f, 0.001f, 0.0005f,
                                0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.05f, 0.05f, 0.005f };
static const float FPS_T[16] = { 10000.0f, 10000.0f, 10000.0f, 10000.0f, 100000.0f, 100000.0f, 100000.0f,
                                 100000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f, 10000.0f,
                                 10000.0f, 10000.0f };
static const float LO_T[16] = { -1e9f, -1e9f, -1e9f, 0.001f, 0.05f, 0.05f, 0.05f, 0.01f,
                                -1e9f, -1e9f, -1e9f, -1e9f, 0.001f, 0.1f, 0.01f, 0.001f };
static const float HI_T[16] = { 1e9f, 1e9f, 1e9f, 2.0f, 1.0f, 1.0f, 1.0f, 0.99f,
                                1e9f, 1e9f, 1e9f, 1e9f, 2.0f, 10.0f, 5.0f, 2.0f };

/*
 * splat_step.wgsl:32-106 for every splat.  *adam_t is the stored step counter;
 * t = *adam_t + 1 is used by every splat (the reference's cross-workgroup race
 * on adam.t, STEP:35-37, is defined away) and written back.
 */
int ora_step_f32(float* splats, int32_t* grads, float* m, float* v, float* accum, int n, float* adam_t) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float t = *adam_t + 1.0f;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        float* s = splats + 16 * (size_t)i;
        float pos_grad_norm2 = 0.0f;
        for (int lp = 0; lp < 16; ++lp) {
            size_t pi = 16 * (size_t)i + (size_t)lp;
            int32_t raw = grads[pi];
            grads[pi] = 0;
            float grad = (float)raw / FPS_T[lp] / 16384.0f;
            float mm = m[pi], vv = v[pi];
            mm = b1 * mm + (1.0f - b1) * grad;
            vv = b2 * vv + ((1.0f - b2) * grad) * grad;
            m[pi] = mm; v[pi] = vv;
            float m_hat = mm / (1.0f - rm_powf(b1, t));
            float v_hat = vv / (1.0f - rm_powf(b2, t));
            float raw_update = (LR_T[lp] * m_hat) / (sqrtf(v_hat) + eps);
            float update = clamp__f32(raw_update, -MU_T[lp], MU_T[lp]);
            s[lp] = clamp__f32(s[lp] - update, LO_T[lp], HI_T[lp]);
            if (lp <= 2) pos_grad_norm2 = pos_grad_norm2 + grad * grad;
        }
        float q_len = fmax__f32(sqrtf(((s[8] * s[8] + s[9] * s[9]) + s[10] * s[10]) + s[11] * s[11]), 1e-8f);
        s[8] = s[8] / q_len; s[9] = s[9] / q_len; s[10] = s[10] / q_len; s[11] = s[11] / q_len;
        accum[i] = accum[i] + sqrtf(pos_grad_norm2);
        float volume = (s[3] * s[12]) * s[15];
        if (s[7] < 0.05f || volume < 1e-7f) s[7] = 0.0f;
    }
    *adam_t = t;
    return 0;
}

/*
 * splat_adc.wgsl:27-87, serial, literally.  adam_t is the counter after the
 * step (adam.t).  stats[0]=qualifiers, [1]=children written, [2]=splits,
 * [3]=clones (may be NULL).
 */
int ora_adc_f32(float* splats, float* m, float* v, float* accum, int n, float adam_t, int32_t* stats) {
    uint32_t next_dead = 0;
    const float PERIOD = 25.0f, TAU = 0.00005f, SPLIT_T = 0.01f, KILL = 0.05f;
    int32_t nq = 0, nc = 0, ns = 0, ncl = 0;
    for (uint32_t i = 0; i < (uint32_t)n; ++i) {
        float s[16];
        memcpy(s, splats + 16 * (size_t)i, sizeof(s));
        if (s[7] < KILL) continue;
        float ga = accum[i] / PERIOD;
        accum[i] = 0.0f;
        if (ga > TAU) {
            nq++;
            float scale_norm = sqrtf((s[3] * s[3] + s[12] * s[12]) + s[15] * s[15]);
            int found = 0;
            uint32_t new_idx = 0;
            for (uint32_t d = next_dead; d < (uint32_t)n; ++d) {
                if (splats[16 * (size_t)d + 7] < KILL) { new_idx = d; found = 1; next_dead = d + 1; break; }
            }
            if (!found) next_dead = (uint32_t)n;
            if (found) {
                float ns_[16];
                memcpy(ns_, s, sizeof(s));
                if (scale_norm > SPLIT_T) {
                    const float f = 0.625f;
                    s[3] = s[3] * f; s[12] = s[12] * f; s[15] = s[15] * f;
                    ns_[3] = ns_[3] * f; ns_[12] = ns_[12] * f; ns_[15] = ns_[15] * f;
                    s[0] = s[0] - s[3] * 0.05f;
                    ns_[0] = ns_[0] + s[3] * 0.05f;
                    ns++;
                } else {
                    float seed = (float)i * 3.14159f + adam_t;
                    ns_[0] = ns_[0] + (rm_fractf(rm_sinf(seed * 12.9898f) * 43758.5453f) - 0.5f) * 0.002f;
                    ns_[1] = ns_[1] + (rm_fractf(rm_sinf(seed * 78.233f) * 43758.5453f) - 0.5f) * 0.002f;
                    ncl++;
                }
                memcpy(splats + 16 * (size_t)i, s, sizeof(s));
                memcpy(splats + 16 * (size_t)new_idx, ns_, sizeof(s));
                for (int p = 0; p < 16; ++p) {
                    m[16 * (size_t)i + p] = 0.0f; v[16 * (size_t)i + p] = 0.0f;
                    m[16 * (size_t)new_idx + p] = 0.0f; v[16 * (size_t)new_idx + p] = 0.0f;
                }
                nc++;
            }
        }
    }
    if (stats) { stats[0] = nq; stats[1] = nc; stats[2] = ns; stats[3] = ncl; }
    return 0;
}

/*
 * splat_forward.wgsl + blend state, brute force per pixel: instances are drawn
 * from splat N-1 down to 0, each covering the pixel centres inside its quad
 * [centre +- extent] (closed), discarding a < 0.001, blending
 *   rgb = src*a + dst*(1-a),  alpha = a + dst_alpha*(1-a)
 * into colour (rgb, a) and depth (enc,enc,enc, a) targets.  color/depth are
 * H*W*4 in/out (cleared to (0.05,0.05,0.05,1)/(1,1,1,1) when clear != 0).
 * quantize != 0 rounds every channel to k/255 after each blend (rgba8unorm).
 */
int ora_render_f32(const float* splats, int n, const float* vp, int W, int H, float* color, float* depth,
                   int clear, int quantize) {
    float aspect = (float)W / (float)H;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py) {
        for (int px = 0; px < W; ++px) {
            size_t pix = (size_t)py * W + px;
            float* c = color + 4 * pix;
            float* d = depth + 4 * pix;
            if (clear) {
                /* the clear value as the target stores it */
                float cv = quantize ? rintf(0.05f * 255.0f) / 255.0f : 0.05f;
                c[0] = c[1] = c[2] = cv; c[3] = 1.0f; d[0] = d[1] = d[2] = d[3] = 1.0f;
            }
            v2_f32 p = pixel_to_p_f32((uint32_t)px, (uint32_t)py, (uint32_t)W, (uint32_t)H, aspect);
            for (int ii = 0; ii < n; ++ii) {
                const float* s = splats + 16 * (size_t)(n - 1 - ii);
                float sx = fmax__f32(s[3], 0.0001f), sy = fmax__f32(s[12], 0.0001f), sz = fmax__f32(s[15], 0.0001f);
                float shape_a = s[13], shape_b = s[14];
                float Rr = rm_powf(15.0f / fmax__f32(shape_b, 0.0001f), 1.0f / shape_a);
                v4_f32 clip = mat_vec_f32(vp, s[0], s[1], s[2], 1.0f);
                float w = clip.w;
                if (w < 0.1f) continue;
                v3_f32 ex = { 1, 0, 0 }, ey = { 0, 1, 0 }, ez = { 0, 0, 1 };
                v2_f32 ax_s = project_axis_f32(vp, quat_rotate_f32(s + 8, ex), clip.x, clip.y, w, aspect);
                v2_f32 ay_s = project_axis_f32(vp, quat_rotate_f32(s + 8, ey), clip.x, clip.y, w, aspect);
                v2_f32 az_s = project_axis_f32(vp, quat_rotate_f32(s + 8, ez), clip.x, clip.y, w, aspect);
                float m0x = ax_s.x * sx, m0y = ax_s.y * sx;
                float m1x = ay_s.x * sy, m1y = ay_s.y * sy;
                float m2x = az_s.x * sz, m2y = az_s.y * sz;
                float cov00 = (m0x * m0x + m1x * m1x) + m2x * m2x;
                float cov01 = (m0x * m0y + m1x * m1y) + m2x * m2y;
                float cov11 = (m0y * m0y + m1y * m1y) + m2y * m2y;
                float filter_std = 0.3f * (2.0f / (float)H);
                float filter2 = filter_std * filter_std;
                cov00 = cov00 + filter2;
                cov11 = cov11 + filter2;
                float det = cov00 * cov11 - cov01 * cov01;
                float inv_det = (fabsf(det) < 1e-10f) ? 0.0f : 1.0f / det;
                float A = cov11 * inv_det, B = -cov01 * inv_det, C = cov00 * inv_det;
                float ext_x = Rr * sqrtf(fmax__f32(cov00, 1e-9f));
                float ext_y = Rr * sqrtf(fmax__f32(cov11, 1e-9f));
                float scx = clip.x / w * aspect, scy = clip.y / w;
                float dx = p.x - scx, dy = p.y - scy;
                if (!(dx >= -ext_x && dx <= ext_x && dy >= -ext_y && dy <= ext_y)) continue;
                float r2 = ((A * dx) * dx + ((2.0f * B) * dx) * dy) + (C * dy) * dy;
                float lr = rm_log2_sqrtf(fmax__f32(r2, 0.0001f));
                float pw = -shape_b * rm_exp2f(shape_a * lr);
                float a = (pw > -15.0f) ? rm_expf(pw) * s[7] : 0.0f;
                a = clamp__f32(a, 0.0f, 0.999f);
                if (a < 0.001f) continue;
                float lin = fmax__f32(w, 0.1f);
                float enc = clamp__f32(1.0f - 0.1f / lin, 0.0f, 1.0f);
                float one_m = 1.0f - a;
                c[0] = s[4] * a + c[0] * one_m;
                c[1] = s[5] * a + c[1] * one_m;
                c[2] = s[6] * a + c[2] * one_m;
                c[3] = a + c[3] * one_m;
                d[0] = enc * a + d[0] * one_m;
                d[1] = enc * a + d[1] * one_m;
                d[2] = enc * a + d[2] * one_m;
                d[3] = a + d[3] * one_m;
                if (quantize) {
                    for (int k = 0; k < 4; ++k) {
                        c[k] = rintf(clamp__f32(c[k], 0.0f, 1.0f) * 255.0f) / 255.0f;
                        d[k] = rintf(clamp__f32(d[k], 0.0f, 1.0f) * 255.0f) / 255.0f;
                    }
                }
            }
        }
    }
    return 0;
}

/* Exposed transcendentals so tests can compare them with the product's. */
float ora_exp2f(float x) { return rm_exp2f(x); }
float ora_log2f(float x) { return rm_log2f(x); }
float ora_powf(float x, float y) { return rm_powf(x, y); }
float ora_expf(float x) { return rm_expf(x); }
float ora_logf(float x) { return rm_logf(x); }
float ora_sinf(float x) { return rm_sinf(x); }
float ora_cosf(float x) { return rm_cosf(x); }
int32_t ora_f2i(float x) { return rm_f2i_sat(x); }

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
