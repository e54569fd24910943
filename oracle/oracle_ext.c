/*
 * oracle_ext.c -- CPU oracle of the EXT splat path (test infrastructure only).
 *
 * EXT is the north-star extension of the reference's splat path, with no
 * reference function to restate (SURVEY.md's a18 row).  Its definition, which
 * the GPU path implements:
 *   splat  s[16] = [x y z opacity_logit][log_sx log_sy log_sz 0][qw qx qy qz][0 0 0 0]
 *          sh[48] = 16 RGB coefficients (degree d uses the first (d+1)^2)
 *   scales exp(log_s); opacity sigmoid(logit); rotation by the normalised
 *   quaternion (quat_rotate of splat_backward.wgsl:42-45); the reference's
 *   projection of the scaled axes (project_axis, BW:52-58) and covariance
 *   M M^T + (0.3 px)^2 I with conic (BW:109-126); cull w < 0.1;
 *   colour = max(SH(d) + 0.5, 0), d = normalise(mean - camera centre), real SH
 *   basis to degree 3;  alpha = min(0.99, opacity exp(-r^2/2)), r^2 = d^T Sigma^-1 d;
 *   a pixel sees a splat only if its centre lies in the splat's 3-sigma box;
 *   pairs with alpha < 1/255 skipped; front-to-back in increasing view depth w
 *   (ties by id); a pixel stops before the splat that would bring T below 1e-4;
 *   background 0.05; loss L1 = sum |C - target|;
 *   gradients: exact derivative of that loss (conic and SH direction depend on
 *   the mean as well), FP64 accumulation in the oracle.
 * Transcendentals in FP32 mode: ref_math.h.  Both modes are compiled from
 * oracle_ext_core.inc.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "ref_math.h"

/* FP32 mode: the GPU's operations.  Pair falloff exp(-r2/2) = exp2(r2 * c) with
 * c = -0.5 log2(e) (exact in FP32), products and sums of r2, the transmittance
 * and the colour as the explicit fused multiply-adds the kernel uses. */
#define EXT_NEG_HALF_LOG2E (-0.5f * 0x1.715476p+0f)
#define R float
#define SFX(n) n##_xf32
#define REXP rm_expf
#define RSQRT sqrtf
#define RFMA fmaf
#define RGAUSS(r2) rm_exp2f((r2) * EXT_NEG_HALF_LOG2E)
#include "oracle_ext_core.inc"
#undef R
#undef SFX
#undef REXP
#undef RSQRT
#undef RFMA
#undef RGAUSS

#define R double
#define SFX(n) n##_xf64
#define REXP exp
#define RSQRT sqrt
#define RFMA fma
#define RGAUSS(r2) exp(-0.5 * (r2))
#include "oracle_ext_core.inc"
#undef R
#undef SFX
#undef REXP
#undef RSQRT
#undef RFMA
#undef RGAUSS

/* Camera centre from the inverse view-projection: VP^-1 (0,0,1,0) is the
 * homogeneous centre, i.e. column 2 of vp_inv, dehomogenised in double. */
static void cam_from_inv(const float* vi, double* c) {
    c[0] = (double)vi[8] / (double)vi[11];
    c[1] = (double)vi[9] / (double)vi[11];
    c[2] = (double)vi[10] / (double)vi[11];
}

int ora_ext_bin(const float* splats, const float* sh, int n, int deg, const float* vp, const float* vp_inv, int W,
                int H, int tile, const int32_t* tiles, int ntiles, int stride, uint32_t* lists, int32_t* counts) {
    double c[3];
    cam_from_inv(vp_inv, c);
    float cam[3] = { (float)c[0], (float)c[1], (float)c[2] };
    ext_bin_xf32(splats, sh, n, deg, vp, cam, (uint32_t)W, (uint32_t)H, (uint32_t)tile, tiles, ntiles, stride, lists,
                 counts);
    return 0;
}

int ora_ext_backward_f32(const float* splats, const float* sh, int n, int deg, const float* vp, const float* vp_inv,
                         const float* tgt, int W, int H, int tile, const int32_t* tiles, int ntiles, const uint32_t* lists,
                         const int32_t* counts, int stride, double* grads, double* gsh, float* c_out, double* loss) {
    double c[3];
    cam_from_inv(vp_inv, c);
    float cam[3] = { (float)c[0], (float)c[1], (float)c[2] };
    ext_backward_xf32(splats, sh, n, deg, vp, cam, tgt, (uint32_t)W, (uint32_t)H, (uint32_t)tile, tiles, ntiles, lists,
                      counts, stride, NULL, grads, gsh, c_out, loss);
    return 0;
}

/* Whole-view binning in O(keys) (CSR): returns the key count; with ids == NULL
 * only offs[T + 1] is filled. */
int64_t ora_ext_bin_all(const float* splats, const float* sh, int n, int deg, const float* vp, const float* vp_inv,
                        int W, int H, int tile, int64_t* offs, uint32_t* ids) {
    double c[3];
    cam_from_inv(vp_inv, c);
    float cam[3] = { (float)c[0], (float)c[1], (float)c[2] };
    return ext_bin_all_xf32(splats, sh, n, deg, vp, cam, (uint32_t)W, (uint32_t)H, (uint32_t)tile, offs, ids);
}

/* ora_ext_backward_f32 over CSR lists (from ora_ext_bin_all): every tile of the view. */
int ora_ext_backward_csr_f32(const float* splats, const float* sh, int n, int deg, const float* vp,
                             const float* vp_inv, const float* tgt, int W, int H, int tile, const uint32_t* ids,
                             const int64_t* offs, double* grads, double* gsh, float* c_out, double* loss) {
    double c[3];
    cam_from_inv(vp_inv, c);
    float cam[3] = { (float)c[0], (float)c[1], (float)c[2] };
    const int ntx = (W + tile - 1) / tile, nty = (H + tile - 1) / tile, T = ntx * nty;
    int32_t* tiles = (int32_t*)malloc(sizeof(int32_t) * (size_t)T);
    int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)T);
    for (int t = 0; t < T; ++t) { tiles[t] = t; counts[t] = (int32_t)(offs[t + 1] - offs[t]); }
    ext_backward_xf32(splats, sh, n, deg, vp, cam, tgt, (uint32_t)W, (uint32_t)H, (uint32_t)tile, tiles, T, ids,
                      counts, 0, offs, grads, gsh, c_out, loss);
    free(tiles); free(counts);
    return 0;
}

/* The FP64 twin over CSR lists: every tile of the view, FP64 splats, SH, VP
 * and camera centre (for the error statement of the FP32 paths; the lists are
 * the FP32 binning's). */
int ora_ext_backward_csr_f64(const double* splats, const double* sh, int n, int deg, const double* vp,
                             const double* cam, const float* tgt, int W, int H, int tile, const uint32_t* ids,
                             const int64_t* offs, double* grads, double* gsh, double* c_out, double* loss) {
    const int ntx = (W + tile - 1) / tile, nty = (H + tile - 1) / tile, T = ntx * nty;
    int32_t* tiles = (int32_t*)malloc(sizeof(int32_t) * (size_t)T);
    int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)T);
    for (int t = 0; t < T; ++t) { tiles[t] = t; counts[t] = (int32_t)(offs[t + 1] - offs[t]); }
    ext_backward_xf64(splats, sh, n, deg, vp, cam, tgt, (uint32_t)W, (uint32_t)H, (uint32_t)tile, tiles, T, ids,
                      counts, 0, offs, grads, gsh, c_out, loss);
    free(tiles); free(counts);
    return 0;
}

int ora_ext_backward_f64(const double* splats, const double* sh, int n, int deg, const double* vp,
                         const double* cam, const float* tgt, int W, int H, int tile, const int32_t* tiles, int ntiles,
                         const uint32_t* lists, const int32_t* counts, int stride, double* grads, double* gsh,
                         double* c_out, double* loss) {
    ext_backward_xf64(splats, sh, n, deg, vp, cam, tgt, (uint32_t)W, (uint32_t)H, (uint32_t)tile, tiles, ntiles, lists,
                      counts, stride, NULL, grads, gsh, c_out, loss);
    return 0;
}

/* EXT Adam (FP32): 64 parameters per splat (16 of the splat record + 48 SH),
 * learning rates per parameter group, bias correction with t = *adam_t + 1,
 * eps 1e-15; parameters without a learning rate (pads, SH above the degree)
 * are left alone.  grads are FP32 [n][64] (record then SH). */
static float ext_lr(int p, int deg) {
    if (p < 3) return 1.6e-4f;
    if (p == 3) return 5e-2f;
    if (p >= 4 && p < 7) return 5e-3f;
    if (p >= 8 && p < 12) return 1e-3f;
    if (p >= 16) {
        int k = (p - 16) / 3;
        if (k >= (deg + 1) * (deg + 1)) return 0.0f;
        return k == 0 ? 2.5e-3f : 2.5e-3f / 20.0f;
    }
    return 0.0f;
}

int ora_ext_adam_f32(float* splats, float* sh, const float* grads, float* m, float* v, int n, int deg, float* adam_t) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-15f;
    float t = *adam_t + 1.0f;
    float bc1 = 1.0f - rm_powf(b1, t), bc2 = 1.0f - rm_powf(b2, t);
    for (int i = 0; i < n; ++i) {
        for (int p = 0; p < 64; ++p) {
            float lr = ext_lr(p, deg);
            if (lr == 0.0f) continue;
            size_t k = 64 * (size_t)i + p;
            float g = grads[k];
            float mm = b1 * m[k] + (1.0f - b1) * g;
            float vv = b2 * v[k] + ((1.0f - b2) * g) * g;
            m[k] = mm; v[k] = vv;
            float upd = (lr * (mm / bc1)) / (sqrtf(vv / bc2) + eps);
            float* x = p < 16 ? splats + 16 * (size_t)i + p : sh + 48 * (size_t)i + (p - 16);
            *x = *x - upd;
        }
    }
    *adam_t = t;
    return 0;
}

/* ---------------------------------------------------------------- EXT ADC
 * Densification of the EXT model (no reference function; 3D Gaussian
 * splatting's clone/split/prune made deterministic):
 *   op = sigmoid(logit); a slot is dead if op < prune_opacity;
 *   a live slot qualifies if its view count c > 0 and g / c >= threshold
 *   (g = summed |dL/d ndc| of its centre since the last ADC);
 *   the k-th qualifier (ascending index) takes the k-th dead slot (ascending);
 *   qualifiers without a dead slot are left alone;
 *   split if max_k exp(log_s_k) > percent_dense * extent, else clone;
 *   split: two children replace the parent (child 0 in the parent's slot,
 *     child 1 in the dead slot): centre + sum_k axis_k * (z_k * s_k), z_k
 *     standard normal (Box-Muller on a counter hash of (seed, slot, child, k)),
 *     log-scales reduced by ln 1.6, every other parameter and the SH copied;
 *   clone: an exact copy in the dead slot;
 *   Adam m, v of every child slot are zeroed (the clone's parent keeps its own);
 *   dead slots that receive nothing get logit -1e4 and zero m, v;
 *   accumulators are reset.
 * stats[5] = qualifiers, children placed, splits, clones, dead slots. */
static uint32_t lowbias32(uint32_t x) {   /* Wellons' lowbias32 integer hash */
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
static float adc_normal(uint32_t seed, uint32_t slot, uint32_t child, uint32_t k) {
    uint32_t h1 = lowbias32(seed ^ lowbias32(slot * 8u + child * 4u + k));
    uint32_t h2 = lowbias32(h1 ^ 0x9e3779b9u);
    float u1 = (float)((h1 >> 8) + 1u) * 0x1p-24f;   /* (0, 1] */
    float u2 = (float)(h2 >> 8) * 0x1p-24f;          /* [0, 1) */
    float r = sqrtf(-2.0f * rm_logf(u1));
    return r * rm_sinf(6.2831855f * u2 + 1.5707964f);
}
#define EXT_LN_1_6 0.47000363f
#define EXT_DEAD_LOGIT -10000.0f

int ora_ext_adc_f32(float* splats, float* sh, float* m, float* v, float* accum, int n, float threshold, float extent,
                    float prune_opacity, float percent_dense, uint32_t seed, int64_t* stats) {
    int32_t* dead = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* qual = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int nd = 0, nq = 0;
    for (int i = 0; i < n; ++i) {
        float op = 1.0f / (1.0f + rm_expf(-splats[16 * (size_t)i + 3]));
        if (op < prune_opacity) { dead[nd++] = i; continue; }
        float c = accum[2 * (size_t)i + 1];
        if (c > 0.0f && accum[2 * (size_t)i] / c >= threshold) qual[nq++] = i;
    }
    int P = nq < nd ? nq : nd;
    int64_t nsplit = 0, nclone = 0;
    for (int k = 0; k < P; ++k) {
        int q = qual[k], d = dead[k];
        float* sp = splats + 16 * (size_t)q;
        float* sd = splats + 16 * (size_t)d;
        float s[3];
        for (int a = 0; a < 3; ++a) s[a] = rm_expf(sp[4 + a]);
        float smax = s[0] > s[1] ? s[0] : s[1];
        smax = smax > s[2] ? smax : s[2];
        memcpy(sh + 48 * (size_t)d, sh + 48 * (size_t)q, 48 * sizeof(float));
        if (smax > percent_dense * extent) {
            float qq = ((sp[8] * sp[8] + sp[9] * sp[9]) + sp[10] * sp[10]) + sp[11] * sp[11];
            float ql = sqrtf(qq > 1e-24f ? qq : 1e-24f);
            float qw = sp[8] / ql, qx = sp[9] / ql, qy = sp[10] / ql, qz = sp[11] / ql;
            float ax[3][3] = {
                { 1.0f - 2.0f * (qy * qy + qz * qz), 2.0f * (qx * qy + qw * qz), 2.0f * (qx * qz - qw * qy) },
                { 2.0f * (qx * qy - qw * qz), 1.0f - 2.0f * (qx * qx + qz * qz), 2.0f * (qy * qz + qw * qx) },
                { 2.0f * (qx * qz + qw * qy), 2.0f * (qy * qz - qw * qx), 1.0f - 2.0f * (qx * qx + qy * qy) } };
            float parent[16];
            memcpy(parent, sp, sizeof(parent));
            for (int c = 0; c < 2; ++c) {
                float* out = c == 0 ? sp : sd;
                uint32_t slot = (uint32_t)(c == 0 ? q : d);
                float o[3];
                for (int a = 0; a < 3; ++a) o[a] = adc_normal(seed, slot, (uint32_t)c, (uint32_t)a) * s[a];
                memcpy(out, parent, sizeof(parent));
                for (int j = 0; j < 3; ++j)
                    out[j] = parent[j] + ((ax[0][j] * o[0] + ax[1][j] * o[1]) + ax[2][j] * o[2]);
                for (int a = 0; a < 3; ++a) out[4 + a] = parent[4 + a] - EXT_LN_1_6;
            }
            memset(m + 64 * (size_t)q, 0, 64 * sizeof(float));
            memset(v + 64 * (size_t)q, 0, 64 * sizeof(float));
            nsplit++;
        } else {
            memcpy(sd, sp, 16 * sizeof(float));
            nclone++;
        }
        memset(m + 64 * (size_t)d, 0, 64 * sizeof(float));
        memset(v + 64 * (size_t)d, 0, 64 * sizeof(float));
    }
    for (int k = P; k < nd; ++k) {
        int d = dead[k];
        splats[16 * (size_t)d + 3] = EXT_DEAD_LOGIT;
        memset(m + 64 * (size_t)d, 0, 64 * sizeof(float));
        memset(v + 64 * (size_t)d, 0, 64 * sizeof(float));
    }
    memset(accum, 0, 2 * (size_t)n * sizeof(float));
    if (stats) { stats[0] = nq; stats[1] = P; stats[2] = nsplit; stats[3] = nclone; stats[4] = nd; }
    free(dead); free(qual);
    return 0;
}

/* the normal samples of a split, for tests of their distribution */
float ora_ext_adc_normal(uint32_t seed, uint32_t slot, uint32_t child, uint32_t k) {
    return adc_normal(seed, slot, child, k);
}
