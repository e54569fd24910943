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

#define R float
#define SFX(n) n##_xf32
#define REXP rm_expf
#define RSQRT sqrtf
#include "oracle_ext_core.inc"
#undef R
#undef SFX
#undef REXP
#undef RSQRT

#define R double
#define SFX(n) n##_xf64
#define REXP exp
#define RSQRT sqrt
#include "oracle_ext_core.inc"
#undef R
#undef SFX
#undef REXP
#undef RSQRT

/* Camera centre from the inverse view-projection: VP^-1 (0,0,1,0) is the
 * homogeneous centre, i.e. column 2 of vp_inv, dehomogenised in double. */
static void cam_from_inv(const float* vi, double* c) {
    c[0] = (double)vi[8] / (double)vi[11];
    c[1] = (double)vi[9] / (double)vi[11];
    c[2] = (double)vi[10] / (double)vi[11];
}

int ora_ext_bin(const float* splats, const float* sh, int n, int deg, const float* vp, const float* vp_inv, int W,
                int H, const int32_t* tiles, int ntiles, int stride, uint32_t* lists, int32_t* counts) {
    double c[3];
    cam_from_inv(vp_inv, c);
    float cam[3] = { (float)c[0], (float)c[1], (float)c[2] };
    ext_bin_xf32(splats, sh, n, deg, vp, cam, (uint32_t)W, (uint32_t)H, tiles, ntiles, stride, lists, counts);
    return 0;
}

int ora_ext_backward_f32(const float* splats, const float* sh, int n, int deg, const float* vp, const float* vp_inv,
                         const float* tgt, int W, int H, const int32_t* tiles, int ntiles, const uint32_t* lists,
                         const int32_t* counts, int stride, double* grads, double* gsh, float* c_out, double* loss) {
    double c[3];
    cam_from_inv(vp_inv, c);
    float cam[3] = { (float)c[0], (float)c[1], (float)c[2] };
    ext_backward_xf32(splats, sh, n, deg, vp, cam, tgt, (uint32_t)W, (uint32_t)H, tiles, ntiles, lists, counts, stride,
                      grads, gsh, c_out, loss);
    return 0;
}

int ora_ext_backward_f64(const double* splats, const double* sh, int n, int deg, const double* vp,
                         const double* cam, const float* tgt, int W, int H, const int32_t* tiles, int ntiles,
                         const uint32_t* lists, const int32_t* counts, int stride, double* grads, double* gsh,
                         double* c_out, double* loss) {
    ext_backward_xf64(splats, sh, n, deg, vp, cam, tgt, (uint32_t)W, (uint32_t)H, tiles, ntiles, lists, counts, stride,
                      grads, gsh, c_out, loss);
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
