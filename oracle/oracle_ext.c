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
