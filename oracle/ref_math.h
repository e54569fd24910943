/*
 * ref_math.h -- oracle-side deterministic transcendentals (plain C, FP32).
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links this file.
 *
 * The reference evaluates exp/pow/log/sqrt/sin with WGSL built-ins, whose results
 * are implementation-defined to within ULP bounds (WGSL spec: pow(x,y) is
 * "inherited from exp2(y * log2(x))").  To compare a GPU path against the CPU
 * bit-for-bit, both sides must evaluate the SAME function.  This header defines
 * that function from published methods, using only IEEE-754 correctly rounded
 * +,-,*,/, sqrt and explicit fmaf, with no implicit contraction (build with
 * -ffp-contract=off):
 *
 *   exp2f: Cody & Waite argument reduction x = k + f, |f| <= 1/2, then the
 *          Taylor series of e^(f ln2) to degree 7 (truncation < 2^-27 relative),
 *          Horner with fused multiply-add (fmaf, exactly specified by IEEE-754),
 *          then exact scaling by 2^k.  Results below 2^-126 flush to 0.
 *   log2f: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log2 m = f P(f), f = m-1,
 *          P the degree-9 Chebyshev interpolant of log2(1+f)/f, fmaf Horner.
 *   powf(x,y) = exp2f(y * log2f(x))   (the WGSL definition).
 *   log2 of a square root: log2_sqrt(x) = log2(x)/2.
 *   logf(x)   = log2f(x) * ln2,  expf(x) = exp2f(x * log2 e).
 *   sinf:     double-precision Cody-Waite reduction by pi/2 (three-part pi/2,
 *             exact for |x| < 2^27), Taylor polynomials, rounded to float.
 *             Used only by the clone-jitter hash (splat_adc.wgsl:70-72).
 *   cosf:     the same reduction, quadrant + 1 (Bezier spawn orientation).
 *
 * The product carries its own implementation of the same methods; a GPU test
 * sweeps both and requires identical bits.
 */
#ifndef NPS_REF_MATH_H
#define NPS_REF_MATH_H

#include <math.h>
#include <stdint.h>

#define RM_LN2F   0x1.62e430p-1f
#define RM_LOG2EF 0x1.715476p+0f

static inline float rm_exp2f(float x) {
    if (x != x) return x;
    if (x >= 128.0f) return INFINITY;
    if (x < -126.0f) return 0.0f;
    float k = rintf(x);
    float f = x - k; /* exact: |f| <= 0.5 */
    float p = 0x1.ffcbfcp-17f;         /* ln2^7/7! */
    p = fmaf(p, f, 0x1.430912p-13f);   /* ln2^6/6! */
    p = fmaf(p, f, 0x1.5d87fep-10f);   /* ln2^5/5! */
    p = fmaf(p, f, 0x1.3b2ab6p-7f);    /* ln2^4/4! */
    p = fmaf(p, f, 0x1.c6b08ep-5f);    /* ln2^3/3! */
    p = fmaf(p, f, 0x1.ebfbe0p-3f);    /* ln2^2/2! */
    p = fmaf(p, f, 0x1.62e430p-1f);    /* ln2      */
    p = fmaf(p, f, 1.0f);
    float r = ldexpf(p, (int)k);
    if (r < 0x1p-126f) return 0.0f;
    return r;
}

/* log2(m) = f * P(f), f = m - 1, m in [sqrt(1/2), sqrt(2)): P is the degree-9
 * Chebyshev interpolant of log2(1+f)/f on that interval (error < 6e-9
 * relative), coefficients rounded to float, Horner with fmaf. */
static inline float rm_log2f(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e;
    float m = frexpf(x, &e); /* m in [0.5, 1) */
    if (m < 0x1.6a09e6p-1f) { m = m * 2.0f; e = e - 1; }
    float f = m - 1.0f; /* exact (Sterbenz) */
    float p = -0x1.b84fb6p-4f;
    p = fmaf(p, f, 0x1.7a63a0p-3f);
    p = fmaf(p, f, -0x1.87f6a2p-3f);
    p = fmaf(p, f, 0x1.a38c68p-3f);
    p = fmaf(p, f, -0x1.eab7aap-3f);
    p = fmaf(p, f, 0x1.277a52p-2f);
    p = fmaf(p, f, -0x1.715a70p-2f);
    p = fmaf(p, f, 0x1.ec70aap-2f);
    p = fmaf(p, f, -0x1.715470p-1f);
    p = fmaf(p, f, 0x1.715476p+0f);
    return fmaf(f, p, (float)e);
}

static inline float rm_powf(float x, float y) { return rm_exp2f(y * rm_log2f(x)); }
/* log2(sqrt(x)) evaluated as log2(x)/2 (no intermediate rounding of the root) */
static inline float rm_log2_sqrtf(float x) { return 0.5f * rm_log2f(x); }
static inline float rm_logf(float x) { return rm_log2f(x) * RM_LN2F; }
static inline float rm_expf(float x) { return rm_exp2f(x * RM_LOG2EF); }

static inline float rm_sinf(float xf) {
    double x = (double)xf;
    if (!(x - x == 0.0)) return NAN; /* inf or nan */
    double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb5p+0;
    r = r - k * 0x1.110b46p-26;
    r = r - k * 0x1.1a62633145c07p-54;
    int64_t ki = (int64_t)k;
    int q = (int)(ki & 3);
    double r2 = r * r;
    double s = -0x1.ae7f3e733b81fp-41;
    s = s * r2 + 0x1.6124613a86d09p-33;
    s = s * r2 - 0x1.ae64567f544e4p-26;
    s = s * r2 + 0x1.71de3a556c734p-19;
    s = s * r2 - 0x1.a01a01a01a01ap-13;
    s = s * r2 + 0x1.1111111111111p-7;
    s = s * r2 - 0x1.5555555555555p-3;
    double sinr = r + (r * r2) * s;
    double c = 0x1.ae7f3e733b81fp-45;
    c = c * r2 - 0x1.93974a8c07c9dp-37;
    c = c * r2 + 0x1.1eed8eff8d898p-29;
    c = c * r2 - 0x1.27e4fb7789f5cp-22;
    c = c * r2 + 0x1.a01a01a01a01ap-16;
    c = c * r2 - 0x1.6c16c16c16c17p-10;
    c = c * r2 + 0x1.5555555555555p-5;
    c = c * r2 - 0x1.0p-1;
    double cosr = 1.0 + r2 * c;
    double out;
    switch (q) {
        case 0: out = sinr; break;
        case 1: out = cosr; break;
        case 2: out = -sinr; break;
        default: out = -cosr; break;
    }
    return (float)out;
}

/* cos with the same reduction and polynomials as rm_sinf, quadrant shifted by
 * one (cos x = sin(x + pi/2) in exact arithmetic).  Used only by the spawn
 * orientation of the Bezier densification (bezier_adc.wgsl:226-228). */
static inline float rm_cosf(float xf) {
    double x = (double)xf;
    if (!(x - x == 0.0)) return NAN;
    double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb5p+0;
    r = r - k * 0x1.110b46p-26;
    r = r - k * 0x1.1a62633145c07p-54;
    int64_t ki = (int64_t)k;
    int q = (int)((ki + 1) & 3);
    double r2 = r * r;
    double s = -0x1.ae7f3e733b81fp-41;
    s = s * r2 + 0x1.6124613a86d09p-33;
    s = s * r2 - 0x1.ae64567f544e4p-26;
    s = s * r2 + 0x1.71de3a556c734p-19;
    s = s * r2 - 0x1.a01a01a01a01ap-13;
    s = s * r2 + 0x1.1111111111111p-7;
    s = s * r2 - 0x1.5555555555555p-3;
    double sinr = r + (r * r2) * s;
    double c = 0x1.ae7f3e733b81fp-45;
    c = c * r2 - 0x1.93974a8c07c9dp-37;
    c = c * r2 + 0x1.1eed8eff8d898p-29;
    c = c * r2 - 0x1.27e4fb7789f5cp-22;
    c = c * r2 + 0x1.a01a01a01a01ap-16;
    c = c * r2 - 0x1.6c16c16c16c17p-10;
    c = c * r2 + 0x1.5555555555555p-5;
    c = c * r2 - 0x1.0p-1;
    double cosr = 1.0 + r2 * c;
    double out;
    switch (q) {
        case 0: out = sinr; break;
        case 1: out = cosr; break;
        case 2: out = -sinr; break;
        default: out = -cosr; break;
    }
    return (float)out;
}

/* atan of t in [0,1] in double: t > tan(pi/8) -> pi/4 + atan((t-1)/(t+1));
 * then atan(u) = 2 atan(u / (1 + sqrt(1 + u^2))) and the odd Taylor series of
 * atan to the u^21 term (|u| < tan(pi/16)).  Same double operations as the
 * kernels' definition, so the rounded float results agree bit for bit. */
static inline double rm_atan01(double t) {
    double base = 0.0, u = t;
    if (t > 0x1.a827999fcef32p-2) {
        base = 0x1.921fb54442d18p-1;
        u = (t - 1.0) / (t + 1.0);
    }
    double h = u / (1.0 + sqrt(1.0 + u * u));
    double h2 = h * h;
    static const double c[10] = { 1.0 / 3.0, -1.0 / 5.0, 1.0 / 7.0, -1.0 / 9.0, 1.0 / 11.0,
                                  -1.0 / 13.0, 1.0 / 15.0, -1.0 / 17.0, 1.0 / 19.0, -1.0 / 21.0 };
    /* q = c0 + h2 (c1 + h2 (... + h2 c9)) evaluated from the inside out */
    double q = c[9];
    for (int k = 8; k >= 0; --k) q = q * h2 + c[k];
    q = 1.0 - h2 * q;
    return base + 2.0 * (h * q);
}

static inline double rm_atan2d(double y, double x) {
    if (y != y || x != x) return NAN;
    double ax = fabs(x), ay = fabs(y);
    if (ax == 0.0 && ay == 0.0) return 0.0;
    int swap = ay > ax;
    double a = rm_atan01(swap ? ax / ay : ay / ax);
    if (swap) a = 0x1.921fb54442d18p+0 - a;
    if (x < 0.0) a = 0x1.921fb54442d18p+1 - a;
    return y < 0.0 ? -a : a;
}

static inline float rm_atan2f(float y, float x) { return (float)rm_atan2d((double)y, (double)x); }
/* asin z = atan2(z, sqrt((1 - z)(1 + z))), z in [-1, 1] */
static inline float rm_asinf(float zf) {
    double z = (double)zf;
    return (float)rm_atan2d(z, sqrt((1.0 - z) * (1.0 + z)));
}

/* WGSL fract(e) = e - floor(e) */
static inline float rm_fractf(float x) { return x - floorf(x); }

/* WGSL i32(f32): truncate toward zero, saturate; NaN -> 0 (defined here). */
static inline int32_t rm_f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

#endif
