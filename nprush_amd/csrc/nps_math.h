// nps_math.h -- deterministic FP32 transcendentals for the gfx950 kernels.
//
// WGSL leaves exp/pow/log/sin implementation-defined within ULP bounds
// (pow is defined as exp2(y * log2(x))).  The kernels need a definition that
// is identical on every device and on a CPU, so these are built only from
// correctly rounded IEEE-754 +,-,*,/ and explicit fmaf (hipcc's default for
// gfx950: division and sqrt are correctly rounded, fp32 denormals kept) and
// compiled with -ffp-contract=off:
//   exp2: Cody-Waite split x = k + f (|f| <= 1/2), degree-7 Taylor series of
//         e^(f ln 2), fmaf Horner, exact scaling by 2^k; results below
//         2^-126 flush to zero.
//   log2: x = m 2^e, m in [sqrt(1/2), sqrt(2)); log2 m = f P(f), f = m - 1,
//         P = degree-9 Chebyshev interpolant of log2(1+f)/f, fmaf Horner.
//   log2_sqrt(x) = log2(x)/2 (log2 of a square root without rounding the root).
//   pow(x,y) = exp2(y log2 x); log(x) = log2(x) ln 2; exp(x) = exp2(x log2 e).
//   sin: double-precision Cody-Waite reduction by pi/2 (three-part constant,
//        exact for |x| < 2^27) and Taylor polynomials, rounded to float.
//   atan2, asin: double precision (octant reduction, argument halving, odd
//        Taylor series), rounded to float.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define NPS_HD __host__ __device__ __forceinline__

namespace nps {

constexpr float kLn2 = 0x1.62e430p-1f;
constexpr float kLog2e = 0x1.715476p+0f;

NPS_HD float exp2_det(float x) {
    if (x != x) return x;
    if (x >= 128.0f) return INFINITY;
    if (x < -126.0f) return 0.0f;
    float k = rintf(x);
    float f = x - k;
    float p = 0x1.ffcbfcp-17f;
    p = fmaf(p, f, 0x1.430912p-13f);
    p = fmaf(p, f, 0x1.5d87fep-10f);
    p = fmaf(p, f, 0x1.3b2ab6p-7f);
    p = fmaf(p, f, 0x1.c6b08ep-5f);
    p = fmaf(p, f, 0x1.ebfbe0p-3f);
    p = fmaf(p, f, 0x1.62e430p-1f);
    p = fmaf(p, f, 1.0f);
    float r = ldexpf(p, (int)k);
    return r < 0x1p-126f ? 0.0f : r;
}

// exp2_det for x in [-125, 127], where none of its special cases applies
// (no NaN, no overflow, results >= 2^-125.5, never flushed): the same
// operations, so the same bits.  Callers guarantee the range.
NPS_HD float exp2_det_in_range(float x) {
    float k = rintf(x);
    float f = x - k;
    float p = 0x1.ffcbfcp-17f;
    p = fmaf(p, f, 0x1.430912p-13f);
    p = fmaf(p, f, 0x1.5d87fep-10f);
    p = fmaf(p, f, 0x1.3b2ab6p-7f);
    p = fmaf(p, f, 0x1.c6b08ep-5f);
    p = fmaf(p, f, 0x1.ebfbe0p-3f);
    p = fmaf(p, f, 0x1.62e430p-1f);
    p = fmaf(p, f, 1.0f);
    return ldexpf(p, (int)k);
}

NPS_HD float log2_det(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e;
    float m = frexpf(x, &e);
    if (m < 0x1.6a09e6p-1f) {
        m = m * 2.0f;
        e = e - 1;
    }
    float f = m - 1.0f;
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

// log2(sqrt(x)) as log2(x)/2: no intermediate rounding of the root.
NPS_HD float log2_sqrt_det(float x) { return 0.5f * log2_det(x); }

NPS_HD float pow_det(float x, float y) { return exp2_det(y * log2_det(x)); }
NPS_HD float log_det(float x) { return log2_det(x) * kLn2; }
NPS_HD float exp_det(float x) { return exp2_det(x * kLog2e); }

NPS_HD float sin_det(float xf) {
    double x = (double)xf;
    if (!(x - x == 0.0)) return NAN;
    double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb5p+0;
    r = r - k * 0x1.110b46p-26;
    r = r - k * 0x1.1a62633145c07p-54;
    int q = (int)(((int64_t)k) & 3);
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
    double out = (q == 0) ? sinr : (q == 1) ? cosr : (q == 2) ? -sinr : -cosr;
    return (float)out;
}

// cos with the reduction and polynomials of sin_det, quadrant shifted by one.
NPS_HD float cos_det(float xf) {
    double x = (double)xf;
    if (!(x - x == 0.0)) return NAN;
    double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = x - k * 0x1.921fb5p+0;
    r = r - k * 0x1.110b46p-26;
    r = r - k * 0x1.1a62633145c07p-54;
    int q = (int)((((int64_t)k) + 1) & 3);
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
    double out = (q == 0) ? sinr : (q == 1) ? cosr : (q == 2) ? -sinr : -cosr;
    return (float)out;
}

// atan of t in [0, 1] in double: above tan(pi/8), atan t = pi/4 + atan((t-1)/(t+1));
// then atan u = 2 atan(u / (1 + sqrt(1 + u^2))) brings |u| below tan(pi/16) < 0.2,
// where the odd Taylor series to the u^21 term leaves an error below 1e-17.
NPS_HD double atan01_det(double t) {
    double base = 0.0, u = t;
    if (t > 0x1.a827999fcef32p-2) {  // tan(pi/8) = sqrt(2) - 1
        base = 0x1.921fb54442d18p-1;  // pi/4
        u = (t - 1.0) / (t + 1.0);
    }
    const double v = u / (1.0 + sqrt(1.0 + u * u));
    const double v2 = v * v;
    double p = -1.0 / 21.0;
    p = p * v2 + 1.0 / 19.0;
    p = p * v2 - 1.0 / 17.0;
    p = p * v2 + 1.0 / 15.0;
    p = p * v2 - 1.0 / 13.0;
    p = p * v2 + 1.0 / 11.0;
    p = p * v2 - 1.0 / 9.0;
    p = p * v2 + 1.0 / 7.0;
    p = p * v2 - 1.0 / 5.0;
    p = p * v2 + 1.0 / 3.0;
    p = 1.0 - v2 * p;
    return base + 2.0 * (v * p);
}

// atan2 in double by octant: t = min(|x|,|y|) / max(|x|,|y|), then the quadrant.
// (0, 0) gives 0; a NaN operand gives NaN.
NPS_HD double atan2_det_d(double y, double x) {
    if (y != y || x != x) return NAN;
    const double ax = fabs(x), ay = fabs(y);
    if (ax == 0.0 && ay == 0.0) return 0.0;
    const bool swap = ay > ax;
    double a = atan01_det(swap ? ax / ay : ay / ax);
    if (swap) a = 0x1.921fb54442d18p+0 - a;  // pi/2
    if (x < 0.0) a = 0x1.921fb54442d18p+1 - a;  // pi
    return y < 0.0 ? -a : a;
}

// FP32 atan2 and asin: the double-precision value rounded to float.
// asin z = atan2(z, sqrt((1 - z)(1 + z))) for z in [-1, 1].
NPS_HD float atan2_det(float y, float x) { return (float)atan2_det_d((double)y, (double)x); }
NPS_HD float asin_det(float zf) {
    const double z = (double)zf;
    return (float)atan2_det_d(z, sqrt((1.0 - z) * (1.0 + z)));
}

NPS_HD float fract_det(float x) { return x - floorf(x); }

// WGSL i32(f32): truncation toward zero, saturating; NaN -> 0.
NPS_HD int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

// WGSL max/min (IEEE maxNum: a NaN operand loses) and clamp = min(max(e, lo), hi).
// Written as selects (no branches): a NaN `a` falls through the compare to b,
// a NaN `b` alone gives a, two NaNs give b.
NPS_HD float fmax_w(float a, float b) {
    const float r = a > b ? a : b;
    return ((b != b) & (a == a)) ? a : r;
}
NPS_HD float fmin_w(float a, float b) {
    const float r = a < b ? a : b;
    return ((b != b) & (a == a)) ? a : r;
}
NPS_HD float clamp_w(float e, float lo, float hi) { return fmin_w(fmax_w(e, lo), hi); }

}  // namespace nps
