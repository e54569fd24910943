// nps_ext.h -- per-splat projection of the EXT path (3D Gaussian splats with
// spherical-harmonic colour), shared by the EXT preprocess and the per-splat
// chain-rule kernel.
//
// EXT splat record (16 fp32) + SH block (48 fp32):
//   [x y z opacity_logit] [log_sx log_sy log_sz 0] [qw qx qy qz] [0 0 0 0]
//   sh[3k + c] = coefficient k (real SH basis, degree <= 3) of channel c
// Activations: scale exp(log_s), opacity sigmoid(logit), rotation by the
// normalised quaternion (the polynomial form of quat_rotate,
// splat_backward.wgsl:42-45).  Projection of the scaled axes as the reference's
// project_axis (splat_backward.wgsl:52-58), covariance M M^T + (0.3 px)^2 I and
// its conic (splat_backward.wgsl:109-126); cull w < 0.1 or det <= 1e-30;
// 3-sigma extents per axis; colour max(SH(d) + 0.5, 0) with d the unit vector
// from the camera centre to the mean.  Every float expression is written in
// the order the FP32 CPU oracle evaluates it (compiled with -ffp-contract=off),
// so projected records, colours and composited images are bit-identical.
#pragma once
#include "nps_common.h"

namespace nps {

constexpr float kShC0 = (float)0.28209479177387814;
constexpr float kShC1 = (float)0.4886025119029199;
constexpr float kShC2_0 = (float)1.0925484305920792;
constexpr float kShC2_1 = (float)-1.0925484305920792;
constexpr float kShC2_2 = (float)0.31539156525252005;
constexpr float kShC2_3 = (float)-1.0925484305920792;
constexpr float kShC2_4 = (float)0.5462742152960396;
constexpr float kShC3_0 = (float)-0.5900435899266435;
constexpr float kShC3_1 = (float)2.890611442640554;
constexpr float kShC3_2 = (float)-0.4570457994644658;
constexpr float kShC3_3 = (float)0.3731763325901154;
constexpr float kShC3_4 = (float)-0.4570457994644658;
constexpr float kShC3_5 = (float)1.445305721320277;
constexpr float kShC3_6 = (float)-0.5900435899266435;

constexpr float kExtMinAlpha = (float)(1.0 / 255.0);  // pairs below are skipped
constexpr float kExtMaxAlpha = 0.99f;
constexpr float kExtTMin = 0.0001f;                    // early termination threshold
constexpr float kExtBg = 0.05f;

// The pixel test of the 3-sigma box, on the same dx, dy as r^2.
__device__ __forceinline__ bool in_box(float dx, float dy, float ex, float ey) {
    return fabsf(dx) <= ex && fabsf(dy) <= ey;
}

// Per-splat record of the EXT composite (48 B).
struct alignas(16) ExtRec {
    float scr_x, scr_y;  // projected centre (aspect-corrected NDC)
    float A, B, C;       // conic
    float op;            // sigmoid(logit)
    float r, g, b;       // max(SH + 0.5, 0)
    float rcut2;         // r^2 beyond which alpha < 1/255 is guaranteed (-inf: never drawn)
    float ext_x, ext_y;  // 3-sigma box half-extents: a pixel sees the splat only inside it
};

// Binning box = 3-sigma box * kExtBinScale + kExtBinPad: wide enough that a
// tile list holds every splat whose exact box contains one of the tile's pixel
// centres (the pixel test then makes images independent of the tile size).
constexpr float kExtBinScale = 1.00001f;
constexpr float kExtBinPad = 1e-6f;
static_assert(sizeof(ExtRec) == 48, "ExtRec layout");

// Gaussian falloff exponent: exp(-r^2/2) = exp2(r^2 * kExtNegHalfLog2e)
// (-0.5 log2 e; -0.5 * log2e is exact in FP32).
constexpr float kExtNegHalfLog2e = -0.5f * kLog2e;

// EXT pair of the pixel centre (pxf, pyf) and one record, given as its three
// float4 pieces p0 = (scr_x, scr_y, A, B), p1 = (C, op, r, g),
// p2 = (b, rcut2, ext_x, ext_y):
//   dx, dy = pixel - centre;  r^2 = fma(dx, fma(A, dx, B dy), dy fma(B, dx, C dy));
//   G = exp2(r^2 * kExtNegHalfLog2e);  alpha = min(op G, 0.99).
// Returns whether the pixel takes the pair: `use`, the centre of the pixel in the
// 3-sigma box, r^2 <= rcut2 (beyond it alpha < 1/255 is guaranteed) and
// alpha >= 1/255.  a and G are set to alpha and G; they are meaningful only
// where the pair is taken (r^2 <= rcut2 < 12 keeps exp2's argument in range
// there).  Branch-free.
__device__ __forceinline__ bool ext_pair_eval(float pxf, float pyf, const float4& p0, const float4& p1,
                                              const float4& p2, bool use, float& a, float& G) {
    const float dx = pxf - p0.x, dy = pyf - p0.y;
    const float spx = fmaf(p0.z, dx, p0.w * dy);
    const float spy = fmaf(p0.w, dx, p1.x * dy);
    const float r2 = fmaf(dx, spx, dy * spy);
    G = exp2_det_in_range(r2 * kExtNegHalfLog2e);
    a = fminf(p1.y * G, kExtMaxAlpha);  // op G is never NaN where the pair is taken
    // non-short-circuit &: every operand is evaluated, no branch
    return (int)use & (int)(fabsf(dx) <= p2.z) & (int)(fabsf(dy) <= p2.w) & (int)!(r2 > p2.y) &
           (int)!(a < kExtMinAlpha);
}

// Wave ballot of a predicate (no round trip through a 0/1 integer).
__device__ __forceinline__ uint64_t ballot_b(bool p) { return __builtin_amdgcn_ballot_w64(p); }

struct ExtProj {
    bool valid;
    float s[3], qn[4], qlen, op;
    float cx, cy, w;
    float scr[2];
    float axw[3][3];
    float axs[3][2];
    float m[3][2];
    float cov[3];
    float det, A, B, C;
    float ext[2];
    float dir[3], dlen;
    float sh_raw[3];
    float rgb[3];
};

// Real SH basis Y_k(d), k < (deg+1)^2.
NPS_HD void sh_basis(float x, float y, float z, int deg, float* Y) {
    Y[0] = kShC0;
    if (deg < 1) return;
    Y[1] = -kShC1 * y;
    Y[2] = kShC1 * z;
    Y[3] = -kShC1 * x;
    if (deg < 2) return;
    float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = kShC2_0 * xy;
    Y[5] = kShC2_1 * yz;
    Y[6] = kShC2_2 * (2.0f * zz - xx - yy);
    Y[7] = kShC2_3 * xz;
    Y[8] = kShC2_4 * (xx - yy);
    if (deg < 3) return;
    Y[9] = kShC3_0 * y * (3.0f * xx - yy);
    Y[10] = kShC3_1 * xy * z;
    Y[11] = kShC3_2 * y * (4.0f * zz - xx - yy);
    Y[12] = kShC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    Y[13] = kShC3_4 * x * (4.0f * zz - xx - yy);
    Y[14] = kShC3_5 * z * (xx - yy);
    Y[15] = kShC3_6 * x * (xx - 3.0f * yy);
}

// g += sum_k w[k] dY_k/d(x, y, z), k < (deg+1)^2: the derivatives of the
// sh_basis polynomials (each evaluated in FP32), the terms that are not
// identically zero, accumulated in FP64.
NPS_HD void sh_grad_weighted(float x, float y, float z, int deg, const double* w, double* g) {
    if (deg < 1) return;
    g[1] += w[1] * (double)(-kShC1);
    g[2] += w[2] * (double)kShC1;
    g[0] += w[3] * (double)(-kShC1);
    if (deg < 2) return;
    g[0] += w[4] * (double)(kShC2_0 * y); g[1] += w[4] * (double)(kShC2_0 * x);
    g[1] += w[5] * (double)(kShC2_1 * z); g[2] += w[5] * (double)(kShC2_1 * y);
    g[0] += w[6] * (double)(kShC2_2 * (-2.0f * x)); g[1] += w[6] * (double)(kShC2_2 * (-2.0f * y));
    g[2] += w[6] * (double)(kShC2_2 * (4.0f * z));
    g[0] += w[7] * (double)(kShC2_3 * z); g[2] += w[7] * (double)(kShC2_3 * x);
    g[0] += w[8] * (double)(kShC2_4 * (2.0f * x)); g[1] += w[8] * (double)(kShC2_4 * (-2.0f * y));
    if (deg < 3) return;
    const float xx = x * x, yy = y * y, zz = z * z;
    g[0] += w[9] * (double)(kShC3_0 * (6.0f * x * y)); g[1] += w[9] * (double)(kShC3_0 * (3.0f * xx - 3.0f * yy));
    g[0] += w[10] * (double)(kShC3_1 * (y * z)); g[1] += w[10] * (double)(kShC3_1 * (x * z));
    g[2] += w[10] * (double)(kShC3_1 * (x * y));
    g[0] += w[11] * (double)(kShC3_2 * (-2.0f * x * y)); g[1] += w[11] * (double)(kShC3_2 * (4.0f * zz - xx - 3.0f * yy));
    g[2] += w[11] * (double)(kShC3_2 * (8.0f * y * z));
    g[0] += w[12] * (double)(kShC3_3 * (-6.0f * x * z)); g[1] += w[12] * (double)(kShC3_3 * (-6.0f * y * z));
    g[2] += w[12] * (double)(kShC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy));
    g[0] += w[13] * (double)(kShC3_4 * (4.0f * zz - 3.0f * xx - yy)); g[1] += w[13] * (double)(kShC3_4 * (-2.0f * x * y));
    g[2] += w[13] * (double)(kShC3_4 * (8.0f * x * z));
    g[0] += w[14] * (double)(kShC3_5 * (2.0f * x * z)); g[1] += w[14] * (double)(kShC3_5 * (-2.0f * y * z));
    g[2] += w[14] * (double)(kShC3_5 * (xx - yy));
    g[0] += w[15] * (double)(kShC3_6 * (3.0f * xx - 3.0f * yy)); g[1] += w[15] * (double)(kShC3_6 * (-6.0f * x * y));
}

// Projection of splat s (16 floats) with SH block sh (48 floats).  cam is the
// camera centre.  On return o.valid says whether the splat is drawn.  With
// sh == nullptr the colour (sh_raw, rgb) is left unset.
NPS_HD void ext_project(const float* s, const float* sh, int deg, const float* vp, const float* cam, uint32_t W,
                        uint32_t H, ExtProj& o) {
    o.valid = false;
    const float aspect = (float)W / (float)H;
    o.op = 1.0f / (1.0f + exp_det(-s[3]));
    for (int k = 0; k < 3; ++k) o.s[k] = exp_det(s[4 + k]);
    float qq = ((s[8] * s[8] + s[9] * s[9]) + s[10] * s[10]) + s[11] * s[11];
    o.qlen = sqrtf(qq > 1e-24f ? qq : 1e-24f);
    for (int k = 0; k < 4; ++k) o.qn[k] = s[8 + k] / o.qlen;
    o.cx = ((vp[0] * s[0] + vp[4] * s[1]) + vp[8] * s[2]) + vp[12];
    o.cy = ((vp[1] * s[0] + vp[5] * s[1]) + vp[9] * s[2]) + vp[13];
    o.w = ((vp[3] * s[0] + vp[7] * s[1]) + vp[11] * s[2]) + vp[15];
    if (!(o.w >= 0.1f)) return;
    o.scr[0] = o.cx / o.w * aspect;
    o.scr[1] = o.cy / o.w;
    const float qw = o.qn[0], qx = o.qn[1], qy = o.qn[2], qz = o.qn[3];
    o.axw[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz);
    o.axw[0][1] = 2.0f * (qx * qy + qw * qz);
    o.axw[0][2] = 2.0f * (qx * qz - qw * qy);
    o.axw[1][0] = 2.0f * (qx * qy - qw * qz);
    o.axw[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz);
    o.axw[1][2] = 2.0f * (qy * qz + qw * qx);
    o.axw[2][0] = 2.0f * (qx * qz + qw * qy);
    o.axw[2][1] = 2.0f * (qy * qz - qw * qx);
    o.axw[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
    const float w2 = o.w * o.w;
    for (int k = 0; k < 3; ++k) {
        const float* a = o.axw[k];
        float acx = (vp[0] * a[0] + vp[4] * a[1]) + vp[8] * a[2];
        float acy = (vp[1] * a[0] + vp[5] * a[1]) + vp[9] * a[2];
        float acw = (vp[3] * a[0] + vp[7] * a[1]) + vp[11] * a[2];
        o.axs[k][0] = (acx * o.w - o.cx * acw) / w2 * aspect;
        o.axs[k][1] = (acy * o.w - o.cy * acw) / w2;
        o.m[k][0] = o.axs[k][0] * o.s[k];
        o.m[k][1] = o.axs[k][1] * o.s[k];
    }
    const float f = 0.3f * (2.0f / (float)H);
    const float f2 = f * f;
    o.cov[0] = ((o.m[0][0] * o.m[0][0] + o.m[1][0] * o.m[1][0]) + o.m[2][0] * o.m[2][0]) + f2;
    o.cov[1] = (o.m[0][0] * o.m[0][1] + o.m[1][0] * o.m[1][1]) + o.m[2][0] * o.m[2][1];
    o.cov[2] = ((o.m[0][1] * o.m[0][1] + o.m[1][1] * o.m[1][1]) + o.m[2][1] * o.m[2][1]) + f2;
    o.det = o.cov[0] * o.cov[2] - o.cov[1] * o.cov[1];
    if (!(o.det > 1e-30f)) return;
    o.A = o.cov[2] / o.det;
    o.B = -o.cov[1] / o.det;
    o.C = o.cov[0] / o.det;
    o.ext[0] = 3.0f * sqrtf(o.cov[0]);
    o.ext[1] = 3.0f * sqrtf(o.cov[2]);
    float v0 = s[0] - cam[0], v1 = s[1] - cam[1], v2 = s[2] - cam[2];
    o.dlen = sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
    o.dir[0] = v0 / o.dlen;
    o.dir[1] = v1 / o.dlen;
    o.dir[2] = v2 / o.dlen;
    o.valid = true;
    if (!sh) return;
    float Y[16];
    sh_basis(o.dir[0], o.dir[1], o.dir[2], deg, Y);
    const int nk = (deg + 1) * (deg + 1);
    for (int c = 0; c < 3; ++c) {
        float acc = 0.0f;
        for (int k = 0; k < nk; ++k) acc = acc + Y[k] * sh[3 * k + c];
        o.sh_raw[c] = acc + 0.5f;
        o.rgb[c] = o.sh_raw[c] > 0.0f ? o.sh_raw[c] : 0.0f;
    }
}

// r^2 of a pair, d = p - centre, sp = conic * d.
struct ExtPair {
    float dx, dy, spx, spy, r2;
};
__device__ __forceinline__ ExtPair ext_pair(const ExtRec& f, float px, float py) {
    ExtPair q;
    q.dx = px - f.scr_x;
    q.dy = py - f.scr_y;
    q.spx = f.A * q.dx + f.B * q.dy;
    q.spy = f.B * q.dx + f.C * q.dy;
    q.r2 = q.dx * q.spx + q.dy * q.spy;
    return q;
}

}  // namespace nps
