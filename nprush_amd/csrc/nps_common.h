// nps_common.h -- shared device-side definitions of the gfx950 splat path.
//
// Geometry helpers restate the reference's WGSL with a fixed evaluation order
// (left-associative operators, no contraction; built with -ffp-contract=off):
//   M * v      = ((c0*v.x + c1*v.y) + c2*v.z) + c3*v.w   (column-major M)
//   dot(a, b)  = (a.x*b.x + a.y*b.y) + a.z*b.z
// so that every value a kernel produces equals the CPU restatement bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nps_math.h"

namespace nps {

constexpr int kTile = 16;        // reference workgroup tile, splat_backward.wgsl:136
constexpr int kTilePx = kTile * kTile;
constexpr int kBatch = 64;       // splat records staged in LDS per batch
constexpr int kChunk = 8;        // transmittance checkpoint stride (entries)
constexpr int kMaxCap = 1024;    // longest tile list the REF composite handles (the reference's cap)

// Per-view constants, passed by value to every kernel.
struct ViewParams {
    float vp[16];      // column-major view-projection (SplatUniforms.vp)
    float vp_inv[16];  // its inverse (SplatUniforms.vp_inv)
    uint32_t W, H;     // target dimensions
    float aspect;      // f32(W) / f32(H)
    uint32_t tiles_x, tiles_y, num_tiles;
    uint32_t tile;     // tile edge in pixels (16 in REF mode)
    uint32_t pad;
};

// Per-splat forward record (64 B): everything the per-pair forward needs.
struct alignas(16) FwdRec {
    float scr_x, scr_y;   // projected centre, aspect-corrected NDC
    float A, B, C;        // conic = inverse 2D covariance
    float op;             // opacity
    float r, g, b;        // colour
    float sa, sb;         // shape a, shape b
    float depth;          // clamp(1 - 0.1/max(w,0.1), 0, 1)
    float rcut2;          // r^2 beyond which power <= -15 is guaranteed (alpha is exactly 0)
    float pad[3];
};

// Per-splat backward record (128 B).
struct alignas(16) BwdRec {
    float m[6];     // m0x m0y m1x m1y m2x m2y   (projected axis * scale)
    float axs[6];   // ax_s.xy ay_s.xy az_s.xy   (projected unit axes)
    float s3[3];    // max(s, 1e-4) per axis
    float J0[3];    // d screen.x / d pos_j = aspect*(vp[j][0] w - clip.x vp[j][3]) / w^2
    float J1[3];    // d screen.y / d pos_j = (vp[j][1] w - clip.y vp[j][3]) / w^2
    float vpd[3];   // vp[j][3] * (0.1 / (w*w))
    float pos[3];   // splat centre
    float q[4];     // qw qx qy qz
    float w;        // clip.w
};

static_assert(sizeof(FwdRec) == 64, "FwdRec layout");
static_assert(sizeof(BwdRec) == 128, "BwdRec layout");

struct F4 {
    float x, y, z, w;
};
struct F3 {
    float x, y, z;
};
struct F2 {
    float x, y;
};

NPS_HD F4 mat_vec(const float* m, float x, float y, float z, float w) {
    F4 o;
    o.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
    o.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
    o.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
    o.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
    return o;
}

// splat_backward.wgsl:34-40, one coordinate at a time.
NPS_HD float px_to_x(uint32_t px, uint32_t W, float aspect) {
    float u = ((float)px + 0.5f) / (float)W;
    float x = u * 2.0f - 1.0f;
    return x * aspect;
}
NPS_HD float py_to_y(uint32_t py, uint32_t H) {
    float v = ((float)py + 0.5f) / (float)H;
    float y = v * 2.0f - 1.0f;
    return -y;
}

NPS_HD F3 cross3(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// splat_backward.wgsl:42-45
NPS_HD F3 quat_rotate(float qw, float qx, float qy, float qz, F3 v) {
    F3 u{qx, qy, qz};
    F3 c = cross3(u, v);
    F3 t{2.0f * c.x, 2.0f * c.y, 2.0f * c.z};
    F3 c2 = cross3(u, t);
    return F3{(v.x + qw * t.x) + c2.x, (v.y + qw * t.y) + c2.y, (v.z + qw * t.z) + c2.z};
}

// splat_backward.wgsl:52-58
NPS_HD F2 project_axis(const float* vp, F3 ax, float cx, float cy, float w, float aspect) {
    F4 ac = mat_vec(vp, ax.x, ax.y, ax.z, 0.0f);
    F2 o;
    o.x = (ac.x * w - cx * ac.w) / (w * w) * aspect;
    o.y = (ac.y * w - cy * ac.w) / (w * w);
    return o;
}

// r^2 of a pair, d = p - centre, sp = conic * d (splat_backward.wgsl:98,127-131).
struct PairR2 {
    float dx, dy, spx, spy, r2;
};
__device__ __forceinline__ PairR2 pair_r2(const FwdRec& f, float px, float py) {
    PairR2 q;
    q.dx = px - f.scr_x;
    q.dy = py - f.scr_y;
    q.spx = f.A * q.dx + f.B * q.dy;
    q.spy = f.B * q.dx + f.C * q.dy;
    q.r2 = q.dx * q.spx + q.dy * q.spy;
    return q;
}

// Saturating float -> int32 conversion in one instruction: gfx950's
// v_cvt_i32_f32 truncates toward zero, clamps out-of-range values and maps
// NaN to 0, which is exactly f2i_sat (checked against it on the GPU).
__device__ __forceinline__ int32_t f2i_hw(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t n) {
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous run of tiles so neighbouring tiles share its L2.
    uint32_t per = n / 8u, rem = n % 8u;
    uint32_t xcd = i % 8u, local = i / 8u;
    return xcd < rem ? xcd * (per + 1u) + local : rem * (per + 1u) + (xcd - rem) * per + local;
}

// First index in [0, n) whose predicate is false, for a predicate that is true
// on a prefix and false after it (n if always true).
template <class P>
__device__ __forceinline__ uint32_t first_false(uint32_t n, P pred) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (pred(mid)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// first_false(n, pred) started from a guess g: pred at g, then at up to two
// neighbours towards the answer, narrows [lo, hi] (pred true below lo, false at
// hi or hi = n); the binary search covers what is left.  Exact for any g; at
// most three evaluations when g is within one of the answer.
template <class P>
__device__ __forceinline__ uint32_t first_false_near(uint32_t n, uint32_t g, P pred) {
    uint32_t lo = 0, hi = n;
    g = min(g, n);
    if (g < n && pred(g)) {
        lo = g + 1;
        for (int k = 0; k < 2 && lo < hi; ++k) {
            if (pred(lo)) {
                ++lo;
            } else {
                hi = lo;
                break;
            }
        }
    } else {
        hi = g;
        for (int k = 0; k < 2 && hi > lo; ++k) {
            if (pred(hi - 1)) {
                lo = hi;
                break;
            }
            --hi;
        }
    }
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (pred(mid)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// A tile index guess from a position in tile units: floor(t) + off clamped to
// [0, n] (NaN and infinities clamp too).
__device__ __forceinline__ uint32_t tile_guess(float t, int off, uint32_t n) {
    const float f = fminf(fmaxf(floorf(t), -1.0f), (float)n) + (float)off;
    return f <= 0.0f ? 0u : (f >= (float)n ? n : (uint32_t)f);
}

// Tiles [tx0, tx_end) x [ty0, ty_end) whose pixel-centre rectangle the screen box
// [minx, maxx] x [miny, maxy] overlaps, with the float overlap test of
// splat_backward.wgsl:148-153: tile tx spans x in [px_to_x(T tx),
// px_to_x(min(T tx + T, W))], tile ty spans y in [py_to_y(min(T ty + T, H)),
// py_to_y(T ty)].  Each predicate is monotone in the tile index, so each bound
// is the first index where it fails; the search starts from that index
// estimated by inverting px_to_x / py_to_y.
__device__ __forceinline__ uint4 tile_rect(float minx, float maxx, float miny, float maxy, uint32_t W, uint32_t H,
                                           uint32_t T, uint32_t tiles_x, uint32_t tiles_y, float aspect) {
    const float fW = (float)W, fH = (float)H, fT = (float)T;
    const float gx0 = (((minx / aspect + 1.0f) * 0.5f) * fW - 0.5f) / fT;
    const float gx1 = (((maxx / aspect + 1.0f) * 0.5f) * fW - 0.5f) / fT;
    const float gy0 = (((1.0f - maxy) * 0.5f) * fH - 0.5f) / fT;
    const float gy1 = (((1.0f - miny) * 0.5f) * fH - 0.5f) / fT;
    const uint32_t tx0 = first_false_near(tiles_x, tile_guess(gx0, 0, tiles_x), [&](uint32_t tx) {
        return minx > px_to_x(min(tx * T + T, W), W, aspect);
    });
    const uint32_t tx_end = first_false_near(tiles_x, tile_guess(gx1, 1, tiles_x),
                                             [&](uint32_t tx) { return !(maxx < px_to_x(tx * T, W, aspect)); });
    const uint32_t ty0 = first_false_near(tiles_y, tile_guess(gy0, 0, tiles_y), [&](uint32_t ty) {
        return maxy < py_to_y(min(ty * T + T, H), H);
    });
    const uint32_t ty_end = first_false_near(tiles_y, tile_guess(gy1, 1, tiles_y),
                                             [&](uint32_t ty) { return !(miny > py_to_y(ty * T, H)); });
    return make_uint4(tx0, tx_end, ty0, ty_end);
}

// Adds the number of active lanes of the wave with `pred` true to *ctr_true and
// with it false to *ctr_false: one atomic per wave and counter instead of one
// per lane (same-address atomics from every lane of a launch serialise).
__device__ __forceinline__ void wave_count2(bool pred, uint32_t* ctr_true, uint32_t* ctr_false) {
    const uint64_t act = __ballot(1);
    const uint64_t yes = __ballot(pred);
    if ((int)(threadIdx.x & 63u) == __ffsll((unsigned long long)act) - 1) {
        const uint32_t t = (uint32_t)__popcll(yes), f = (uint32_t)__popcll(act) - t;
        if (t) atomicAdd(ctr_true, t);
        if (f) atomicAdd(ctr_false, f);
    }
}

// A 64-bit value that is the same in every lane, as a scalar.  readfirstlane
// returns int: each half is taken back to uint32_t before widening, so the low
// half does not sign-extend into the high one.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Minimum over y in [lo, hi] of a x^2 + 2 b x y + c y^2 (c > 0, a c >= b^2) for
// fixed x: the interior minimum (ac - b^2) x^2 / c when -b x / c lies in the
// interval, else the value at the nearer end.  ic ~ 1/c.
__device__ __forceinline__ float conic_edge_min(float a, float b, float c, float det, float ic, float x, float lo,
                                                float hi) {
    const float t = -b * x;
    const bool below = t < c * lo, above = t > c * hi;
    const float y = below ? lo : hi;
    const float fe = ((a * x) * x + ((2.0f * b) * x) * y) + (c * y) * y;
    const float fi = ((det * x) * x) * ic;
    return (below | above) ? fe : fi;
}

// Whether the conic's level set r^2 <= rc reaches the rectangle [x0,x1]x[y0,y1]
// (offsets from the centre): the exact minimum of r^2 over the rectangle (0 when
// it holds the centre, else the least edge minimum) against rc.  The margin
// 1e-5 (A X^2 + 2|B| X Y + C Y^2), X, Y the largest offsets, exceeds by far the
// rounding of this minimum and of the per-pixel r^2 (a few FP32 ulps of the
// same terms), so false means every pixel centre of the rectangle has
// r^2 > rc as the pixel test computes it.
__device__ __forceinline__ bool conic_reaches_rect(float A, float B, float C, float x0, float x1, float y0, float y1,
                                                   float rc) {
    const float det = A * C - B * B;
    const float iA = __builtin_amdgcn_rcpf(A), iC = __builtin_amdgcn_rcpf(C);
    const float m = fminf(fminf(conic_edge_min(A, B, C, det, iC, x0, y0, y1), conic_edge_min(A, B, C, det, iC, x1, y0, y1)),
                          fminf(conic_edge_min(C, B, A, det, iA, y0, x0, x1), conic_edge_min(C, B, A, det, iA, y1, x0, x1)));
    const bool inside = (x0 <= 0.0f) & (x1 >= 0.0f) & (y0 <= 0.0f) & (y1 >= 0.0f);
    const float X = fmaxf(fabsf(x0), fabsf(x1)), Y = fmaxf(fabsf(y0), fabsf(y1));
    const float S = ((A * X) * X + ((2.0f * fabsf(B)) * X) * Y) + (C * Y) * Y;
    return inside | !(m > rc + 1e-5f * S);
}

}  // namespace nps
