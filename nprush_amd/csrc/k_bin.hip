// k_bin.hip -- per-view preprocess, tile-key generation and tile ranges.
//
// Replaces the binning block of splat_backward.wgsl:141-211.  The reference
// tests every splat against every 16x16 tile with an LDS bitmask and keeps the
// first 1024 ascending ids.  Here each splat is projected once:
//   k_preprocess : cull (opacity < 0.005, w < 0), the reference's square
//                  bounding box (R = pow(15/max(b,1e-4), 1/a),
//                  max_r = R*sqrt(max(cov00,cov11,1e-9))), the exact range of
//                  tiles whose rectangle passes the reference's float overlap
//                  test (binary search with the identical float predicate), and
//                  the forward/backward records of the splat;
//   k_emit_keys  : one 64-bit key (tile << 32 | splat id) per (splat, tile),
//                  written in splat order so a stable sort on the tile bits
//                  leaves every tile list in ascending id order;
//   k_tile_ranges: [start, end) of each tile in the sorted keys.
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

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

__global__ __launch_bounds__(256) void k_preprocess(const float4* __restrict__ splats, uint32_t n, ViewParams v,
                                                    FwdRec* __restrict__ fwd, BwdRec* __restrict__ bwd,
                                                    uint32_t* __restrict__ counts, uint4* __restrict__ rects) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 s0 = splats[4 * i + 0], s1 = splats[4 * i + 1], s2 = splats[4 * i + 2], s3 = splats[4 * i + 3];
    counts[i] = 0u;
    // s0 = x y z sx ; s1 = r g b op ; s2 = qw qx qy qz ; s3 = sy a b sz
    if (s1.w < 0.005f) return;
    F4 clip = mat_vec(v.vp, s0.x, s0.y, s0.z, 1.0f);
    float aspect = v.aspect;
    float scr_x = clip.x / clip.w * aspect;
    float scr_y = clip.y / clip.w;
    float w = clip.w;
    if (w < 0.0f) return;
    float sx = fmax_w(s0.w, 0.0001f), sy = fmax_w(s3.x, 0.0001f), sz = fmax_w(s3.w, 0.0001f);
    float safe_sb = fmax_w(s3.z, 0.0001f);
    float Rr = pow_det(15.0f / safe_sb, 1.0f / s3.y);
    F3 ax_w = quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{1.0f, 0.0f, 0.0f});
    F3 ay_w = quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{0.0f, 1.0f, 0.0f});
    F3 az_w = quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{0.0f, 0.0f, 1.0f});
    F2 ax_s = project_axis(v.vp, ax_w, clip.x, clip.y, w, aspect);
    F2 ay_s = project_axis(v.vp, ay_w, clip.x, clip.y, w, aspect);
    F2 az_s = project_axis(v.vp, az_w, clip.x, clip.y, w, aspect);
    float m0x = ax_s.x * sx, m0y = ax_s.y * sx;
    float m1x = ay_s.x * sy, m1y = ay_s.y * sy;
    float m2x = az_s.x * sz, m2y = az_s.y * sz;
    float cov00 = (m0x * m0x + m1x * m1x) + m2x * m2x;
    float cov01 = (m0x * m0y + m1x * m1y) + m2x * m2y;
    float cov11 = (m0y * m0y + m1y * m1y) + m2y * m2y;
    float filter_std = 0.3f * (2.0f / (float)v.H);
    float filter2 = filter_std * filter_std;
    cov00 = cov00 + filter2;
    cov11 = cov11 + filter2;
    float max_r = Rr * sqrtf(fmax_w(fmax_w(cov00, cov11), 1e-9f));
    float sminx = scr_x - max_r, sminy = scr_y - max_r;
    float smaxx = scr_x + max_r, smaxy = scr_y + max_r;

    const uint32_t W = v.W, H = v.H;
    const uint32_t T = (uint32_t)kTile;
    // tile tx spans x in [px_to_x(16 tx), px_to_x(min(16 tx + 16, W))] (BW:148-153)
    uint32_t tx0 = first_false(v.tiles_x, [&](uint32_t tx) {
        uint32_t x1 = min(tx * T + T, W);
        return sminx > px_to_x(x1, W, aspect);
    });
    uint32_t tx_end = first_false(v.tiles_x, [&](uint32_t tx) { return !(smaxx < px_to_x(tx * T, W, aspect)); });
    // tile ty spans y in [py_to_y(min(16 ty + 16, H)), py_to_y(16 ty)]
    uint32_t ty0 = first_false(v.tiles_y, [&](uint32_t ty) {
        uint32_t y1 = min(ty * T + T, H);
        return smaxy < py_to_y(y1, H);
    });
    uint32_t ty_end = first_false(v.tiles_y, [&](uint32_t ty) { return !(sminy > py_to_y(ty * T, H)); });
    if (tx0 >= tx_end || ty0 >= ty_end) return;
    counts[i] = (tx_end - tx0) * (ty_end - ty0);
    rects[i] = make_uint4(tx0, tx_end, ty0, ty_end);

    // records (splat_backward.wgsl:92-134 hoisted out of the per-pair loop)
    float det = cov00 * cov11 - cov01 * cov01;
    float inv_det = (fabsf(det) < 1e-10f) ? 0.0f : 1.0f / det;
    FwdRec f;
    f.scr_x = scr_x;
    f.scr_y = scr_y;
    f.A = cov11 * inv_det;
    f.B = -cov01 * inv_det;
    f.C = cov00 * inv_det;
    f.op = s1.w;
    f.r = s1.x;
    f.g = s1.y;
    f.b = s1.z;
    f.sa = s3.y;
    f.sb = s3.z;
    float lin = fmax_w(w, 0.1f);
    f.depth = clamp_w(1.0f - 0.1f / lin, 0.0f, 1.0f);
    // Pairs with r^2 > rcut2 have power = -b*pow(r, a) <= -15, so alpha is exactly
    // 0 and the pair can skip pow/exp.  rcut = 1.001 * R with R = (15/b)^(1/a)
    // as computed above; for a in [0.1, 64] and b >= 1e-4 the 0.1 % margin
    // (>= 1e-4 relative in b*r^a) exceeds the combined error of the two pow
    // evaluations (< 2e-5 relative), so the float test power > -15 is false.
    float rcut2 = INFINITY;
    if (s3.y >= 0.1f && s3.y <= 64.0f && s3.z >= 0.0001f && s3.z <= 1e30f && Rr > 0.0f && Rr <= 1e15f) {
        float rc = Rr * 1.001f;
        rcut2 = fmax_w(rc * rc, 0.0001f);
    }
    f.rcut2 = rcut2;
    f.pad[0] = f.pad[1] = f.pad[2] = 0.0f;
    fwd[i] = f;

    BwdRec b;
    b.m[0] = m0x; b.m[1] = m0y; b.m[2] = m1x; b.m[3] = m1y; b.m[4] = m2x; b.m[5] = m2y;
    b.axs[0] = ax_s.x; b.axs[1] = ax_s.y; b.axs[2] = ay_s.x; b.axs[3] = ay_s.y; b.axs[4] = az_s.x; b.axs[5] = az_s.y;
    b.s3[0] = sx; b.s3[1] = sy; b.s3[2] = sz;
    float w2 = w * w;
    float dddw = 0.1f / (w * w);
    for (int j = 0; j < 3; ++j) {
        float vp0 = v.vp[4 * j + 0], vp1 = v.vp[4 * j + 1], vp3 = v.vp[4 * j + 3];
        b.J0[j] = aspect * (vp0 * w - clip.x * vp3) / w2;
        b.J1[j] = (vp1 * w - clip.y * vp3) / w2;
        b.vpd[j] = vp3 * dddw;
    }
    b.pos[0] = s0.x; b.pos[1] = s0.y; b.pos[2] = s0.z;
    b.q[0] = s2.x; b.q[1] = s2.y; b.q[2] = s2.z; b.q[3] = s2.w;
    b.w = w;
    bwd[i] = b;
}

__global__ __launch_bounds__(256) void k_emit_keys(uint32_t n, const uint32_t* __restrict__ counts,
                                                   const uint64_t* __restrict__ offsets,
                                                   const uint4* __restrict__ rects, uint32_t tiles_x,
                                                   uint64_t* __restrict__ keys) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (counts[i] == 0u) return;
    uint4 r = rects[i];
    uint64_t o = offsets[i];
    for (uint32_t ty = r.z; ty < r.w; ++ty)
        for (uint32_t tx = r.x; tx < r.y; ++tx) {
            uint64_t tile = (uint64_t)(ty * tiles_x + tx);
            keys[o++] = (tile << 32) | (uint64_t)i;
        }
}

// ranges[2t] = start, ranges[2t+1] = end of tile t in the sorted keys.
// Tiles without keys keep the [0,0) the caller cleared.
__global__ __launch_bounds__(256) void k_tile_ranges(const uint64_t* __restrict__ keys, uint64_t m,
                                                     uint32_t* __restrict__ ranges) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    uint32_t t = (uint32_t)(keys[j] >> 32);
    if (j == 0 || (uint32_t)(keys[j - 1] >> 32) != t) ranges[2 * t] = (uint32_t)j;
    if (j + 1 == m || (uint32_t)(keys[j + 1] >> 32) != t) ranges[2 * t + 1] = (uint32_t)(j + 1);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
}

void launch_preprocess(hipStream_t st, const float* splats, uint32_t n, const ViewParams& v, FwdRec* fwd,
                       BwdRec* bwd, uint32_t* counts, uint4* rects) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_preprocess, dim3((n + 255) / 256), dim3(256), 0, st, (const float4*)splats, n, v, fwd,
                       bwd, counts, rects);
}

void launch_emit_keys(hipStream_t st, uint32_t n, const uint32_t* counts, const uint64_t* offsets,
                      const uint4* rects, uint32_t tiles_x, uint64_t* keys) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_emit_keys, dim3((n + 255) / 256), dim3(256), 0, st, n, counts, offsets, rects, tiles_x,
                       keys);
}

void launch_tile_ranges(hipStream_t st, const uint64_t* keys, uint64_t m, uint32_t* ranges) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, keys, m, ranges);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint64_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}

}  // namespace nps
