// k_render.hip -- the "render" entry: splat_forward.wgsl's instanced-quad
// raster with fixed-function "over" blending, as a tiled compute pass.
//
// The reference draws N quads in reverse index order (instance ii -> splat
// N-1-ii, FWD:44), culls w < 0.1 (FWD:62-66), sizes each quad by per-axis
// extents R*sqrt(cov00), R*sqrt(cov11) (FWD:99-100), discards alpha < 0.001
// (FWD:143-145) and blends src*a + dst*(1-a) (alpha: a + dst*(1-a)) into an
// rgba8unorm colour target and an encoded-depth target (FWDM:54-82).
// Here: one record per splat, the same tile keys/sort as the training path,
// and one thread per pixel walking its tile list from the highest id down.
// A pixel is covered when its centre lies in the closed quad
// [centre - extent, centre + extent] (raster edge rules are not modelled).
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

template <class P>
__device__ __forceinline__ uint32_t first_false_r(uint32_t n, P pred) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (pred(mid)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// rec: [scx scy A B] [C op r g] [b sa sb enc] [ext_x ext_y 0 0]
__global__ __launch_bounds__(256) void k_preprocess_render(const float4* __restrict__ splats, uint32_t n,
                                                           ViewParams v, float4* __restrict__ rec,
                                                           uint32_t* __restrict__ counts,
                                                           uint4* __restrict__ rects) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    counts[i] = 0u;
    float4 s0 = splats[4 * i + 0], s1 = splats[4 * i + 1], s2 = splats[4 * i + 2], s3 = splats[4 * i + 3];
    float sx = fmax_w(s0.w, 0.0001f), sy = fmax_w(s3.x, 0.0001f), sz = fmax_w(s3.w, 0.0001f);
    float shape_a = s3.y, shape_b = s3.z;
    float Rr = pow_det(15.0f / fmax_w(shape_b, 0.0001f), 1.0f / shape_a);
    F4 clip = mat_vec(v.vp, s0.x, s0.y, s0.z, 1.0f);
    float w = clip.w;
    if (w < 0.1f) return;
    float aspect = v.aspect;
    F2 ax_s = project_axis(v.vp, quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{1.0f, 0.0f, 0.0f}), clip.x, clip.y, w, aspect);
    F2 ay_s = project_axis(v.vp, quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{0.0f, 1.0f, 0.0f}), clip.x, clip.y, w, aspect);
    F2 az_s = project_axis(v.vp, quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{0.0f, 0.0f, 1.0f}), clip.x, clip.y, w, aspect);
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
    float det = cov00 * cov11 - cov01 * cov01;
    float inv_det = (fabsf(det) < 1e-10f) ? 0.0f : 1.0f / det;
    float A = cov11 * inv_det, B = -cov01 * inv_det, C = cov00 * inv_det;
    float ext_x = Rr * sqrtf(fmax_w(cov00, 1e-9f));
    float ext_y = Rr * sqrtf(fmax_w(cov11, 1e-9f));
    float scx = clip.x / w * aspect, scy = clip.y / w;
    float lin = fmax_w(w, 0.1f);
    float enc = clamp_w(1.0f - 0.1f / lin, 0.0f, 1.0f);
    float sminx = scx - ext_x, smaxx = scx + ext_x, sminy = scy - ext_y, smaxy = scy + ext_y;
    const uint32_t W = v.W, H = v.H, T = (uint32_t)kTile;
    uint32_t tx0 = first_false_r(v.tiles_x, [&](uint32_t tx) { return sminx > px_to_x(min(tx * T + T, W), W, aspect); });
    uint32_t tx_end = first_false_r(v.tiles_x, [&](uint32_t tx) { return !(smaxx < px_to_x(tx * T, W, aspect)); });
    uint32_t ty0 = first_false_r(v.tiles_y, [&](uint32_t ty) { return smaxy < py_to_y(min(ty * T + T, H), H); });
    uint32_t ty_end = first_false_r(v.tiles_y, [&](uint32_t ty) { return !(sminy > py_to_y(ty * T, H)); });
    if (tx0 >= tx_end || ty0 >= ty_end) return;
    counts[i] = (tx_end - tx0) * (ty_end - ty0);
    rects[i] = make_uint4(tx0, tx_end, ty0, ty_end);
    rec[4 * (size_t)i + 0] = make_float4(scx, scy, A, B);
    rec[4 * (size_t)i + 1] = make_float4(C, s1.w, s1.x, s1.y);
    rec[4 * (size_t)i + 2] = make_float4(s1.z, shape_a, shape_b, enc);
    rec[4 * (size_t)i + 3] = make_float4(ext_x, ext_y, 0.0f, 0.0f);
}

__device__ __forceinline__ float q8(float x) { return rintf(clamp_w(x, 0.0f, 1.0f) * 255.0f) / 255.0f; }

__global__ __launch_bounds__(256) void k_render(ViewParams v, RenderArgs A) {
    __shared__ float4 s_rec[kBatch][4];
    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t start = A.ranges[2 * tile], end = A.ranges[2 * tile + 1];
    const int count = (int)(end - start);
    const int tid = threadIdx.x;
    const uint32_t px = tx * kTile + (tid & 15), py = ty * kTile + (tid >> 4);
    const bool valid = px < v.W && py < v.H;
    const float pxf = px_to_x(px, v.W, v.aspect), pyf = py_to_y(py, v.H);
    const size_t pix = (size_t)py * v.W + px;
    float4 c, d;
    if (A.clear) {
        // clear value (0.05, 0.05, 0.05, 1) as the target stores it
        float cv = A.quantize ? q8(0.05f) : 0.05f;
        c = make_float4(cv, cv, cv, 1.0f);
        d = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    } else if (valid) {
        c = A.color[pix];
        d = A.depth[pix];
    } else {
        c = d = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int nbatch = (count + kBatch - 1) / kBatch;
    for (int bi = nbatch - 1; bi >= 0; --bi) {
        const int b0 = bi * kBatch, nb = min(kBatch, count - b0);
        __syncthreads();
        {
            int j = tid >> 2, part = tid & 3;
            if (j < nb) {
                uint32_t id = (uint32_t)A.keys[start + b0 + j];
                s_rec[j][part] = A.rec[4 * (size_t)id + part];
            }
        }
        __syncthreads();
        for (int j = nb - 1; j >= 0; --j) {
            float4 r0 = s_rec[j][0], r1 = s_rec[j][1], r2_ = s_rec[j][2], r3 = s_rec[j][3];
            float dx = pxf - r0.x, dy = pyf - r0.y;
            if (!(dx >= -r3.x && dx <= r3.x && dy >= -r3.y && dy <= r3.y)) continue;
            float Aa = r0.z, Bb = r0.w, Cc = r1.x;
            float rr2 = ((Aa * dx) * dx + ((2.0f * Bb) * dx) * dy) + (Cc * dy) * dy;
            float lr = log2_sqrt_det(fmax_w(rr2, 0.0001f));
            float pw = -r2_.z * exp2_det(r2_.y * lr);
            float a = (pw > -15.0f) ? exp_det(pw) * r1.y : 0.0f;
            a = clamp_w(a, 0.0f, 0.999f);
            if (a < 0.001f) continue;
            float enc = r2_.w, one_m = 1.0f - a;
            c.x = r1.z * a + c.x * one_m;
            c.y = r1.w * a + c.y * one_m;
            c.z = r2_.x * a + c.z * one_m;
            c.w = a + c.w * one_m;
            d.x = enc * a + d.x * one_m;
            d.y = enc * a + d.y * one_m;
            d.z = enc * a + d.z * one_m;
            d.w = a + d.w * one_m;
            if (A.quantize) {
                c = make_float4(q8(c.x), q8(c.y), q8(c.z), q8(c.w));
                d = make_float4(q8(d.x), q8(d.y), q8(d.z), q8(d.w));
            }
        }
    }
    if (valid) {
        A.color[pix] = c;
        A.depth[pix] = d;
    }
}

void launch_preprocess_render(hipStream_t st, const float* splats, uint32_t n, const ViewParams& v, float4* rec,
                              uint32_t* counts, uint4* rects) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_preprocess_render, dim3((n + 255) / 256), dim3(256), 0, st, (const float4*)splats, n, v, rec,
                       counts, rects);
}

void launch_render(hipStream_t st, const ViewParams& v, const RenderArgs& a) {
    if (v.num_tiles == 0) return;
    hipLaunchKernelGGL(k_render, dim3(v.num_tiles), dim3(256), 0, st, v, a);
}

}  // namespace nps
