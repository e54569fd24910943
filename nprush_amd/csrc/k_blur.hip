// k_blur.hip -- separable Gaussian blur of a target image (REF semantics of
// src/lib/gpu/blur.wgsl:13-48 as GpuBlurPipelineManager.blur runs it,
// GpuBlurPipelineManager.ts:51-99): a horizontal pass (sRGB decode when asked)
// into an 8-bit linear temp, then a vertical pass (sRGB encode when asked)
// into the 8-bit output.  Images are H*W float4 holding what an rgba8unorm
// texture returns (k/255).
//
// One thread per output pixel; the 2*radius+1 taps of a row (or column) are
// read through L1/L2 (the passes run once per target at the optimisation
// resolution and move 2 x 16 B per pixel of HBM traffic).
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

namespace {
constexpr float kInvGamma = (float)(1.0 / 2.4);  // WGSL folds 1.0 / 2.4 exactly, then rounds

__device__ __forceinline__ float srgb_to_linear(float c) {
    return c > 0.04045f ? pow_det((c + 0.055f) / 1.055f, 2.4f) : c / 12.92f;
}
__device__ __forceinline__ float linear_to_srgb(float c) {
    float s = c > 0.0f ? c : 0.0f;
    return s > 0.0031308f ? 1.055f * pow_det(s, kInvGamma) - 0.055f : s * 12.92f;
}
// rgba8unorm store: clamp to [0, 1], round to nearest (ties to even)
__device__ __forceinline__ float unorm8(float x) {
    if (!(x > 0.0f)) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return rintf(x * 255.0f) / 255.0f;
}
}  // namespace

// Texel access of the two passes: float4 texels (k/255 per channel), or the
// 8-bit temp as uchar4 holding k itself.  (float)k / 255 is exactly the value
// unorm8 produced, so both temps give the same bits.
__device__ __forceinline__ float4 load_texel(const float4* p, size_t i) { return p[i]; }
__device__ __forceinline__ float4 load_texel(const uchar4* p, size_t i) {
    const uchar4 k = p[i];
    return make_float4((float)k.x / 255.0f, (float)k.y / 255.0f, (float)k.z / 255.0f, 1.0f);
}
__device__ __forceinline__ void store_texel(float4* p, size_t i, float r0, float r1, float r2) {
    p[i] = make_float4(unorm8(r0), unorm8(r1), unorm8(r2), 1.0f);
}
__device__ __forceinline__ unsigned char unorm8_k(float x) {
    if (!(x > 0.0f)) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return (unsigned char)rintf(x * 255.0f);
}
__device__ __forceinline__ void store_texel(uchar4* p, size_t i, float r0, float r1, float r2) {
    p[i] = make_uchar4(unorm8_k(r0), unorm8_k(r1), unorm8_k(r2), 255);
}

template <class In, class Out>
__global__ __launch_bounds__(256) void k_blur_pass(const In* __restrict__ src, Out* __restrict__ dst,
                                                   int W, int H, int dx, int dy, int radius, float sigma,
                                                   int in_srgb, int out_srgb) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float den = 2.0f * sigma * sigma;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, ws = 0.0f;
    for (int i = -radius; i <= radius; ++i) {
        int cx = min(max(x + dx * i, 0), W - 1);
        int cy = min(max(y + dy * i, 0), H - 1);
        float4 p = load_texel(src, (size_t)cy * W + cx);
        float w = exp_det(-(float)(i * i) / den);
        float c0 = p.x, c1 = p.y, c2 = p.z;
        if (in_srgb) {
            c0 = srgb_to_linear(c0);
            c1 = srgb_to_linear(c1);
            c2 = srgb_to_linear(c2);
        }
        s0 = s0 + c0 * w;
        s1 = s1 + c1 * w;
        s2 = s2 + c2 * w;
        ws = ws + w;
    }
    float r0 = s0 / ws, r1 = s1 / ws, r2 = s2 / ws;
    if (out_srgb) {
        r0 = linear_to_srgb(r0);
        r1 = linear_to_srgb(r1);
        r2 = linear_to_srgb(r2);
    }
    store_texel(dst, (size_t)y * W + x, r0, r1, r2);
}

void launch_blur(hipStream_t st, const float4* src, float4* dst, float4* temp, int W, int H, int radius, float sigma,
                 int is_srgb) {
    if (W <= 0 || H <= 0) return;
    dim3 g((W + 15) / 16, (H + 15) / 16), b(256);
    hipLaunchKernelGGL((k_blur_pass<float4, float4>), g, b, 0, st, src, temp, W, H, 1, 0, radius, sigma, is_srgb, 0);
    hipLaunchKernelGGL((k_blur_pass<float4, float4>), g, b, 0, st, (const float4*)temp, dst, W, H, 0, 1, radius,
                       sigma, 0, is_srgb);
}

void launch_blur_u8temp(hipStream_t st, const float4* src, float4* dst, uchar4* temp, int W, int H, int radius,
                        float sigma, int is_srgb) {
    if (W <= 0 || H <= 0) return;
    dim3 g((W + 15) / 16, (H + 15) / 16), b(256);
    hipLaunchKernelGGL((k_blur_pass<float4, uchar4>), g, b, 0, st, src, temp, W, H, 1, 0, radius, sigma, is_srgb, 0);
    hipLaunchKernelGGL((k_blur_pass<uchar4, float4>), g, b, 0, st, (const uchar4*)temp, dst, W, H, 0, 1, radius,
                       sigma, 0, is_srgb);
}

}  // namespace nps
