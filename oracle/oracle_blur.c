/*
 * oracle_blur.c -- CPU restatement of the reference's separable target blur
 * (src/lib/gpu/blur.wgsl:13-48, driven by GpuBlurPipelineManager.blur,
 * src/lib/gpu/GpuBlurPipelineManager.ts:51-99).  TEST INFRASTRUCTURE ONLY.
 *
 * Images are H*W*4 floats holding what an rgba8unorm texture returns (k/255).
 * Pass 1 (direction (1,0)): decode sRGB -> linear if is_srgb (flags bit 0),
 * Gaussian-weighted mean over i in [-radius, radius] with clamp-to-edge,
 * stored to an rgba8unorm temp (quantised, alpha 1).  Pass 2 (direction
 * (0,1)): the same over the temp, encode linear -> sRGB if is_srgb (flags bit
 * 1), stored rgba8unorm.  FP32 throughout; transcendentals from ref_math.h.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "ref_math.h"

/* WGSL constant 1.0 / 2.4 is folded exactly, then rounded to f32 */
#define INV_GAMMA ((float)(1.0 / 2.4))

static float srgb_to_linear(float c) {                       /* blur.wgsl:13-15 */
    return c > 0.04045f ? rm_powf((c + 0.055f) / 1.055f, 2.4f) : c / 12.92f;
}
static float linear_to_srgb(float c) {                       /* blur.wgsl:17-20 */
    float s = c > 0.0f ? c : 0.0f;
    return s > 0.0031308f ? 1.055f * rm_powf(s, INV_GAMMA) - 0.055f : s * 12.92f;
}
/* rgba8unorm store: clamp to [0,1], round to nearest (ties to even) */
static float unorm8(float x) {
    if (!(x > 0.0f)) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return rintf(x * 255.0f) / 255.0f;
}

static void blur_pass(const float* src, float* dst, int W, int H, int dx, int dy, int radius, float sigma,
                      int in_srgb, int out_srgb) {
    const float den = 2.0f * sigma * sigma;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum[3] = { 0.0f, 0.0f, 0.0f }, ws = 0.0f;
            for (int i = -radius; i <= radius; ++i) {
                int cx = x + dx * i, cy = y + dy * i;
                cx = cx < 0 ? 0 : (cx > W - 1 ? W - 1 : cx);
                cy = cy < 0 ? 0 : (cy > H - 1 ? H - 1 : cy);
                const float* p = src + 4 * ((size_t)cy * W + cx);
                float w = rm_expf(-(float)(i * i) / den);
                for (int c = 0; c < 3; ++c) {
                    float col = in_srgb ? srgb_to_linear(p[c]) : p[c];
                    sum[c] = sum[c] + col * w;
                }
                ws = ws + w;
            }
            float* q = dst + 4 * ((size_t)y * W + x);
            for (int c = 0; c < 3; ++c) {
                float r = sum[c] / ws;
                q[c] = unorm8(out_srgb ? linear_to_srgb(r) : r);
            }
            q[3] = 1.0f;
        }
}

/* GpuBlurPipelineManager.blur(src, dst, temp, w, h, radius, sigma, isSrgb) */
int ora_blur_f32(const float* src, float* dst, float* temp, int W, int H, int radius, float sigma, int is_srgb) {
    if (W <= 0 || H <= 0 || radius < 0) return -1;
    blur_pass(src, temp, W, H, 1, 0, radius, sigma, is_srgb, 0);
    blur_pass(temp, dst, W, H, 0, 1, radius, sigma, 0, is_srgb);
    return 0;
}
