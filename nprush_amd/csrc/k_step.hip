// k_step.hip -- fused Adam step (REF semantics of splat_step.wgsl:32-106).
//
// One thread per splat; 16 params are read and written as float4/int4 vectors:
// read splat 64 B + grads 64 B + m,v 128 B + grad_accum 4 B, write the same
// (grads written as zero) = 520 B per splat, all HBM traffic.
// The reference reads adam.t in every workgroup while one thread increments
// it (STEP:35-37); here t (= previous t + 1) is passed in by the host.
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

// splat_step.wgsl:48-52,85-86
__constant__ float c_lr[16] = {0.0005f, 0.0005f, 0.0005f, 0.01f, 0.02f, 0.02f, 0.02f, 0.01f,
                               0.005f,  0.005f,  0.005f,  0.005f, 0.01f, 0.01f, 0.01f, 0.01f};
__constant__ float c_mu[16] = {0.005f, 0.005f, 0.005f, 0.005f, 0.001f, 0.001f, 0.001f, 0.0005f,
                               0.005f, 0.005f, 0.005f, 0.005f, 0.005f, 0.05f,  0.05f,  0.005f};
__constant__ float c_fps[16] = {10000.0f,  10000.0f,  10000.0f,  10000.0f, 100000.0f, 100000.0f,
                                100000.0f, 100000.0f, 10000.0f,  10000.0f, 10000.0f,  10000.0f,
                                10000.0f,  10000.0f,  10000.0f,  10000.0f};
__constant__ float c_lo[16] = {-1e9f, -1e9f, -1e9f, 0.001f, 0.05f, 0.05f, 0.05f, 0.01f,
                               -1e9f, -1e9f, -1e9f, -1e9f,  0.001f, 0.1f, 0.01f, 0.001f};
__constant__ float c_hi[16] = {1e9f, 1e9f, 1e9f, 2.0f, 1.0f, 1.0f, 1.0f, 0.99f,
                               1e9f, 1e9f, 1e9f, 1e9f, 2.0f, 10.0f, 5.0f, 2.0f};

__global__ __launch_bounds__(256) void k_adam_ref(float4* __restrict__ splats, int4* __restrict__ grads,
                                                  float4* __restrict__ m, float4* __restrict__ v,
                                                  float* __restrict__ accum, uint32_t n, float t) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - pow_det(b1, t);
    const float bc2 = 1.0f - pow_det(b2, t);
    float s[16], mm[16], vv[16];
    int32_t g[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float4 a = splats[4 * (size_t)i + q];
        float4 b = m[4 * (size_t)i + q];
        float4 c = v[4 * (size_t)i + q];
        int4 d = grads[4 * (size_t)i + q];
        s[4 * q + 0] = a.x; s[4 * q + 1] = a.y; s[4 * q + 2] = a.z; s[4 * q + 3] = a.w;
        mm[4 * q + 0] = b.x; mm[4 * q + 1] = b.y; mm[4 * q + 2] = b.z; mm[4 * q + 3] = b.w;
        vv[4 * q + 0] = c.x; vv[4 * q + 1] = c.y; vv[4 * q + 2] = c.z; vv[4 * q + 3] = c.w;
        g[4 * q + 0] = d.x; g[4 * q + 1] = d.y; g[4 * q + 2] = d.z; g[4 * q + 3] = d.w;
    }
    float pos_grad_norm2 = 0.0f;
#pragma unroll
    for (int lp = 0; lp < 16; ++lp) {
        float grad = (float)g[lp] / c_fps[lp] / 16384.0f;
        float m1 = b1 * mm[lp] + (1.0f - b1) * grad;
        float v1 = b2 * vv[lp] + ((1.0f - b2) * grad) * grad;
        mm[lp] = m1;
        vv[lp] = v1;
        float m_hat = m1 / bc1;
        float v_hat = v1 / bc2;
        float raw_update = (c_lr[lp] * m_hat) / (sqrtf(v_hat) + eps);
        float update = clamp_w(raw_update, -c_mu[lp], c_mu[lp]);
        s[lp] = clamp_w(s[lp] - update, c_lo[lp], c_hi[lp]);
        if (lp <= 2) pos_grad_norm2 = pos_grad_norm2 + grad * grad;
    }
    float q_len = fmax_w(sqrtf(((s[8] * s[8] + s[9] * s[9]) + s[10] * s[10]) + s[11] * s[11]), 1e-8f);
    s[8] = s[8] / q_len;
    s[9] = s[9] / q_len;
    s[10] = s[10] / q_len;
    s[11] = s[11] / q_len;
    accum[i] = accum[i] + sqrtf(pos_grad_norm2);
    float volume = (s[3] * s[12]) * s[15];
    if (s[7] < 0.05f || volume < 1e-7f) s[7] = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        splats[4 * (size_t)i + q] = make_float4(s[4 * q + 0], s[4 * q + 1], s[4 * q + 2], s[4 * q + 3]);
        m[4 * (size_t)i + q] = make_float4(mm[4 * q + 0], mm[4 * q + 1], mm[4 * q + 2], mm[4 * q + 3]);
        v[4 * (size_t)i + q] = make_float4(vv[4 * q + 0], vv[4 * q + 1], vv[4 * q + 2], vv[4 * q + 3]);
        grads[4 * (size_t)i + q] = make_int4(0, 0, 0, 0);
    }
}

void launch_adam_ref(hipStream_t st, float* splats, int32_t* grads, float* m, float* v, float* accum, uint32_t n,
                     float t) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_adam_ref, dim3((n + 255) / 256), dim3(256), 0, st, (float4*)splats, (int4*)grads,
                       (float4*)m, (float4*)v, accum, n, t);
}

}  // namespace nps
