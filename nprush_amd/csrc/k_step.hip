// k_step.hip -- fused Adam step (REF semantics of splat_step.wgsl:32-106).
//
// Four threads per splat, one float4/int4 of each array per thread: read
// splat 64 B + grads 64 B + m,v 128 B + grad_accum 4 B, write the same (grads
// written as zero) = 520 B per splat, all HBM traffic.
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

// Four lanes per splat, lane q = tid % 4 owning parameters 4q..4q+3 (one
// float4 of each array): consecutive lanes touch consecutive 16-byte words, so
// every load and store is fully coalesced.  The per-splat parts: the position
// gradient norm (lane 0), the quaternion renormalisation (lane 2 holds all of
// qw..qz), and the kill test of lane 1 (opacity), which takes scale x from
// lane 0 and scales y, z from lane 3 over the quad (DPP quad_perm).
__global__ __launch_bounds__(256) void k_adam_ref(float4* __restrict__ splats, int4* __restrict__ grads,
                                                  float4* __restrict__ m, float4* __restrict__ v,
                                                  float* __restrict__ accum, uint32_t n, float t) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = (int)(threadIdx.x & 3u);
    const size_t i = idx >> 2;
    const bool live = i < n;  // whole quads are live or not (n * 4 threads)
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - pow_det(b1, t);
    const float bc2 = 1.0f - pow_det(b2, t);
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a;
    int4 d = make_int4(0, 0, 0, 0);
    if (live) {
        a = splats[idx];
        b = m[idx];
        c = v[idx];
        d = grads[idx];
    }
    float s[4] = {a.x, a.y, a.z, a.w}, mm[4] = {b.x, b.y, b.z, b.w}, vv[4] = {c.x, c.y, c.z, c.w};
    const int32_t g[4] = {d.x, d.y, d.z, d.w};
    float pos_grad_norm2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int lp = 4 * q + k;
        float grad = (float)g[k] / c_fps[lp] / 16384.0f;
        float m1 = b1 * mm[k] + (1.0f - b1) * grad;
        float v1 = b2 * vv[k] + ((1.0f - b2) * grad) * grad;
        mm[k] = m1;
        vv[k] = v1;
        float m_hat = m1 / bc1;
        float v_hat = v1 / bc2;
        float raw_update = (c_lr[lp] * m_hat) / (sqrtf(v_hat) + eps);
        float update = clamp_w(raw_update, -c_mu[lp], c_mu[lp]);
        s[k] = clamp_w(s[k] - update, c_lo[lp], c_hi[lp]);
        if (lp <= 2) pos_grad_norm2 = pos_grad_norm2 + grad * grad;
    }
    if (q == 2) {  // params 8..11: the quaternion
        float q_len = fmax_w(sqrtf(((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]) + s[3] * s[3]), 1e-8f);
        s[0] = s[0] / q_len;
        s[1] = s[1] / q_len;
        s[2] = s[2] / q_len;
        s[3] = s[3] / q_len;
    }
    // scale x (param 3, lane 0 .w), scales y and z (params 12 and 15, lane 3 .x and .w)
    // to lane 1 of the quad: quad_perm [0,0,0,0] and [3,3,3,3] broadcasts
    const float sx = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s[3]), 0x00, 0xF, 0xF, false));
    const float sy = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s[0]), 0xFF, 0xF, 0xF, false));
    const float sz = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s[3]), 0xFF, 0xF, 0xF, false));
    if (!live) return;
    if (q == 0) accum[i] = accum[i] + sqrtf(pos_grad_norm2);
    if (q == 1) {  // param 7: opacity
        float volume = (sx * sy) * sz;
        if (s[3] < 0.05f || volume < 1e-7f) s[3] = 0.0f;
    }
    splats[idx] = make_float4(s[0], s[1], s[2], s[3]);
    m[idx] = make_float4(mm[0], mm[1], mm[2], mm[3]);
    v[idx] = make_float4(vv[0], vv[1], vv[2], vv[3]);
    grads[idx] = make_int4(0, 0, 0, 0);
}

void launch_adam_ref(hipStream_t st, float* splats, int32_t* grads, float* m, float* v, float* accum, uint32_t n,
                     float t) {
    if (n == 0) return;
    const size_t threads = 4 * (size_t)n;
    hipLaunchKernelGGL(k_adam_ref, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, (float4*)splats,
                       (int4*)grads, (float4*)m, (float4*)v, accum, n, t);
}

}  // namespace nps
