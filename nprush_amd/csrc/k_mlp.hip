// k_mlp.hip -- the densification MLP (32 -> 256 -> 256 -> 256 -> 14, GELU) on
// gfx950 matrix cores: BF16 operands, FP32 accumulation (v_mfma_f32_16x16x32_bf16),
// FP32 master weights and Adam.  No reference counterpart.
//
// Every GEMM is C[M][N] = sum_k A[m][k] * Bt[n][k] with both operands stored
// k-contiguous, so each lane of an MFMA loads its 8 consecutive k values with one
// 16-byte load.  To keep that form for all products, activations and their
// gradients are written in two layouts (row-major [M][F] for the next layer's
// forward or the previous layer's backward, feature-major [F][M] for the weight
// gradients, which reduce over the batch), and the weights in two bf16 copies
// (out-major for the forward, in-major for the backward).  A wave computes a
// 32 x 32 tile (2 x 2 MFMAs), 4 waves per workgroup; the reduction over k runs
// in one wave in a fixed order, so every result is deterministic.
//
// MFMA 16x16x32 bf16 lane layout: A[i][k] in lane i + 16*(k/8), element k%8;
// B[k][j] in lane j + 16*(k/8), element k%8; D[i][j] in lane j + 16*(i/4),
// element i%4.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nps_mlp.h"

namespace nps {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
    // round to nearest even (NaN kept quiet)
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form 0.5 z (1 + tanh(u)), u = c (z + 0.044715 z^3), c = sqrt(2 / pi),
// evaluated as z * sigmoid(2u) = z / (1 + 2^(z (k0 + k1 z^2))) with
// k0 = -2 c log2(e), k1 = 0.044715 k0: one exp2 and one reciprocal, saturating
// without a branch (the power overflows to inf or flushes to 0).
constexpr float kGeluK0 = -2.302208198144325f;
constexpr float kGeluK1 = -0.1029432395800235f;
__device__ __forceinline__ float gelu_f(float z) {
    const float e = exp2f(z * fmaf(kGeluK1, z * z, kGeluK0));
    return __fdividef(z, 1.0f + e);
}
// d/dz: s + z s (1 - s) 2c (1 + 3 * 0.044715 z^2), s = sigmoid(2u)
__device__ __forceinline__ float gelu_d(float z) {
    const float z2 = z * z;
    const float e = exp2f(z * fmaf(kGeluK1, z2, kGeluK0));
    const float s = __fdividef(1.0f, 1.0f + e);
    return s + z * s * (1.0f - s) * (1.5957691216057308f * fmaf(0.134145f, z2, 1.0f));
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// two floats to bf16 (round to nearest even, v_cvt_pk_bf16_f32), a in the low half
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    uint4 q = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, q);
}

__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, const uint16_t* __restrict__ Bt,
                                                 int M, int N, int K, MlpEpi e) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    if (m0 >= M || n0 >= N) return;
    const int r = l & 15, kq = (l >> 4) * 8;
    f32x4 c00 = {0.f, 0.f, 0.f, 0.f}, c01 = c00, c10 = c00, c11 = c00;
    const uint16_t* a0p = A + (size_t)(m0 + r) * K + kq;
    const uint16_t* a1p = A + (size_t)(m0 + 16 + r) * K + kq;
    const uint16_t* b0p = Bt + (size_t)(n0 + r) * K + kq;
    const uint16_t* b1p = Bt + (size_t)(n0 + 16 + r) * K + kq;
    for (int kb = 0; kb < K; kb += 32) {
        bf16x8 a0 = ld8(a0p + kb), a1 = ld8(a1p + kb), b0 = ld8(b0p + kb), b1 = ld8(b1p + kb);
        c00 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c11, 0, 0, 0);
    }
    f32x4 cs[4] = {c00, c01, c10, c11};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int ti = t >> 1, tj = t & 1;
        const int n = n0 + 16 * tj + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = m0 + 16 * ti + (l >> 4) * 4 + i;
            const float acc = cs[t][i];
            const size_t rm = (size_t)m * N + n;    // row-major [M][N]
            const size_t fm = (size_t)n * M + m;    // feature-major [N][M]
            switch (e.mode) {
                case kEpiHidden: {                  // z = acc + b; h = gelu(z)
                    const float z = acc + e.bias[n];
                    e.z[rm] = z;
                    const uint16_t h = f2bf(m < e.m_valid ? gelu_f(z) : 0.0f);
                    e.out_rm[rm] = h;
                    e.out_fm[fm] = h;
                    break;
                }
                case kEpiOutput:                    // y = acc + b
                    e.y[rm] = acc + e.bias[n];
                    break;
                case kEpiBackHidden: {              // dz_prev = acc * gelu'(z_prev)
                    const float g = m < e.m_valid ? acc * gelu_d(e.z[rm]) : 0.0f;
                    const uint16_t h = f2bf(g);
                    e.out_rm[rm] = h;
                    e.out_fm[fm] = h;
                    break;
                }
                default:                            // plain FP32 store of the first m_valid rows (weight gradients)
                    if (m < e.m_valid) e.y[rm] = acc;
                    break;
            }
        }
    }
}

void launch_gemm_nt(hipStream_t st, const uint16_t* A, const uint16_t* Bt, int M, int N, int K, const MlpEpi& e) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, st, A, Bt, M, N, K, e);
}

// Standardised inputs in both layouts: x_rm[m][f], x_fm[f][m] (bf16); padded rows zero.
__global__ void k_mlp_input(const float* __restrict__ x, const float* __restrict__ mean, const float* __restrict__ stdv,
                            int n, int M, uint16_t* __restrict__ x_rm, uint16_t* __restrict__ x_fm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * kMlpIn) return;
    const int m = i / kMlpIn, f = i - m * kMlpIn;
    float v = 0.0f;
    if (m < n) v = (x[(size_t)m * kMlpIn + f] - mean[f]) / stdv[f];
    const uint16_t h = f2bf(v);
    x_rm[(size_t)m * kMlpIn + f] = h;
    x_fm[(size_t)f * M + m] = h;
}

void launch_mlp_input(hipStream_t st, const float* x, const float* mean, const float* stdv, int n, int M,
                      uint16_t* x_rm, uint16_t* x_fm) {
    const int tot = M * kMlpIn;
    hipLaunchKernelGGL(k_mlp_input, dim3((tot + 255) / 256), dim3(256), 0, st, x, mean, stdv, n, M, x_rm, x_fm);
}

// Outputs de-standardised: out[m][j] = y[m][j] * std[j] + mean[j], j < 14.
__global__ void k_mlp_output(const float* __restrict__ y, const float* __restrict__ mean,
                             const float* __restrict__ stdv, int n, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kMlpOut) return;
    const int m = i / kMlpOut, j = i - m * kMlpOut;
    out[i] = y[(size_t)m * kMlpOutPad + j] * stdv[j] + mean[j];
}

void launch_mlp_output(hipStream_t st, const float* y, const float* mean, const float* stdv, int n, float* out) {
    const int tot = n * kMlpOut;
    if (tot == 0) return;
    hipLaunchKernelGGL(k_mlp_output, dim3((tot + 255) / 256), dim3(256), 0, st, y, mean, stdv, n, out);
}

// MSE against standardised targets: d = 2 (y - t) / (n * 14) for real rows and
// outputs (0 elsewhere) in both bf16 layouts; per-block loss partials in a
// fixed order.
__global__ __launch_bounds__(256) void k_mlp_loss(const float* __restrict__ y, const float* __restrict__ t,
                                                  const float* __restrict__ mean, const float* __restrict__ stdv,
                                                  int n, int M, uint16_t* __restrict__ d_rm,
                                                  uint16_t* __restrict__ d_fm, double* __restrict__ partial) {
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0;
    if (i < M * kMlpOutPad) {
        const int m = i / kMlpOutPad, j = i - m * kMlpOutPad;
        float g = 0.0f;
        if (m < n && j < kMlpOut) {
            const float tn = (t[(size_t)m * kMlpOut + j] - mean[j]) / stdv[j];
            const float r = y[i] - tn;
            sq = (double)r * (double)r;
            g = 2.0f * r / ((float)n * (float)kMlpOut);
        }
        const uint16_t h = f2bf(g);
        d_rm[i] = h;
        d_fm[(size_t)j * M + m] = h;
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

int launch_mlp_loss(hipStream_t st, const float* y, const float* t, const float* mean, const float* stdv, int n, int M,
                    uint16_t* d_rm, uint16_t* d_fm, double* partial) {
    const int tot = M * kMlpOutPad, nb = (tot + 255) / 256;
    hipLaunchKernelGGL(k_mlp_loss, dim3(nb), dim3(256), 0, st, y, t, mean, stdv, n, M, d_rm, d_fm, partial);
    return nb;
}

// db[j] = sum over the batch of the feature-major gradient row j (fixed order).
__global__ void k_mlp_bias_grad(const uint16_t* __restrict__ d_fm, int M, int N, float* __restrict__ db) {
    const int j = blockIdx.x;
    if (j >= N) return;
    __shared__ float red[256];
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += 256) s += bf2f(d_fm[(size_t)j * M + m]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[j] = red[0];
}

void launch_mlp_bias_grad(hipStream_t st, const uint16_t* d_fm, int M, int N, float* db) {
    hipLaunchKernelGGL(k_mlp_bias_grad, dim3(N), dim3(256), 0, st, d_fm, M, N, db);
}

// Adam over the FP32 master parameters; refreshes both bf16 copies of each
// weight matrix.  Layer l has W [out][in] then b [out] in the flat order.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, int total, float lr, float bc1, float bc2, MlpLayout L) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float gi = g[i];
    const float mm = 0.9f * m[i] + 0.1f * gi;
    const float vv = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mm;
    v[i] = vv;
    const float w = p[i] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
    p[i] = w;
#pragma unroll
    for (int l = 0; l < kMlpLayers; ++l) {
        const int off = L.w_off[l], nin = L.in[l], nout = L.out[l];
        if (i >= off && i < off + nin * nout) {
            const int o = (i - off) / nin, k = (i - off) - o * nin;
            const uint16_t h = f2bf(w);
            L.wo[l][(size_t)o * nin + k] = h;     // [out_pad][in]: padded rows stay zero
            L.wi[l][(size_t)k * L.out_pad[l] + o] = h;  // [in][out_pad]
        }
    }
}

void launch_mlp_adam(hipStream_t st, float* p, const float* g, float* m, float* v, int total, float lr, float bc1,
                     float bc2, const MlpLayout& L) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((total + 255) / 256), dim3(256), 0, st, p, g, m, v, total, lr, bc1, bc2, L);
}


// Inference: all four layers in one launch.  A workgroup takes 64 rows; their
// activations stay in LDS (two bf16 buffers, rows padded to 264 elements so
// the 16 rows an MFMA operand reads fall on different banks) and the bf16
// out-major weights stream from L2.  Each product is computed transposed,
// D[n][m] = sum_k W[n][k] act[m][k], so a lane's four results are consecutive
// output features of one row: one 8-byte LDS store after bias + GELU.  Wave w
// owns output features [64w, 64w + 64) of the hidden layers (4 x 4 MFMA tiles)
// and rows [16w, 16w + 16) of the output layer.
constexpr int kInfRows = 64;
constexpr int kInfStride = kMlpHidden + 8;

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) {
    uint4 q = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, q);
}

template <int K>
__device__ __forceinline__ void infer_hidden(const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                             const uint16_t* src, uint16_t* dst, int w, int l) {
    const int r = l & 15, kq = (l >> 4) * 8;
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* wp = W + (size_t)(64 * w + r) * K + kq;
#pragma unroll 2
    for (int kb = 0; kb < K; kb += 32) {
        bf16x8 wa[4], xb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) wa[t] = ld8(wp + (size_t)16 * t * K + kb);
#pragma unroll
        for (int t = 0; t < 4; ++t) xb[t] = lds8(src + (16 * t + r) * kInfStride + kb + kq);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[a], xb[b], acc[a][b], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int n = 64 * w + 16 * a + 4 * (l >> 4);
        const float4 bv = *reinterpret_cast<const float4*>(bias + n);
        const float bb[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            uint2 pk;
            pk.x = pack_bf16(gelu_f(acc[a][b][0] + bb[0]), gelu_f(acc[a][b][1] + bb[1]));
            pk.y = pack_bf16(gelu_f(acc[a][b][2] + bb[2]), gelu_f(acc[a][b][3] + bb[3]));
            *reinterpret_cast<uint2*>(dst + (16 * b + r) * kInfStride + n) = pk;
        }
    }
}

__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ x, int n, const float* __restrict__ norm,
                                                   MlpInfer P, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t act[2][kInfRows * kInfStride];
    const int row0 = blockIdx.x * kInfRows;
    for (int i = threadIdx.x; i < kInfRows * kMlpIn; i += 256) {
        const int m = i / kMlpIn, f = i - m * kMlpIn;
        float v = 0.0f;
        if (row0 + m < n) v = (x[(size_t)(row0 + m) * kMlpIn + f] - norm[f]) / norm[32 + f];
        act[0][m * kInfStride + f] = f2bf(v);
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    infer_hidden<kMlpIn>(P.wo[0], P.bias[0], act[0], act[1], w, l);
    __syncthreads();
    infer_hidden<kMlpHidden>(P.wo[1], P.bias[1], act[1], act[0], w, l);
    __syncthreads();
    infer_hidden<kMlpHidden>(P.wo[2], P.bias[2], act[0], act[1], w, l);
    __syncthreads();
    // output layer: 16 (of 32 padded) output rows of W3 x the wave's 16 rows
    const int r = l & 15, kq = (l >> 4) * 8;
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* src = act[1] + (16 * w + r) * kInfStride + kq;
    const uint16_t* wp = P.wo[3] + (size_t)r * kMlpHidden + kq;
#pragma unroll
    for (int kb = 0; kb < kMlpHidden; kb += 32)
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(wp + kb), lds8(src + kb), c, 0, 0, 0);
    const int m = row0 + 16 * w + r;
    if (m < n) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int j = 4 * (l >> 4) + e;
            if (j < kMlpOut) out[(size_t)m * kMlpOut + j] = (c[e] + P.bias[3][j]) * norm[64 + kMlpOut + j] + norm[64 + j];
        }
    }
}

void launch_mlp_infer(hipStream_t st, const float* x, int n, const float* norm, const MlpInfer& P, float* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + kInfRows - 1) / kInfRows), dim3(256), 0, st, x, n, norm, P, out);
}

}  // namespace nps
