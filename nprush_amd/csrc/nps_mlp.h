// nps_mlp.h -- the densification MLP: sizes, flat parameter layout, GEMM
// epilogues and launchers (k_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nps {

constexpr int kMlpIn = 32;      // parent features
constexpr int kMlpHidden = 256;
constexpr int kMlpOut = 14;     // 2 children x (3 position + 3 log-scale + 1 opacity logit)
constexpr int kMlpOutPad = 32;  // output rows padded for the 32-wide MFMA k steps
constexpr int kMlpLayers = 4;
// 32*256+256 + 2*(256*256+256) + 256*14+14
constexpr int kMlpParams = 143630;

enum MlpEpiMode { kEpiStore = 0, kEpiHidden = 1, kEpiOutput = 2, kEpiBackHidden = 3 };

struct MlpEpi {
    int mode;
    int m_valid;          // rows of real samples (hidden / backward) or of real outputs (store)
    const float* bias;    // hidden / output
    float* z;             // hidden: pre-activation written; backward: previous layer's read
    float* y;             // output / store
    uint16_t* out_rm;     // hidden / backward: bf16 row-major [M][N]
    uint16_t* out_fm;     // hidden / backward: bf16 feature-major [N][M]
};

struct MlpLayout {
    int w_off[kMlpLayers];    // offset of W_l [out][in] in the flat parameters; b_l follows
    int in[kMlpLayers], out[kMlpLayers], out_pad[kMlpLayers];
    uint16_t* wo[kMlpLayers];  // bf16 [out_pad][in]
    uint16_t* wi[kMlpLayers];  // bf16 [in][out_pad]
};

// Weights and biases the fused inference reads: wo[l] bf16 [out_pad][in], bias[l]
// FP32 (bias[3] padded to 32 entries, 16-byte aligned).
struct MlpInfer {
    const uint16_t* wo[kMlpLayers];
    const float* bias[kMlpLayers];
};

// C[M][N] = sum_k A[m][k] Bt[n][k]; M, N, K multiples of 32.
void launch_gemm_nt(hipStream_t st, const uint16_t* A, const uint16_t* Bt, int M, int N, int K, const MlpEpi& e);
void launch_mlp_input(hipStream_t st, const float* x, const float* mean, const float* stdv, int n, int M,
                      uint16_t* x_rm, uint16_t* x_fm);
void launch_mlp_output(hipStream_t st, const float* y, const float* mean, const float* stdv, int n, float* out);
// returns the number of loss partials written
int launch_mlp_loss(hipStream_t st, const float* y, const float* t, const float* mean, const float* stdv, int n, int M,
                    uint16_t* d_rm, uint16_t* d_fm, double* partial);
void launch_mlp_bias_grad(hipStream_t st, const uint16_t* d_fm, int M, int N, float* db);
// De-standardised outputs [n][14] of n feature rows [n][32] in one launch.
void launch_mlp_infer(hipStream_t st, const float* x, int n, const float* norm, const MlpInfer& P, float* out);
void launch_mlp_adam(hipStream_t st, float* p, const float* g, float* m, float* v, int total, float lr, float bc1,
                     float bc2, const MlpLayout& L);

}  // namespace nps
