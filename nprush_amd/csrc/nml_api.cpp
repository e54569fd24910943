// nml_api.cpp -- C ABI of the densification MLP (nml_*, include/nprush_hip.h).
// Forward: standardised features -> 3 x (BF16 GEMM + bias + GELU) -> BF16 GEMM
// + bias -> de-standardised deltas.  Training: standardised-target MSE, the
// backward as BF16 GEMMs with FP32 accumulation, Adam on FP32 master weights.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/nprush_hip.h"
#include "nps_mlp.h"

using namespace nps;

namespace {

constexpr int kIn[kMlpLayers] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden};
constexpr int kOut[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpOut};
constexpr int kOutPad[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpOutPad};

__global__ void k_mlp_refresh(const float* __restrict__ p, MlpLayout L) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int l = 0; l < kMlpLayers; ++l) {
        const int off = L.w_off[l], nin = L.in[l], nout = L.out[l];
        if (i >= off && i < off + nin * nout) {
            const int o = (i - off) / nin, k = (i - off) - o * nin;
            uint32_t u = __float_as_uint(p[i]);
            uint16_t h;
            if ((u & 0x7FFFFFFFu) > 0x7F800000u) h = (uint16_t)((u >> 16) | 0x0040u);
            else h = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
            L.wo[l][(size_t)o * nin + k] = h;
            L.wi[l][(size_t)k * L.out_pad[l] + o] = h;
        }
    }
}

}  // namespace

struct nml_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int max_m = 0;  // batch capacity, a multiple of 32
    float *params = nullptr, *grads = nullptr, *am = nullptr, *av = nullptr;
    float* bias4 = nullptr;  // output bias padded to 32
    float* norm = nullptr;   // in_mean[32], in_std[32], out_mean[14], out_std[14]
    uint16_t *wo[kMlpLayers] = {nullptr}, *wi[kMlpLayers] = {nullptr};
    uint16_t *x_rm = nullptr, *x_fm = nullptr;
    float* z[3] = {nullptr};
    uint16_t *h_rm[3] = {nullptr}, *h_fm[3] = {nullptr};
    uint16_t *d_rm[4] = {nullptr}, *d_fm[4] = {nullptr};  // [3]: output gradient, [0..2]: hidden
    float* y = nullptr;
    float *x_tmp = nullptr, *t_tmp = nullptr, *y_tmp = nullptr;
    double* partial = nullptr;
    float adam_t = 0.0f;
    MlpLayout L;
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float timing_ms[2] = {0.0f, 0.0f};  // last inference launch, last training step
    std::string err;
};

namespace {

int ml_err(nml_ctx* h, int code, const std::string& m) {
    if (h) h->err = m;
    return code;
}
#define NML_HIP(h, expr)                                                                                  \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return ml_err((h), NPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int w_offset(int l) {
    int off = 0;
    for (int k = 0; k < l; ++k) off += kIn[k] * kOut[k] + kOut[k];
    return off;
}

void refresh(nml_ctx* h) {
    hipLaunchKernelGGL(k_mlp_refresh, dim3((kMlpParams + 255) / 256), dim3(256), 0, h->stream, h->params, h->L);
    // padded output bias
    (void)hipMemsetAsync(h->bias4, 0, 4 * kMlpOutPad, h->stream);
    (void)hipMemcpyAsync(h->bias4, h->params + w_offset(3) + kMlpOut * kMlpHidden, 4 * kMlpOut,
                         hipMemcpyDeviceToDevice, h->stream);
}

// Forward of n <= max_m standardised-input rows already in x_rm / x_fm; M = n rounded up to 32.
void forward_core(nml_ctx* h, int n, int M) {
    const uint16_t* in_rm = h->x_rm;
    for (int l = 0; l < 3; ++l) {
        MlpEpi e = {};
        e.mode = kEpiHidden;
        e.m_valid = n;
        e.bias = h->params + w_offset(l) + kIn[l] * kOut[l];
        e.z = h->z[l];
        e.out_rm = h->h_rm[l];
        e.out_fm = h->h_fm[l];
        launch_gemm_nt(h->stream, in_rm, h->wo[l], M, kMlpHidden, kIn[l], e);
        in_rm = h->h_rm[l];
    }
    MlpEpi e = {};
    e.mode = kEpiOutput;
    e.m_valid = n;
    e.bias = h->bias4;
    e.y = h->y;
    launch_gemm_nt(h->stream, in_rm, h->wo[3], M, kMlpOutPad, kMlpHidden, e);
}

// Gradient of the standardised MSE of the last forward into h->grads; returns the loss partial count.
int backward_core(nml_ctx* h, const float* t_dev, int n, int M) {
    const float* nm = h->norm;
    int np = launch_mlp_loss(h->stream, h->y, t_dev, nm + 64, nm + 64 + kMlpOut, n, M, h->d_rm[3], h->d_fm[3],
                             h->partial);
    // layer 4
    {
        MlpEpi e = {};
        e.mode = kEpiStore;
        e.m_valid = kMlpOut;
        e.y = h->grads + w_offset(3);
        launch_gemm_nt(h->stream, h->d_fm[3], h->h_fm[2], kMlpOutPad, kMlpHidden, M, e);
        launch_mlp_bias_grad(h->stream, h->d_fm[3], M, kMlpOut, h->grads + w_offset(3) + kMlpOut * kMlpHidden);
    }
    // hidden layers 3, 2, 1 (index 2, 1, 0): dZ_l from the layer above, then dW_l, db_l
    const uint16_t* up_rm = h->d_rm[3];
    int up_k = kMlpOutPad;
    for (int l = 2; l >= 0; --l) {
        MlpEpi e = {};
        e.mode = kEpiBackHidden;
        e.m_valid = n;
        e.z = h->z[l];
        e.out_rm = h->d_rm[l];
        e.out_fm = h->d_fm[l];
        launch_gemm_nt(h->stream, up_rm, h->wi[l + 1], M, kMlpHidden, up_k, e);
        MlpEpi s = {};
        s.mode = kEpiStore;
        s.m_valid = kMlpHidden;
        s.y = h->grads + w_offset(l);
        const uint16_t* below_fm = l > 0 ? h->h_fm[l - 1] : h->x_fm;
        launch_gemm_nt(h->stream, h->d_fm[l], below_fm, kMlpHidden, kIn[l], M, s);
        launch_mlp_bias_grad(h->stream, h->d_fm[l], M, kMlpHidden, h->grads + w_offset(l) + kIn[l] * kMlpHidden);
        up_rm = h->d_rm[l];
        up_k = kMlpHidden;
    }
    return np;
}

int read_loss(nml_ctx* h, int np, int n, double* loss) {
    if (!loss) {
        NML_HIP(h, hipStreamSynchronize(h->stream));
        return NPS_OK;
    }
    std::vector<double> p(np);
    NML_HIP(h, hipMemcpyAsync(p.data(), h->partial, 8 * (size_t)np, hipMemcpyDeviceToHost, h->stream));
    NML_HIP(h, hipStreamSynchronize(h->stream));
    double s = 0.0;
    for (double v : p) s += v;
    *loss = n > 0 ? s / ((double)n * kMlpOut) : 0.0;
    return NPS_OK;
}

}  // namespace

extern "C" {

int nml_create(int32_t device, int32_t max_batch, nml_handle* out) {
    if (!out || max_batch <= 0) return NPS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NPS_ERR_HIP;
    if (device < 0 || device >= ndev) return NPS_ERR_ARG;
    nml_ctx* h = new nml_ctx();
    h->device = device;
    h->max_m = (max_batch + 31) / 32 * 32;
    auto fail = [&](int code) {
        nml_destroy(h);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(NPS_ERR_HIP);
    auto al = [&](void** p, size_t bytes) {
        if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) return false;
        return hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 256), h->stream) == hipSuccess;
    };
    const size_t M = (size_t)h->max_m;
    bool ok = al((void**)&h->params, 4 * kMlpParams) && al((void**)&h->grads, 4 * kMlpParams) &&
              al((void**)&h->am, 4 * kMlpParams) && al((void**)&h->av, 4 * kMlpParams) &&
              al((void**)&h->bias4, 4 * kMlpOutPad) && al((void**)&h->norm, 4 * 128) &&
              al((void**)&h->x_rm, 2 * M * kMlpIn) && al((void**)&h->x_fm, 2 * M * kMlpIn) &&
              al((void**)&h->y, 4 * M * kMlpOutPad) && al((void**)&h->x_tmp, 4 * M * kMlpIn) &&
              al((void**)&h->t_tmp, 4 * M * kMlpOut) && al((void**)&h->y_tmp, 4 * M * kMlpOut) &&
              al((void**)&h->partial, 8 * (M * kMlpOutPad / 256 + 1)) &&
              al((void**)&h->d_rm[3], 2 * M * kMlpOutPad) && al((void**)&h->d_fm[3], 2 * M * kMlpOutPad);
    for (int l = 0; l < 3 && ok; ++l)
        ok = al((void**)&h->z[l], 4 * M * kMlpHidden) && al((void**)&h->h_rm[l], 2 * M * kMlpHidden) &&
             al((void**)&h->h_fm[l], 2 * M * kMlpHidden) && al((void**)&h->d_rm[l], 2 * M * kMlpHidden) &&
             al((void**)&h->d_fm[l], 2 * M * kMlpHidden);
    for (int l = 0; l < kMlpLayers && ok; ++l)
        ok = al((void**)&h->wo[l], 2 * (size_t)kOutPad[l] * kIn[l]) && al((void**)&h->wi[l], 2 * (size_t)kIn[l] * kOutPad[l]);
    if (!ok) return fail(NPS_ERR_HIP);
    for (hipEvent_t& e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(NPS_ERR_HIP);
    for (int l = 0; l < kMlpLayers; ++l) {
        h->L.w_off[l] = w_offset(l);
        h->L.in[l] = kIn[l];
        h->L.out[l] = kOut[l];
        h->L.out_pad[l] = kOutPad[l];
        h->L.wo[l] = h->wo[l];
        h->L.wi[l] = h->wi[l];
    }
    std::vector<float> nrm(128, 0.0f);
    for (int k = 0; k < 32; ++k) nrm[32 + k] = 1.0f;
    for (int k = 0; k < kMlpOut; ++k) nrm[64 + kMlpOut + k] = 1.0f;
    if (hipMemcpyAsync(h->norm, nrm.data(), 4 * 128, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        return fail(NPS_ERR_HIP);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    *out = h;
    return NPS_OK;
}

void nml_destroy(nml_handle h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* ps[] = {h->params, h->grads, h->am, h->av, h->bias4, h->norm, h->x_rm, h->x_fm, h->y, h->x_tmp,
                  h->t_tmp, h->y_tmp, h->partial, h->d_rm[3], h->d_fm[3]};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (int l = 0; l < 3; ++l) {
        void* q[] = {h->z[l], h->h_rm[l], h->h_fm[l], h->d_rm[l], h->d_fm[l]};
        for (void* p : q)
            if (p) (void)hipFree(p);
    }
    for (int l = 0; l < kMlpLayers; ++l) {
        if (h->wo[l]) (void)hipFree(h->wo[l]);
        if (h->wi[l]) (void)hipFree(h->wi[l]);
    }
    for (hipEvent_t e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* nml_last_error(nml_handle h) { return h ? h->err.c_str() : "null handle"; }

int32_t nml_num_params(void) { return kMlpParams; }

int nml_write_params(nml_handle h, const float* params) {
    if (!h || !params) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    NML_HIP(h, hipMemcpyAsync(h->params, params, 4 * (size_t)kMlpParams, hipMemcpyHostToDevice, h->stream));
    NML_HIP(h, hipMemsetAsync(h->am, 0, 4 * (size_t)kMlpParams, h->stream));
    NML_HIP(h, hipMemsetAsync(h->av, 0, 4 * (size_t)kMlpParams, h->stream));
    h->adam_t = 0.0f;
    refresh(h);
    NML_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nml_read_params(nml_handle h, float* params) {
    if (!h || !params) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    NML_HIP(h, hipMemcpyAsync(params, h->params, 4 * (size_t)kMlpParams, hipMemcpyDeviceToHost, h->stream));
    NML_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nml_set_norm(nml_handle h, const float* in_mean, const float* in_std, const float* out_mean,
                 const float* out_std) {
    if (!h || !in_mean || !in_std || !out_mean || !out_std) return NPS_ERR_ARG;
    std::vector<float> n(128, 0.0f);
    memcpy(&n[0], in_mean, 4 * 32);
    memcpy(&n[32], in_std, 4 * 32);
    memcpy(&n[64], out_mean, 4 * kMlpOut);
    memcpy(&n[64 + kMlpOut], out_std, 4 * kMlpOut);
    NML_HIP(h, hipSetDevice(h->device));
    NML_HIP(h, hipMemcpyAsync(h->norm, n.data(), 4 * 128, hipMemcpyHostToDevice, h->stream));
    NML_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nml_forward_dev(nml_handle h, const float* x_dev, int32_t n, float* y_dev) {
    if (!h || n < 0 || (n > 0 && (!x_dev || !y_dev))) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    // one fused launch over all n rows: activations live in LDS, so the batch
    // capacity does not apply
    MlpInfer P;
    for (int l = 0; l < kMlpLayers; ++l) {
        P.wo[l] = h->wo[l];
        P.bias[l] = l < 3 ? h->params + w_offset(l) + kIn[l] * kOut[l] : h->bias4;
    }
    if (h->timing) NML_HIP(h, hipEventRecord(h->ev[0], h->stream));
    launch_mlp_infer(h->stream, x_dev, n, h->norm, P, y_dev);
    if (h->timing) NML_HIP(h, hipEventRecord(h->ev[1], h->stream));
    NML_HIP(h, hipGetLastError());
    // the caller's next use of y_dev may be on another stream
    NML_HIP(h, hipStreamSynchronize(h->stream));
    if (h->timing) NML_HIP(h, hipEventElapsedTime(&h->timing_ms[0], h->ev[0], h->ev[1]));
    return NPS_OK;
}

int nml_forward(nml_handle h, const float* x, int32_t n, float* y) {
    if (!h || n < 0 || (n > 0 && (!x || !y))) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    for (int32_t b = 0; b < n; b += h->max_m) {
        const int nb = std::min<int32_t>(h->max_m, n - b);
        NML_HIP(h, hipMemcpyAsync(h->x_tmp, x + (size_t)b * kMlpIn, 4 * (size_t)nb * kMlpIn, hipMemcpyHostToDevice,
                                  h->stream));
        int rc = nml_forward_dev(h, h->x_tmp, nb, h->y_tmp);
        if (rc != NPS_OK) return rc;
        NML_HIP(h, hipMemcpyAsync(y + (size_t)b * kMlpOut, h->y_tmp, 4 * (size_t)nb * kMlpOut, hipMemcpyDeviceToHost,
                                  h->stream));
        NML_HIP(h, hipStreamSynchronize(h->stream));
    }
    return NPS_OK;
}

int nml_grads(nml_handle h, const float* x, const float* t, int32_t n, float* grads, double* loss) {
    if (!h || !x || !t || !grads || n <= 0) return NPS_ERR_ARG;
    if (n > h->max_m) return ml_err(h, NPS_ERR_ARG, "batch above the handle's capacity");
    NML_HIP(h, hipSetDevice(h->device));
    const int M = (n + 31) / 32 * 32;
    NML_HIP(h, hipMemcpyAsync(h->x_tmp, x, 4 * (size_t)n * kMlpIn, hipMemcpyHostToDevice, h->stream));
    NML_HIP(h, hipMemcpyAsync(h->t_tmp, t, 4 * (size_t)n * kMlpOut, hipMemcpyHostToDevice, h->stream));
    launch_mlp_input(h->stream, h->x_tmp, h->norm, h->norm + 32, n, M, h->x_rm, h->x_fm);
    forward_core(h, n, M);
    int np = backward_core(h, h->t_tmp, n, M);
    NML_HIP(h, hipGetLastError());
    NML_HIP(h, hipMemcpyAsync(grads, h->grads, 4 * (size_t)kMlpParams, hipMemcpyDeviceToHost, h->stream));
    return read_loss(h, np, n, loss);
}

int nml_train_dev(nml_handle h, const float* x_dev, const float* t_dev, int32_t n, float lr, double* loss) {
    if (!h || !x_dev || !t_dev || n <= 0) return NPS_ERR_ARG;
    if (n > h->max_m) return ml_err(h, NPS_ERR_ARG, "batch above the handle's capacity");
    NML_HIP(h, hipSetDevice(h->device));
    const int M = (n + 31) / 32 * 32;
    if (h->timing) NML_HIP(h, hipEventRecord(h->ev[2], h->stream));
    launch_mlp_input(h->stream, x_dev, h->norm, h->norm + 32, n, M, h->x_rm, h->x_fm);
    forward_core(h, n, M);
    int np = backward_core(h, t_dev, n, M);
    h->adam_t += 1.0f;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)h->adam_t));
    const float bc2 = (float)(1.0 - std::pow(0.999, (double)h->adam_t));
    launch_mlp_adam(h->stream, h->params, h->grads, h->am, h->av, kMlpParams, lr, bc1, bc2, h->L);
    NML_HIP(h, hipMemcpyAsync(h->bias4, h->params + w_offset(3) + kMlpOut * kMlpHidden, 4 * kMlpOut,
                              hipMemcpyDeviceToDevice, h->stream));
    if (h->timing) NML_HIP(h, hipEventRecord(h->ev[3], h->stream));
    NML_HIP(h, hipGetLastError());
    const int rc = read_loss(h, np, n, loss);
    if (rc == NPS_OK && h->timing) NML_HIP(h, hipEventElapsedTime(&h->timing_ms[1], h->ev[2], h->ev[3]));
    return rc;
}

int nml_train(nml_handle h, const float* x, const float* t, int32_t n, float lr, double* loss) {
    if (!h || !x || !t || n <= 0) return NPS_ERR_ARG;
    if (n > h->max_m) return ml_err(h, NPS_ERR_ARG, "batch above the handle's capacity");
    NML_HIP(h, hipSetDevice(h->device));
    NML_HIP(h, hipMemcpyAsync(h->x_tmp, x, 4 * (size_t)n * kMlpIn, hipMemcpyHostToDevice, h->stream));
    NML_HIP(h, hipMemcpyAsync(h->t_tmp, t, 4 * (size_t)n * kMlpOut, hipMemcpyHostToDevice, h->stream));
    return nml_train_dev(h, h->x_tmp, h->t_tmp, n, lr, loss);
}

int nml_enable_timing(nml_handle h, int32_t on) {
    if (!h) return NPS_ERR_ARG;
    h->timing = on != 0;
    return NPS_OK;
}

int nml_read_timing(nml_handle h, float ms[2]) {
    if (!h || !ms) return NPS_ERR_ARG;
    ms[0] = h->timing_ms[0];
    ms[1] = h->timing_ms[1];
    return NPS_OK;
}

int nml_dev_alloc(nml_handle h, int64_t bytes, void** out) {
    if (!h || !out || bytes < 0) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    *out = nullptr;
    NML_HIP(h, hipMalloc(out, (size_t)std::max<int64_t>(bytes, 1)));
    return NPS_OK;
}

int nml_dev_free(nml_handle h, void* p) {
    if (!h) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    NML_HIP(h, hipStreamSynchronize(h->stream));
    if (p) NML_HIP(h, hipFree(p));
    return NPS_OK;
}

int nml_dev_upload(nml_handle h, void* dst, const void* src, int64_t bytes) {
    if (!h || !dst || !src || bytes < 0) return NPS_ERR_ARG;
    NML_HIP(h, hipSetDevice(h->device));
    NML_HIP(h, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, h->stream));
    NML_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

}  // extern "C"
