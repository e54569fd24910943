// k_debug.hip -- evaluates the deterministic transcendentals of nps_math.h on
// the GPU so tests can compare them bit for bit with the CPU oracle.
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

__global__ void k_debug_math(int op, const float* x, const float* y, float* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = x[i];
    float r;
    switch (op) {
        case 0: r = exp2_det(a); break;
        case 1: r = log2_det(a); break;
        case 2: r = pow_det(a, y[i]); break;
        case 3: r = exp_det(a); break;
        case 4: r = log_det(a); break;
        case 6: r = __int_as_float(f2i_hw(a)); break;
        case 5: r = sin_det(a); break;
        case 7: r = cos_det(a); break;
        case 8: r = atan2_det(a, y[i]); break;
        case 9: r = asin_det(a); break;
        default: r = 0.0f; break;
    }
    out[i] = r;
}

void launch_debug_math(hipStream_t st, int op, const float* x, const float* y, float* out, uint32_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_debug_math, dim3((n + 255) / 256), dim3(256), 0, st, op, x, y, out, n);
}

}  // namespace nps
