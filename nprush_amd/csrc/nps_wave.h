// nps_wave.h -- wave-wide integer transpose reductions (gfx950 wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nps {

// Wave-wide sum of 16 values; lane l ends with the total of value
// ((l>>5)&1)*8 + ((l>>4)&1)*4 + ((l>>3)&1)*2 + ((l>>2)&1).
// Transpose reduction: each step pairs lanes that differ in one "keep" bit, each
// lane keeps half of its values and adds its partner's copy of that half.
//   16 -> 8  v_permlane32_swap (half-waves), keep bit 5
//    8 -> 4  v_permlane16_swap (odd/even rows), keep bit 4
//    4 -> 2  DPP row_mirror (lane ^ 15 in a row), keep bit 3
//    2 -> 1  DPP row_half_mirror (lane ^ 7), keep bit 2
//    then DPP quad_perm sums over lane bits 0 and 1.
// Integer adds, so the totals do not depend on the pairing.
__device__ __forceinline__ uint32_t wave_reduce16(uint32_t (&v)[16], int lane) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        auto r = __builtin_amdgcn_permlane32_swap(v[k], v[k + 8], false, false);
        v[k] = r[0] + r[1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        auto r = __builtin_amdgcn_permlane16_swap(v[k], v[k + 4], false, false);
        v[k] = r[0] + r[1];
    }
    {
        const bool hi = lane & 8;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            uint32_t send = hi ? v[k] : v[k + 2];
            uint32_t keep = hi ? v[k + 2] : v[k];
            v[k] = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x140, 0xF, 0xF, false);
        }
    }
    {
        const bool hi = lane & 4;
        uint32_t send = hi ? v[0] : v[1];
        uint32_t keep = hi ? v[1] : v[0];
        v[0] = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x141, 0xF, 0xF, false);
    }
    v[0] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], 0xB1, 0xF, 0xF, false);
    v[0] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], 0x4E, 0xF, 0xF, false);
    return v[0];
}

// Wave-wide sum of 8 values; lane l ends with the total of value
// ((l>>5)&1)*4 + ((l>>4)&1)*2 + ((l>>3)&1).  Same transpose steps as
// wave_reduce16 for lane bits 5, 4 and 3, then sums over lane bits 0-2
// (quad_perm lane^1, lane^2, then row_half_mirror pairs quad 0 with quad 1).
__device__ __forceinline__ uint32_t wave_reduce8(uint32_t (&v)[8], int lane) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        auto r = __builtin_amdgcn_permlane32_swap(v[k], v[k + 4], false, false);
        v[k] = r[0] + r[1];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        auto r = __builtin_amdgcn_permlane16_swap(v[k], v[k + 2], false, false);
        v[k] = r[0] + r[1];
    }
    {
        const bool hi = lane & 8;
        uint32_t send = hi ? v[0] : v[1];
        uint32_t keep = hi ? v[1] : v[0];
        v[0] = keep + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send, 0x140, 0xF, 0xF, false);
    }
    v[0] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], 0xB1, 0xF, 0xF, false);
    v[0] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], 0x4E, 0xF, 0xF, false);
    v[0] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], 0x141, 0xF, 0xF, false);
    return v[0];
}

}  // namespace nps
