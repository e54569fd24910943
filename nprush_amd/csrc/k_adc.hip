// k_adc.hip -- adaptive density control as parallel stream compaction, with
// the exact result of the reference's single-thread scan (splat_adc.wgsl:27-87).
//
// Serial semantics being reproduced: for i = 0..N-1, a splat alive at its visit
// (opacity >= 0.05) resets grad_accum[i]; if accum/25 > 5e-5 it takes the next
// dead slot at or after a forward-only cursor and writes a split/clone child
// there.  Because the cursor only advances and a dead slot becomes alive only
// when it is taken, the k-th qualifying parent (in visit order) receives the
// k-th initially-dead slot D[k].  A dead slot D[k] holds a child when the loop
// reaches it iff K(D[k]) > k, where K(i) = number of qualifiers before i; such a
// child is visited (its stale grad_accum is tested, so it can spawn in turn).
// Parallel form:
//   1. flags: alive0, qa = accum/25 > tau, dead; D = compaction of dead slots.
//   2. Q = alive0 & qa; repeat { K = scan(Q); Q |= dead & qa & (K > rank) }
//      until nothing changes.  Q only grows and every addition is forced by
//      the current K, so the fixed point is the serial qualifier set.
//   3. qualifier with rank k < |D| -> parent of slot D[k].  Its value at visit
//      is its original value, or (for a child visited alive) its parent's
//      child value; chains are followed upwards and replayed downwards.
//   4. results go to a temporary buffer first, then are written, so no thread
//      reads a splat another thread has already replaced.
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

constexpr float kAdcPeriod = 25.0f;     // SPLAT_ADC_PERIOD (constants.ts:30)
constexpr float kTau = 0.00005f;        // SPLAT_GRAD_THRESH (constants.ts:31)
constexpr float kKill = 0.05f;          // SPLAT_OPACITY_KILL_THRESH (constants.ts:32)
constexpr float kSplitScale = 0.01f;    // splat_adc.wgsl:32
constexpr int kMaxChain = 256;

__global__ void k_adc_flags(const float4* __restrict__ splats, const float* __restrict__ accum, uint32_t n,
                            uint32_t* alive0, uint32_t* qa, uint32_t* dead, uint32_t* Q) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float op = splats[4 * (size_t)i + 1].w;
    bool is_dead = op < kKill;
    float ga = accum[i] / kAdcPeriod;
    bool q = ga > kTau;
    alive0[i] = is_dead ? 0u : 1u;
    dead[i] = is_dead ? 1u : 0u;
    qa[i] = q ? 1u : 0u;
    Q[i] = (!is_dead && q) ? 1u : 0u;
}

__global__ void k_adc_dead_list(const uint32_t* dead, const uint32_t* dead_rank, uint32_t n, uint32_t* dead_list) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dead[i]) dead_list[dead_rank[i]] = i;
}

__global__ void k_adc_cascade(const uint32_t* dead, const uint32_t* qa, const uint32_t* dead_rank,
                              const uint32_t* K, uint32_t n, uint32_t* Q, uint32_t* flag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dead[i] && qa[i] && !Q[i] && K[i] > dead_rank[i]) {
        Q[i] = 1u;
        flag[0] = 1u;
    }
}

__global__ void k_adc_qual_list(const uint32_t* Q, const uint32_t* K, uint32_t n, uint32_t num_dead,
                                uint32_t* qual_list) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (Q[i] && K[i] < num_dead) qual_list[K[i]] = i;
}

struct S16 {
    float v[16];
};

__device__ __forceinline__ S16 load16(const float4* p, uint32_t i) {
    S16 s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float4 a = p[4 * (size_t)i + q];
        s.v[4 * q] = a.x; s.v[4 * q + 1] = a.y; s.v[4 * q + 2] = a.z; s.v[4 * q + 3] = a.w;
    }
    return s;
}

// splat_adc.wgsl:59-73 applied to the visit value `s` of parent i: updates the
// parent in place and returns the child.  *split reports the branch.
__device__ __forceinline__ S16 adc_transform(S16& s, uint32_t i, float adam_t, bool* split) {
    float scale_norm = sqrtf((s.v[3] * s.v[3] + s.v[12] * s.v[12]) + s.v[15] * s.v[15]);
    S16 c = s;
    if (scale_norm > kSplitScale) {
        const float f = 0.625f;
        s.v[3] = s.v[3] * f; s.v[12] = s.v[12] * f; s.v[15] = s.v[15] * f;
        c.v[3] = c.v[3] * f; c.v[12] = c.v[12] * f; c.v[15] = c.v[15] * f;
        s.v[0] = s.v[0] - s.v[3] * 0.05f;
        c.v[0] = c.v[0] + s.v[3] * 0.05f;
        *split = true;
    } else {
        float seed = (float)i * 3.14159f + adam_t;
        c.v[0] = c.v[0] + (fract_det(sin_det(seed * 12.9898f) * 43758.5453f) - 0.5f) * 0.002f;
        c.v[1] = c.v[1] + (fract_det(sin_det(seed * 78.233f) * 43758.5453f) - 0.5f) * 0.002f;
        *split = false;
    }
    return c;
}

// One thread per assignment k: parent qual_list[k] -> slot dead_list[k].
constexpr int kChainLds = 32;

__global__ __launch_bounds__(256) void k_adc_compute(const float4* __restrict__ splats, const uint32_t* alive0,
                                                     const uint32_t* dead_rank, const uint32_t* qual_list,
                                                     uint32_t num_assign, float adam_t, float* tmp, uint32_t* flag) {
    __shared__ uint32_t s_chain[kChainLds][256];
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= num_assign) return;
    // chain(0) = qual_list[k]; while chain(d) was dead at the scan, it is the
    // dead slot of assignment dead_rank[chain(d)], whose parent is chain(d+1).
    // The first kChainLds entries are kept in LDS ([d][thread]: no bank
    // conflicts), deeper ones re-walked from there (a per-thread array indexed
    // at run time would live in scratch).
    auto chain = [&](int d) {
        if (d < kChainLds) return s_chain[d][threadIdx.x];
        uint32_t c = s_chain[kChainLds - 1][threadIdx.x];
        for (int t = kChainLds - 1; t < d; ++t) c = qual_list[dead_rank[c]];
        return c;
    };
    int depth = 1;
    uint32_t cur = qual_list[k];
    s_chain[0][threadIdx.x] = cur;
    while (!alive0[cur]) {
        cur = qual_list[dead_rank[cur]];
        if (depth >= kMaxChain) {
            flag[1] = 1u;
            return;
        }
        if (depth < kChainLds) s_chain[depth][threadIdx.x] = cur;
        ++depth;
    }
    // replay from the root (cur = chain(depth - 1)): visit value of chain(d-1) is original
    S16 s = load16(splats, cur);
    bool split = false;
    for (int d = depth - 1; d >= 1; --d) {
        S16 c = adc_transform(s, chain(d), adam_t, &split);
        s = c;  // child value = visit value of chain(d-1)
    }
    S16 child = adc_transform(s, chain(0), adam_t, &split);
    float* o = tmp + 32 * (size_t)k;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        o[q] = s.v[q];
        o[16 + q] = child.v[q];
    }
    wave_count2(split, &flag[2], &flag[3]);
}

__global__ void k_adc_write(float4* __restrict__ splats, float4* __restrict__ m, float4* __restrict__ v,
                            const uint32_t* Q, const uint32_t* K, uint32_t num_dead, const uint32_t* qual_list,
                            const uint32_t* dead_list, uint32_t num_assign, const float* tmp) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= num_assign) return;
    uint32_t i = qual_list[k], slot = dead_list[k];
    const float* o = tmp + 32 * (size_t)k;
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        splats[4 * (size_t)i + q] = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
        m[4 * (size_t)i + q] = z; v[4 * (size_t)i + q] = z;
        m[4 * (size_t)slot + q] = z; v[4 * (size_t)slot + q] = z;
    }
    // The slot's final value is its own parent result when it later qualified
    // with a free slot of its own; that thread writes it.
    bool slot_is_parent = Q[slot] && K[slot] < num_dead;
    if (!slot_is_parent) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            splats[4 * (size_t)slot + q] =
                make_float4(o[16 + 4 * q], o[16 + 4 * q + 1], o[16 + 4 * q + 2], o[16 + 4 * q + 3]);
    }
}

// grad_accum[i] = 0 for every splat alive at its visit.
__global__ void k_adc_reset(float* accum, const uint32_t* alive0, const uint32_t* dead, const uint32_t* dead_rank,
                            const uint32_t* K, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool visited_alive = alive0[i] || (dead[i] && K[i] > dead_rank[i]);
    if (visited_alive) accum[i] = 0.0f;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

int run_adc_ref(hipStream_t st, float* splats, float* m, float* v, float* accum, uint32_t n, float adam_t,
                const AdcBuffers& b, int64_t stats[5]) {
    for (int k = 0; k < 5; ++k) stats[k] = 0;
    if (n == 0) return 0;
    dim3 grid((n + 255) / 256), blk(256);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, b.flag, 4u);
    hipLaunchKernelGGL(k_adc_flags, grid, blk, 0, st, (const float4*)splats, (const float*)accum, n, b.alive0, b.qa,
                       b.dead, b.Q);
    exclusive_scan_u32(st, b.dead, b.dead_rank, n, b.scan_tmp);
    hipLaunchKernelGGL(k_adc_dead_list, grid, blk, 0, st, (const uint32_t*)b.dead, (const uint32_t*)b.dead_rank, n,
                       b.dead_list);
    uint32_t num_dead = 0;
    hipMemcpyAsync(&num_dead, b.dead_rank + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    // Cascade rounds until one changes nothing.  A round only turns qualifier
    // flags on, so a round after convergence is a no-op: rounds run in groups
    // of kCascadeGroup per host readback of the last round's flag.
    constexpr int kCascadeGroup = 4;
    int rounds = 0;
    for (;;) {
        for (int g = 0; g < kCascadeGroup; ++g) {
            exclusive_scan_u32(st, b.Q, b.K, n, b.scan_tmp);
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, b.flag, 1u);
            hipLaunchKernelGGL(k_adc_cascade, grid, blk, 0, st, (const uint32_t*)b.dead, (const uint32_t*)b.qa,
                               (const uint32_t*)b.dead_rank, (const uint32_t*)b.K, n, b.Q, b.flag);
        }
        uint32_t changed = 0;
        hipMemcpyAsync(&changed, b.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        rounds += kCascadeGroup;
        if (!changed) break;
    }
    uint32_t total_q = 0;
    hipMemcpyAsync(&total_q, b.K + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    uint32_t num_assign = total_q < num_dead ? total_q : num_dead;
    if (num_assign > 0) {
        hipLaunchKernelGGL(k_adc_qual_list, grid, blk, 0, st, (const uint32_t*)b.Q, (const uint32_t*)b.K, n, num_dead,
                           b.qual_list);
        dim3 ga((num_assign + 255) / 256);
        hipLaunchKernelGGL(k_adc_compute, ga, blk, 0, st, (const float4*)splats, (const uint32_t*)b.alive0,
                           (const uint32_t*)b.dead_rank, (const uint32_t*)b.qual_list, num_assign, adam_t, b.tmp,
                           b.flag);
        hipLaunchKernelGGL(k_adc_write, ga, blk, 0, st, (float4*)splats, (float4*)m, (float4*)v,
                           (const uint32_t*)b.Q, (const uint32_t*)b.K, num_dead, (const uint32_t*)b.qual_list,
                           (const uint32_t*)b.dead_list, num_assign, (const float*)b.tmp);
    }
    hipLaunchKernelGGL(k_adc_reset, grid, blk, 0, st, accum, (const uint32_t*)b.alive0, (const uint32_t*)b.dead,
                       (const uint32_t*)b.dead_rank, (const uint32_t*)b.K, n);
    uint32_t fl[4] = {0, 0, 0, 0};
    hipMemcpyAsync(fl, b.flag, sizeof(fl), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    stats[0] = total_q;
    stats[1] = num_assign;
    stats[2] = fl[2];
    stats[3] = fl[3];
    stats[4] = rounds;
    return fl[1] ? -1 : 0;
}

}  // namespace nps
