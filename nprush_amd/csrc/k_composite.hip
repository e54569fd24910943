// k_composite.hip -- fused per-tile forward composite, L2 loss seeds and
// reverse pass with fixed-point gradients (REF semantics of
// splat_backward.wgsl:213-455).
//
// One 256-thread workgroup per 16x16 tile, one thread per pixel, each wave an
// 8x8 pixel block.  The tile list (ascending splat ids, first `cap` of them)
// is walked in batches of 64 splat records staged in LDS.
//   forward : per pixel front-to-back alpha compositing.  Every 8 entries the
//             transmittance T is checkpointed ([tile][chunk][pixel], coalesced),
//             and each wave records one bit per entry: does any of its 64 pixels
//             reach alpha >= 0.001 (the reference's reverse-pass condition,
//             BW:263)?  Pairs with r^2 beyond the splat's cut radius skip
//             pow/exp: their alpha is provably exactly 0.
//   reverse : batches and 8-entry chunks in reverse order.  Chunks without an
//             active bit are skipped by the whole wave.  Otherwise alpha, T,
//             log2 r and r^a are recomputed from the checkpoint into registers up
//             to the last active entry (same expressions, same bits as the
//             forward), and the active entries are walked backwards.  The
//             reference keeps alphas[1024]/Ts[1025] per thread instead.
//   grads   : per pair the 16 gradients are converted to i32 exactly as the
//             reference's atomicAdd(i32(g * FP)) does; a wave sums them with a
//             permlane-swap/DPP transpose reduction (integer adds, so the sum
//             equals the reference's order-independent atomics), 16 lanes add the
//             totals into an LDS accumulator per batch, and each batch ends with
//             one global atomic per non-zero (splat, parameter).
#include "nps_common.h"
#include "nps_kernels.h"
#include "nps_wave.h"

namespace nps {

__device__ __forceinline__ uint32_t gi(float g, float fp) { return (uint32_t)f2i_hw(g * fp); }

struct PixelState {
    float dCr, dCg, dCb, dD;
    float dT_C, dT_D;
    float sx, sy, sz;  // surface point for the 3D pull
    bool fg;
};

// Reverse-pass body for one (pixel, splat) pair: splat_backward.wgsl:264-455
// with the per-splat factors taken from the records.  Writes 16 i32 values.
__device__ __forceinline__ void pair_grads(const FwdRec& f, const BwdRec& b, const PairR2& g, float log2r,
                                           float rpa, float a, float T_prev, PixelState& st, uint32_t (&out)[16]) {
    const float cr = f.r, cg = f.g, cb = f.b, opacity = f.op;
    float Ta = T_prev * a;
    float dColor_r = st.dCr * Ta, dColor_g = st.dCg * Ta, dColor_b = st.dCb * Ta;
    float depth = f.depth;
    float da_C = st.dT_C * (-T_prev) + ((st.dCr * (T_prev * cr) + st.dCg * (T_prev * cg)) + st.dCb * (T_prev * cb));
    float da_D = st.dT_D * (-T_prev) + (st.dD * T_prev) * depth;
    float da = da_C + da_D;
    st.dT_C = st.dT_C * (1.0f - a) + ((st.dCr * (a * cr) + st.dCg * (a * cg)) + st.dCb * (a * cb));
    st.dT_D = st.dT_D * (1.0f - a) + (st.dD * a) * depth;
    float d_depth = (st.dD * T_prev) * a;
    const float sx = b.s3[0], sy = b.s3[1], sz = b.s3[2];
    const float shape_a = f.sa, shape_b = f.sb;
    float power = -shape_b * rpa;
    bool above = power > -15.0f;
    float a_un = above ? exp_det(power) : 0.0f;
    float d_opacity = above ? da * a_un : 0.0f;
    float d_power = above ? (da * opacity) * a_un : 0.0f;
    float r_pow_a = rpa;
    float r_pow_a_m2 = exp2_det((shape_a - 2.0f) * log2r);
    float log_r = log2r * kLn2;
    float d_shape_a = above ? d_power * ((-shape_b * r_pow_a) * log_r) : 0.0f;
    float d_shape_b = above ? d_power * (-r_pow_a) : 0.0f;
    float kk = d_power * ((-shape_b * shape_a) * r_pow_a_m2);
    float d_dx = above ? kk * g.spx : 0.0f;
    float d_dy = above ? kk * g.spy : 0.0f;
    float d_screen_x = -d_dx, d_screen_y = -d_dy;

    float sp_dot_m0 = g.spx * b.m[0] + g.spy * b.m[1];
    float sp_dot_m1 = g.spx * b.m[2] + g.spy * b.m[3];
    float sp_dot_m2 = g.spx * b.m[4] + g.spy * b.m[5];
    float d_m0x = -sp_dot_m0 * d_dx, d_m0y = -sp_dot_m0 * d_dy;
    float d_m1x = -sp_dot_m1 * d_dx, d_m1y = -sp_dot_m1 * d_dy;
    float d_m2x = -sp_dot_m2 * d_dx, d_m2y = -sp_dot_m2 * d_dy;
    float d_axs_x = d_m0x * sx, d_axs_y = d_m0y * sx;
    float d_ays_x = d_m1x * sy, d_ays_y = d_m1y * sy;
    float d_azs_x = d_m2x * sz, d_azs_y = d_m2y * sz;
    float d_sx = d_m0x * b.axs[0] + d_m0y * b.axs[1];
    float d_sy = d_m1x * b.axs[2] + d_m1y * b.axs[3];
    float d_sz = d_m2x * b.axs[4] + d_m2y * b.axs[5];

    bool w_ok = b.w > 0.1f;
    float d_pos[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float dp = d_screen_x * b.J0[ax] + d_screen_y * b.J1[ax];
        d_pos[ax] = dp + (w_ok ? b.vpd[ax] * d_depth : 0.0f);
    }
    if (st.fg) {
        const float wgt = 0.3f * 2.0f;
        d_pos[0] = d_pos[0] + (wgt * (b.pos[0] - st.sx)) * Ta;
        d_pos[1] = d_pos[1] + (wgt * (b.pos[1] - st.sy)) * Ta;
        d_pos[2] = d_pos[2] + (wgt * (b.pos[2] - st.sz)) * Ta;
    }
    float axw[3], ayw[3], azw[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        axw[k] = d_axs_x * b.J0[k] + d_axs_y * b.J1[k];
        ayw[k] = d_ays_x * b.J0[k] + d_ays_y * b.J1[k];
        azw[k] = d_azs_x * b.J0[k] + d_azs_y * b.J1[k];
    }
    const float qw = b.q[0], qx = b.q[1], qy = b.q[2], qz = b.q[3];
    const float T2 = 2.0f, T4 = 4.0f;
    float d_qw_ax = (axw[1] * T2) * qz - (axw[2] * T2) * qy;
    float d_qw_ay = ((-ayw[0]) * T2) * qz + (ayw[2] * T2) * qx;
    float d_qw_az = (azw[0] * T2) * qy - (azw[1] * T2) * qx;
    float d_qx_ax = (axw[1] * T2) * qy + (axw[2] * T2) * qz;
    float d_qx_ay = ((ayw[0] * T2) * qy - (ayw[1] * T4) * qx) + (ayw[2] * T2) * qw;
    float d_qx_az = ((azw[0] * T2) * qz - (azw[1] * T2) * qw) - (azw[2] * T4) * qx;
    float d_qy_ax = (((-axw[0]) * T4) * qy + (axw[1] * T2) * qx) + axw[2] * (-T2 * qw);
    float d_qy_ay = (ayw[0] * T2) * qx + (ayw[2] * T2) * qz;
    float d_qy_az = ((azw[0] * T2) * qw + (azw[1] * T2) * qz) - (azw[2] * T4) * qy;
    float d_qz_ax = (((-axw[0]) * T4) * qz + (axw[1] * T2) * qw) + (axw[2] * T2) * qx;
    float d_qz_ay = (((-ayw[0]) * T2) * qw - (ayw[1] * T4) * qz) + (ayw[2] * T2) * qy;
    float d_qz_az = (azw[0] * T2) * qx + (azw[1] * T2) * qy;
    float d_qw = (d_qw_ax + d_qw_ay) + d_qw_az;
    float d_qx = (d_qx_ax + d_qx_ay) + d_qx_az;
    float d_qy = (d_qy_ax + d_qy_ay) + d_qy_az;
    float d_qz = (d_qz_ax + d_qz_ay) + d_qz_az;
    const float reg = 0.005f * 2.0f;
    d_shape_a = d_shape_a + reg * (shape_a - 6.0f);
    d_shape_b = d_shape_b + reg * (shape_b - 0.3f);

    const float FPP = 10000.0f, FPC = 100000.0f;
    out[0] = gi(d_pos[0], FPP);
    out[1] = gi(d_pos[1], FPP);
    out[2] = gi(d_pos[2], FPP);
    out[3] = gi(d_sx, FPP);
    out[4] = gi(dColor_r, FPC);
    out[5] = gi(dColor_g, FPC);
    out[6] = gi(dColor_b, FPC);
    out[7] = gi(d_opacity, FPC);
    out[8] = gi(d_qw, FPP);
    out[9] = gi(d_qx, FPP);
    out[10] = gi(d_qy, FPP);
    out[11] = gi(d_qz, FPP);
    out[12] = gi(d_sy, FPP);
    out[13] = gi(d_shape_a, FPP);
    out[14] = gi(d_shape_b, FPP);
    out[15] = gi(d_sz, FPP);
}

// alpha of a pair from its r^2 (BW:232-237); log2 r and r^a are returned for
// the reverse pass.  Beyond rcut2 the float power is <= -15, so alpha is 0.
__device__ __forceinline__ float alpha_of(const FwdRec& f, float r2, float& log2r, float& rpa) {
    if (r2 > f.rcut2) {
        log2r = 0.0f;
        rpa = 0.0f;
        return 0.0f;
    }
    log2r = log2_sqrt_det(fmax_w(r2, 0.0001f));  // log2 of r = sqrt(max(r2, 1e-4))
    rpa = exp2_det(f.sa * log2r);                 // pow(r, shape_a)
    float power = -f.sb * rpa;
    float a = (power > -15.0f) ? exp_det(power) * f.op : 0.0f;
    return clamp_w(a, 0.0f, 0.999f);
}

__device__ __forceinline__ void stage_records(const FwdRec* __restrict__ fwd, const BwdRec* __restrict__ bwd,
                                              const uint64_t* __restrict__ keys, uint32_t base, int nb,
                                              FwdRec* s_fwd, BwdRec* s_bwd, uint32_t* s_ids, bool with_bwd) {
    const int t = threadIdx.x;
    {
        int j = t >> 2, part = t & 3;
        if (j < nb) {
            uint32_t id = (uint32_t)keys[base + j];
            reinterpret_cast<float4*>(&s_fwd[j])[part] = reinterpret_cast<const float4*>(&fwd[id])[part];
            if (part == 0) s_ids[j] = id;
        }
    }
    if (with_bwd) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int e = t + h * 256;
            int j = e >> 3, part = e & 7;
            if (j < nb) {
                uint32_t id = (uint32_t)keys[base + j];
                reinterpret_cast<float4*>(&s_bwd[j])[part] = reinterpret_cast<const float4*>(&bwd[id])[part];
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_composite_ref(ViewParams v, CompositeArgs A) {
    __shared__ FwdRec s_fwd[kBatch];
    __shared__ BwdRec s_bwd[kBatch];
    __shared__ uint32_t s_ids[kBatch];
    __shared__ uint32_t s_grad[kBatch][16];
    __shared__ uint32_t s_act[4][kMaxCap / 32];
    __shared__ double s_red[2][4];
    __shared__ uint32_t s_cnt[3][4];

    const uint32_t tile = xcd_remap(blockIdx.x, v.num_tiles);
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t start = A.ranges[2 * tile], end = A.ranges[2 * tile + 1];
    const uint32_t len = end - start;
    const int count = (int)min(len, A.cap);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t px = tx * kTile + (wid & 1) * 8 + (lane & 7);
    const uint32_t py = ty * kTile + (wid >> 1) * 8 + (lane >> 3);
    const bool valid = px < v.W && py < v.H;
    const float pxf = px_to_x(px, v.W, v.aspect);
    const float pyf = py_to_y(py, v.H);
    float* ckpt = A.ckpt + (size_t)tile * A.nchunk_max * kTilePx;

    // ---------------- forward (BW:216-248) ----------------
    // An entry whose cut radius misses every pixel centre of the wave's 8x8
    // block has alpha exactly 0 there (T * (1 - 0) = T, the sums gain +0): lane j
    // of each wave tests entry j of the batch (conic_reaches_rect against rcut2)
    // and the wave walks only the entries that can reach it, in order.  The T
    // checkpoint of chunk c (entries 8c..8c+7) is written when the walk first
    // reaches the chunk (T is unchanged over the skipped entries); activity words
    // (alpha >= 0.001 on some pixel of the wave) per 32 entries as before.
    float T = 1.0f, Cr = 0.0f, Cg = 0.0f, Cb = 0.0f, D = 0.0f;
    uint32_t n_feval = 0u;  // wave-entries that evaluate pow/exp in the forward
    int ck_next = 0;        // first chunk whose checkpoint is not written (uniform)
    const uint32_t bx0 = tx * kTile + (wid & 1) * 8, by0 = ty * kTile + (wid >> 1) * 8;
    const float blk_x0 = px_to_x(bx0, v.W, v.aspect), blk_x1 = px_to_x(bx0 + 7u, v.W, v.aspect);
    const float blk_y0 = py_to_y(by0 + 7u, v.H), blk_y1 = py_to_y(by0, v.H);
    const int ngroups = (count + 31) >> 5;
    static_assert(kBatch == 64, "one 64-bit hit mask per batch");
    for (int b0 = 0; b0 < count; b0 += kBatch) {
        int nb = min(kBatch, count - b0);
        __syncthreads();
        stage_records(A.fwd, A.bwd, A.keys, start + (uint32_t)b0, nb, s_fwd, s_bwd, s_ids, false);
        __syncthreads();
        bool h = false;
        if (lane < nb) {
            const FwdRec& f = s_fwd[lane];
            h = conic_reaches_rect(f.A, f.B, f.C, blk_x0 - f.scr_x, blk_x1 - f.scr_x, blk_y0 - f.scr_y,
                                   blk_y1 - f.scr_y, f.rcut2);
        }
        uint64_t hits = uniform_u64(__ballot(h));
        uint64_t act = 0ull;
        while (hits != 0ull) {
            const int j = (int)__builtin_ctzll(hits);
            hits &= hits - 1ull;
            const int e = b0 + j;
            for (; ck_next <= (e >> 3); ++ck_next) ckpt[(size_t)ck_next * kTilePx + tid] = T;
            const FwdRec& f = s_fwd[j];
            PairR2 q = pair_r2(f, pxf, pyf);
            float l2r, rpa;
            n_feval += (uint32_t)(__ballot(valid && !(q.r2 > f.rcut2)) != 0ull);
            float a = alpha_of(f, q.r2, l2r, rpa);
            act |= (__ballot(valid && !(a < 0.001f)) != 0ull ? 1ull : 0ull) << j;
            float Ta = T * a;
            Cr = Cr + Ta * f.r;
            Cg = Cg + Ta * f.g;
            Cb = Cb + Ta * f.b;
            D = D + Ta * f.depth;
            T = T * (1.0f - a);
        }
        if (lane == 0) {
            s_act[wid][b0 >> 5] = (uint32_t)act;
            if ((b0 >> 5) + 1 < ngroups) s_act[wid][(b0 >> 5) + 1] = (uint32_t)(act >> 32);
        }
    }
    for (; ck_next < (count + kChunk - 1) / kChunk; ++ck_next) ckpt[(size_t)ck_next * kTilePx + tid] = T;
    const float bg = 0.05f;
    Cr = Cr + T * bg;
    Cg = Cg + T * bg;
    Cb = Cb + T * bg;
    D = D + T * 1.0f;

    // ---------------- loss seeds (BW:250-257) ----------------
    PixelState st;
    float4 tg = make_float4(0.f, 0.f, 0.f, 0.f);
    float tdepth = 1.0f;
    if (valid) {
        size_t pix = (size_t)py * v.W + px;
        tg = A.tgt_rgba[pix];
        tdepth = A.tgt_depth[pix];
        A.c_img[3 * pix + 0] = Cr;
        A.c_img[3 * pix + 1] = Cg;
        A.c_img[3 * pix + 2] = Cb;
        A.d_img[pix] = D;
    }
    st.dCr = 2.0f * (Cr - tg.x);
    st.dCg = 2.0f * (Cg - tg.y);
    st.dCb = 2.0f * (Cb - tg.z);
    st.fg = tdepth < 0.99f;
    st.dD = st.fg ? 2.0f * (D - tdepth) : 0.0f;
    double l2 = 0.0, l1 = 0.0;
    if (valid) {
        double er = (double)Cr - (double)tg.x, eg = (double)Cg - (double)tg.y, eb = (double)Cb - (double)tg.z;
        double ed = st.fg ? (double)D - (double)tdepth : 0.0;
        l2 = er * er + eg * eg + eb * eb + ed * ed;
        l1 = fabs(er) + fabs(eg) + fabs(eb);
    }
    st.dT_C = (st.dCr * bg + st.dCg * bg) + st.dCb * bg;
    st.dT_D = st.dD * 1.0f;
    st.sx = st.sy = st.sz = 0.0f;
    if (st.fg) {
        // reconstruct_world_pos (BW:64-75)
        float w = 0.1f / fmax_w(1.0f - tdepth, 1e-5f);
        float ndc_x = pxf / v.aspect;
        float ndc_y = pyf;
        F4 world = mat_vec(v.vp_inv, ndc_x * w, ndc_y * w, w, w);
        st.sx = world.x / world.w;
        st.sy = world.y / world.w;
        st.sz = world.z / world.w;
    }
    // per-tile loss partial (fixed reduction tree: deterministic)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        l2 += __shfl_xor(l2, o, 64);
        l1 += __shfl_xor(l1, o, 64);
    }
    if (lane == 0) {
        s_red[0][wid] = l2;
        s_red[1][wid] = l1;
    }

    // ---------------- reverse pass (BW:259-455) ----------------
    const int nbatch = (count + kBatch - 1) / kBatch;
    uint32_t n_active = 0;  // wave-uniform count of pairs taken by the reverse pass
    uint32_t n_went = 0;    // wave-entries walked by the reverse pass
    for (int bi = nbatch - 1; bi >= 0; --bi) {
        const int b0 = bi * kBatch;
        const int nb = min(kBatch, count - b0);
        __syncthreads();
        stage_records(A.fwd, A.bwd, A.keys, start + (uint32_t)b0, nb, s_fwd, s_bwd, s_ids, true);
        for (int e = tid; e < kBatch * 16; e += 256) (&s_grad[0][0])[e] = 0u;
        __syncthreads();
        const int nch = (nb + kChunk - 1) / kChunk;
        for (int c = nch - 1; c >= 0; --c) {
            const int e0 = b0 + c * kChunk;
            const uint32_t bits = (s_act[wid][e0 >> 5] >> (e0 & 31)) & ((1u << kChunk) - 1u);
            if (bits == 0u) continue;
            const int hi = 31 - __clz((int)bits);
            float Tc = ckpt[(size_t)(e0 / kChunk) * kTilePx + tid];
            float ra[kChunk], rT[kChunk], rl[kChunk], rp[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
                if (j <= hi) {
                    const FwdRec& f = s_fwd[c * kChunk + j];
                    PairR2 q = pair_r2(f, pxf, pyf);
                    float a = alpha_of(f, q.r2, rl[j], rp[j]);
                    ra[j] = a;
                    rT[j] = Tc;
                    Tc = Tc * (1.0f - a);
                }
            }
#pragma unroll
            for (int j = kChunk - 1; j >= 0; --j) {
                if ((bits >> j) & 1u) {
                    const int jj = c * kChunk + j;
                    const bool active = valid && !(ra[j] < 0.001f);
                    n_active += (uint32_t)__popcll(__ballot(active));
                    n_went += 1u;
                    uint32_t gv[16];
                    if (active) {
                        const FwdRec& f = s_fwd[jj];
                        PairR2 q = pair_r2(f, pxf, pyf);
                        pair_grads(f, s_bwd[jj], q, rl[j], rp[j], ra[j], rT[j], st, gv);
                    } else {
#pragma unroll
                        for (int k = 0; k < 16; ++k) gv[k] = 0u;
                    }
                    uint32_t tot = wave_reduce16(gv, lane);
                    if ((lane & 3) == 0 && tot != 0u) {
                        int idx = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                                  ((lane >> 2) & 1);
                        atomicAdd(&s_grad[jj][idx], tot);
                    }
                }
            }
        }
        __syncthreads();
        for (int e = tid; e < nb * 16; e += 256) {
            uint32_t val = (&s_grad[0][0])[e];
            if (val != 0u) atomicAdd((unsigned int*)&A.grads[16 * (size_t)s_ids[e >> 4] + (e & 15)], val);
        }
    }
    if (lane == 0) {
        s_cnt[0][wid] = n_active;
        s_cnt[1][wid] = n_went;
        s_cnt[2][wid] = n_feval;
    }
    __syncthreads();
    if (tid == 0) {
        double a2 = (s_red[0][0] + s_red[0][1]) + (s_red[0][2] + s_red[0][3]);
        double a1 = (s_red[1][0] + s_red[1][1]) + (s_red[1][2] + s_red[1][3]);
        A.tile_loss[2 * tile] = a2;
        A.tile_loss[2 * tile + 1] = a1;
        uint32_t vw = min(v.W - tx * kTile, (uint32_t)kTile), vh = min(v.H - ty * kTile, (uint32_t)kTile);
        // the tile's counters in its slot of A.part (summed by launch_sum_parts)
        unsigned long long* pc = A.part + 4 * (size_t)tile;
        pc[0] = (unsigned long long)count * (unsigned long long)(vw * vh);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            pc[k + 1] = (unsigned long long)s_cnt[k][0] + s_cnt[k][1] + s_cnt[k][2] + s_cnt[k][3];
    }
}

void launch_composite_ref(hipStream_t st, const ViewParams& v, const CompositeArgs& a) {
    if (v.num_tiles == 0) return;
    hipLaunchKernelGGL(k_composite_ref, dim3(v.num_tiles), dim3(256), 0, st, v, a);
}

}  // namespace nps
