// k_bin.hip -- per-view preprocess, tile-key generation and tile ranges.
//
// Replaces the binning block of splat_backward.wgsl:141-211.  The reference
// tests every splat against every 16x16 tile with an LDS bitmask and keeps the
// first 1024 ascending ids.  Here each splat is projected once:
//   k_preprocess : cull (opacity < 0.005, w < 0), the reference's square
//                  bounding box (R = pow(15/max(b,1e-4), 1/a),
//                  max_r = R*sqrt(max(cov00,cov11,1e-9))), the exact range of
//                  tiles whose rectangle passes the reference's float overlap
//                  test (binary search with the identical float predicate), and
//                  the forward/backward records of the splat;
//   k_tile_cut   : per tile, the id block in which its count of covering
//                  splats reaches the cap (2D difference arrays per block of
//                  splat ids, prefix-summed); later blocks cannot be among the
//                  tile's first `cap` ascending ids, so they emit no key;
//   k_keys_wave  : one 64-bit key (tile << 32 | splat id) per kept (splat, tile),
//                  written in splat order so a stable sort on the tile bits
//                  leaves every tile list in ascending id order;
//   k_tile_ranges: [start, end) of each tile in the sorted keys.
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

__global__ __launch_bounds__(256) void k_preprocess(const float4* __restrict__ splats, uint32_t n, ViewParams v,
                                                    FwdRec* __restrict__ fwd, BwdRec* __restrict__ bwd,
                                                    uint32_t* __restrict__ counts, uint4* __restrict__ rects) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 s0 = splats[4 * i + 0], s1 = splats[4 * i + 1], s2 = splats[4 * i + 2], s3 = splats[4 * i + 3];
    counts[i] = 0u;
    // s0 = x y z sx ; s1 = r g b op ; s2 = qw qx qy qz ; s3 = sy a b sz
    if (s1.w < 0.005f) return;
    F4 clip = mat_vec(v.vp, s0.x, s0.y, s0.z, 1.0f);
    float aspect = v.aspect;
    float scr_x = clip.x / clip.w * aspect;
    float scr_y = clip.y / clip.w;
    float w = clip.w;
    if (w < 0.0f) return;
    float sx = fmax_w(s0.w, 0.0001f), sy = fmax_w(s3.x, 0.0001f), sz = fmax_w(s3.w, 0.0001f);
    float safe_sb = fmax_w(s3.z, 0.0001f);
    float Rr = pow_det(15.0f / safe_sb, 1.0f / s3.y);
    F3 ax_w = quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{1.0f, 0.0f, 0.0f});
    F3 ay_w = quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{0.0f, 1.0f, 0.0f});
    F3 az_w = quat_rotate(s2.x, s2.y, s2.z, s2.w, F3{0.0f, 0.0f, 1.0f});
    F2 ax_s = project_axis(v.vp, ax_w, clip.x, clip.y, w, aspect);
    F2 ay_s = project_axis(v.vp, ay_w, clip.x, clip.y, w, aspect);
    F2 az_s = project_axis(v.vp, az_w, clip.x, clip.y, w, aspect);
    float m0x = ax_s.x * sx, m0y = ax_s.y * sx;
    float m1x = ay_s.x * sy, m1y = ay_s.y * sy;
    float m2x = az_s.x * sz, m2y = az_s.y * sz;
    float cov00 = (m0x * m0x + m1x * m1x) + m2x * m2x;
    float cov01 = (m0x * m0y + m1x * m1y) + m2x * m2y;
    float cov11 = (m0y * m0y + m1y * m1y) + m2y * m2y;
    float filter_std = 0.3f * (2.0f / (float)v.H);
    float filter2 = filter_std * filter_std;
    cov00 = cov00 + filter2;
    cov11 = cov11 + filter2;
    float max_r = Rr * sqrtf(fmax_w(fmax_w(cov00, cov11), 1e-9f));
    float sminx = scr_x - max_r, sminy = scr_y - max_r;
    float smaxx = scr_x + max_r, smaxy = scr_y + max_r;

    const uint32_t W = v.W, H = v.H;
    const uint32_t T = (uint32_t)kTile;
    // the tiles the reference's per-tile box test accepts (BW:148-153)
    const uint4 tr = tile_rect(sminx, smaxx, sminy, smaxy, W, H, T, v.tiles_x, v.tiles_y, aspect);
    const uint32_t tx0 = tr.x, tx_end = tr.y, ty0 = tr.z, ty_end = tr.w;
    if (tx0 >= tx_end || ty0 >= ty_end) return;
    counts[i] = (tx_end - tx0) * (ty_end - ty0);
    rects[i] = make_uint4(tx0, tx_end, ty0, ty_end);

    // records (splat_backward.wgsl:92-134 hoisted out of the per-pair loop)
    float det = cov00 * cov11 - cov01 * cov01;
    float inv_det = (fabsf(det) < 1e-10f) ? 0.0f : 1.0f / det;
    FwdRec f;
    f.scr_x = scr_x;
    f.scr_y = scr_y;
    f.A = cov11 * inv_det;
    f.B = -cov01 * inv_det;
    f.C = cov00 * inv_det;
    f.op = s1.w;
    f.r = s1.x;
    f.g = s1.y;
    f.b = s1.z;
    f.sa = s3.y;
    f.sb = s3.z;
    float lin = fmax_w(w, 0.1f);
    f.depth = clamp_w(1.0f - 0.1f / lin, 0.0f, 1.0f);
    // Pairs with r^2 > rcut2 have power = -b*pow(r, a) <= -15, so alpha is exactly
    // 0 and the pair can skip pow/exp.  rcut = 1.001 * R with R = (15/b)^(1/a)
    // as computed above; for a in [0.1, 64] and b >= 1e-4 the 0.1 % margin
    // (>= 1e-4 relative in b*r^a) exceeds the combined error of the two pow
    // evaluations (< 2e-5 relative), so the float test power > -15 is false.
    float rcut2 = INFINITY;
    if (s3.y >= 0.1f && s3.y <= 64.0f && s3.z >= 0.0001f && s3.z <= 1e30f && Rr > 0.0f && Rr <= 1e15f) {
        float rc = Rr * 1.001f;
        rcut2 = fmax_w(rc * rc, 0.0001f);
    }
    f.rcut2 = rcut2;
    f.pad[0] = f.pad[1] = f.pad[2] = 0.0f;
    fwd[i] = f;

    BwdRec b;
    b.m[0] = m0x; b.m[1] = m0y; b.m[2] = m1x; b.m[3] = m1y; b.m[4] = m2x; b.m[5] = m2y;
    b.axs[0] = ax_s.x; b.axs[1] = ax_s.y; b.axs[2] = ay_s.x; b.axs[3] = ay_s.y; b.axs[4] = az_s.x; b.axs[5] = az_s.y;
    b.s3[0] = sx; b.s3[1] = sy; b.s3[2] = sz;
    float w2 = w * w;
    float dddw = 0.1f / (w * w);
    for (int j = 0; j < 3; ++j) {
        float vp0 = v.vp[4 * j + 0], vp1 = v.vp[4 * j + 1], vp3 = v.vp[4 * j + 3];
        b.J0[j] = aspect * (vp0 * w - clip.x * vp3) / w2;
        b.J1[j] = (vp1 * w - clip.y * vp3) / w2;
        b.vpd[j] = vp3 * dddw;
    }
    b.pos[0] = s0.x; b.pos[1] = s0.y; b.pos[2] = s0.z;
    b.q[0] = s2.x; b.q[1] = s2.y; b.q[2] = s2.z; b.q[3] = s2.w;
    b.w = w;
    bwd[i] = b;
}

constexpr uint32_t kStripRows = 16;  // diff-grid rows per workgroup of k_diff_build
constexpr uint32_t kDiffSubs = 4;    // id sub-ranges per (strip, id block), summed by k_diff_cols

// 2D difference array of the rectangles of one id block sb, restricted to a
// strip of kStripRows grid rows, accumulated with LDS atomics and prefix-summed
// along x: after k_diff_cols, cnt[sb][ty][tx] = splats of block sb covering tile
// (tx, ty).  Grid: (strips, id blocks).
__global__ __launch_bounds__(256) void k_diff_build(const uint4* __restrict__ rects,
                                                    const uint32_t* __restrict__ counts, uint32_t n,
                                                    uint32_t splat_block, uint32_t gx, uint32_t gy, uint32_t nsb,
                                                    int32_t* __restrict__ diff) {
    extern __shared__ int32_t s_strip[];  // kStripRows * gx
    const uint32_t sb = blockIdx.y, y0 = blockIdx.x * kStripRows, sub = blockIdx.z;
    for (uint32_t k = threadIdx.x; k < kStripRows * gx; k += blockDim.x) s_strip[k] = 0;
    __syncthreads();
    const uint32_t part = (splat_block + kDiffSubs - 1) / kDiffSubs;
    const uint32_t b0 = sb * splat_block, b1 = min(n, b0 + splat_block);
    const uint32_t i0 = min(b1, b0 + sub * part), i1 = min(b1, i0 + part);
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        if (counts[i] == 0u) continue;
        uint4 r = rects[i];
        if (r.z >= y0 && r.z < y0 + kStripRows) {
            atomicAdd(&s_strip[(r.z - y0) * gx + r.x], 1);
            atomicAdd(&s_strip[(r.z - y0) * gx + r.y], -1);
        }
        if (r.w >= y0 && r.w < y0 + kStripRows) {
            atomicAdd(&s_strip[(r.w - y0) * gx + r.x], -1);
            atomicAdd(&s_strip[(r.w - y0) * gx + r.y], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < kStripRows) {
        int32_t acc = 0;
        int32_t* row = s_strip + threadIdx.x * gx;
        for (uint32_t x = 0; x < gx; ++x) {
            acc += row[x];
            row[x] = acc;
        }
    }
    __syncthreads();
    int32_t* out = diff + (((size_t)sub * nsb + sb) * gy + y0) * gx;
    for (uint32_t k = threadIdx.x; k < kStripRows * gx; k += blockDim.x)
        if (y0 + k / gx < gy) out[k] = s_strip[k];
}

// Sums the kDiffSubs partial grids of each id block and prefix-sums along y;
// the result lands in sub-grid 0.
__global__ void k_diff_cols(int32_t* diff, uint32_t nsb, uint32_t gx, uint32_t gy) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsb * gx) return;
    uint32_t sb = k / gx, x = k % gx;
    const size_t sub_stride = (size_t)nsb * gx * gy;
    int32_t* base = diff + (size_t)sb * gx * gy + x;
    int32_t acc = 0;
    for (uint32_t y = 0; y < gy; ++y) {
        int32_t v = 0;
#pragma unroll
        for (uint32_t q = 0; q < kDiffSubs; ++q) v += base[q * sub_stride + (size_t)y * gx];
        acc += v;
        base[(size_t)y * gx] = acc;
    }
}

// cut[t] = first id block at which the running count of splats covering tile t
// reaches `cap` (the last block if it never does); full[t] = total count.  The
// first `cap` ascending ids of the tile all lie in blocks <= cut[t], so keys of
// later blocks are never composited and are not emitted.
__global__ void k_tile_cut(const int32_t* diff, uint32_t nsb, uint32_t gx, uint32_t gy, uint32_t tiles_x,
                           uint32_t num_tiles, uint32_t cap, uint32_t* cut, uint32_t* full,
                           unsigned long long* over_cap) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tiles) return;
    uint32_t tx = t % tiles_x, ty = t / tiles_x;
    uint32_t cum = 0, c = nsb - 1u;
    bool found = false;
    for (uint32_t sb = 0; sb < nsb; ++sb) {
        cum += (uint32_t)diff[((size_t)sb * gy + ty) * gx + tx];
        if (!found && cum >= cap) {
            c = sb;
            found = true;
        }
    }
    cut[t] = c;
    full[t] = cum;
    if (cum > cap) atomicAdd(over_cap, 1ull);
}

// Keys of one splat per wave iteration: the 64 lanes take 64 tiles of the
// splat's rectangle at a time, drop tiles whose cut is below the splat's id
// block, and (emit mode) write the kept keys contiguously, row-major in tile
// order, at the splat's offset; count mode writes the number of kept keys.  One
// wave per 64 consecutive splats, so a large rectangle costs its wave a few
// more iterations instead of stalling one lane.
// P[b][y][x] = number of tiles in [0,x) x [0,y) whose cut is >= b: a 2D prefix
// table per id block b, so the keys a splat of block b keeps in its rectangle
// [x0,x1) x [y0,y1) are P[b][y1][x1] - P[b][y0][x1] - P[b][y1][x0] + P[b][y0][x0].
__global__ void k_cut_rows(const uint32_t* cut, uint32_t nsb, uint32_t tiles_x, uint32_t tiles_y, int32_t* P) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t gx = tiles_x + 1u, gy = tiles_y + 1u;
    if (k >= nsb * gy) return;
    const uint32_t b = k / gy, y = k % gy;
    int32_t* row = P + ((size_t)b * gy + y) * gx;
    int32_t acc = 0;
    row[0] = 0;
    for (uint32_t x = 0; x < tiles_x; ++x) {
        if (y > 0u) acc += (cut[(y - 1u) * tiles_x + x] >= b) ? 1 : 0;
        row[x + 1u] = acc;
    }
}
__global__ void k_cut_cols(uint32_t nsb, uint32_t tiles_x, uint32_t tiles_y, int32_t* P) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t gx = tiles_x + 1u, gy = tiles_y + 1u;
    if (k >= nsb * gx) return;
    const uint32_t b = k / gx, x = k % gx;
    int32_t* col = P + (size_t)b * gy * gx + x;
    int32_t acc = 0;
    for (uint32_t y = 0; y < gy; ++y) {
        acc += col[(size_t)y * gx];
        col[(size_t)y * gx] = acc;
    }
}
__global__ __launch_bounds__(256) void k_count_kept(uint32_t n, const uint32_t* __restrict__ counts,
                                                    const uint4* __restrict__ rects, uint32_t splat_block,
                                                    uint32_t tiles_x, uint32_t tiles_y, const int32_t* __restrict__ P,
                                                    uint32_t* __restrict__ kept) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (counts[i] == 0u) {
        kept[i] = 0u;
        return;
    }
    const uint4 r = rects[i];
    const uint32_t gx = tiles_x + 1u, gy = tiles_y + 1u;
    const int32_t* T = P + (size_t)(i / splat_block) * gy * gx;
    int32_t c = T[r.w * gx + r.y] - T[r.z * gx + r.y] - T[r.w * gx + r.x] + T[r.z * gx + r.x];
    kept[i] = (uint32_t)c;
}

// The cut table is staged in LDS as bytes (id blocks < 256) when it fits, so
// the per-splat lookups cost LDS rather than L2 latency.
constexpr uint32_t kCutLdsMax = 48 * 1024;

template <bool kEmit>
__global__ __launch_bounds__(256) void k_keys_wave(uint32_t n, const uint32_t* __restrict__ counts,
                                                   const uint4* __restrict__ rects, uint32_t tiles_x,
                                                   uint32_t num_tiles, const uint32_t* __restrict__ cut,
                                                   uint32_t splat_block, const uint64_t* __restrict__ offsets,
                                                   uint32_t* __restrict__ kept_counts,
                                                   uint64_t* __restrict__ keys) {
    extern __shared__ uint8_t s_cut[];
    const bool lds_cut = cut && num_tiles <= kCutLdsMax;
    if (lds_cut) {
        for (uint32_t t = threadIdx.x; t < num_tiles; t += blockDim.x) s_cut[t] = (uint8_t)cut[t];
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wbase = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // each lane loads one splat of the wave; the wave then walks them in order
    const uint32_t mine = wbase + lane;
    const uint32_t my_cnt = mine < n ? counts[mine] : 0u;
    uint4 my_r = make_uint4(0u, 0u, 0u, 0u);
    uint64_t my_off = 0ull;
    if (my_cnt) {
        my_r = rects[mine];
        if (kEmit) my_off = offsets[mine];
    }
    uint32_t my_kept = 0u;
    for (uint32_t s = 0; s < 64u; ++s) {
        const uint32_t cnt = __shfl(my_cnt, (int)s, 64);
        if (cnt == 0u) continue;
        const uint32_t i = wbase + s;
        const uint32_t rx0 = __shfl(my_r.x, (int)s, 64), rx1 = __shfl(my_r.y, (int)s, 64);
        const uint32_t ry0 = __shfl(my_r.z, (int)s, 64), ry1 = __shfl(my_r.w, (int)s, 64);
        const uint32_t rw = rx1 - rx0, total = rw * (ry1 - ry0);
        const uint32_t sb = i / splat_block;
        const uint64_t pos = kEmit ? __shfl(my_off, (int)s, 64) : 0ull;
        uint32_t kept = 0u;
        for (uint32_t k0 = 0; k0 < total; k0 += 64u) {
            const uint32_t k = k0 + lane;
            bool keep = false;
            uint32_t tile = 0u;
            if (k < total) {
                const uint32_t ty = ry0 + k / rw, tx = rx0 + k % rw;
                tile = ty * tiles_x + tx;
                keep = !cut || sb <= (lds_cut ? (uint32_t)s_cut[tile] : cut[tile]);
            }
            const uint64_t bal = __ballot(keep);
            if (kEmit && keep)
                keys[pos + kept + (uint32_t)__popcll(bal & lt_mask)] = ((uint64_t)tile << 32) | (uint64_t)i;
            kept += (uint32_t)__popcll(bal);
        }
        if (lane == s) my_kept = kept;
    }
    if (!kEmit && mine < n) kept_counts[mine] = my_kept;
}

// ranges[2t] = start, ranges[2t+1] = end of tile t in the sorted keys.
// Tiles without keys keep the [0,0) the caller cleared.
__global__ __launch_bounds__(256) void k_tile_ranges(const uint64_t* __restrict__ keys, uint64_t m,
                                                     const uint64_t* __restrict__ m_dev,
                                                     uint32_t* __restrict__ ranges) {
    if (m_dev) m = min(*m_dev, m);
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    uint32_t t = (uint32_t)(keys[j] >> 32);
    if (j == 0 || (uint32_t)(keys[j - 1] >> 32) != t) ranges[2 * t] = (uint32_t)j;
    if (j + 1 == m || (uint32_t)(keys[j + 1] >> 32) != t) ranges[2 * t + 1] = (uint32_t)(j + 1);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
}

// 16-byte stores: p 16-byte aligned, n4 = words / 4
__global__ void k_fill_u32x4(uint4* p, uint32_t v, uint64_t n4) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n4) p[j] = make_uint4(v, v, v, v);
}

void launch_preprocess(hipStream_t st, const float* splats, uint32_t n, const ViewParams& v, FwdRec* fwd,
                       BwdRec* bwd, uint32_t* counts, uint4* rects) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_preprocess, dim3((n + 255) / 256), dim3(256), 0, st, (const float4*)splats, n, v, fwd,
                       bwd, counts, rects);
}

void launch_tile_cut(hipStream_t st, const ViewParams& v, uint32_t n, int32_t* diff, uint32_t splat_block,
                     uint32_t cap, uint32_t* cut, uint32_t* full, unsigned long long* over_cap,
                     const uint32_t* counts, const uint4* rects, uint32_t* kept_counts) {
    if (n == 0) return;
    const uint32_t nsb = (n + splat_block - 1) / splat_block;
    const uint32_t gx = v.tiles_x + 1u, gy = v.tiles_y + 1u;
    const uint32_t strips = (gy + kStripRows - 1) / kStripRows;
    hipLaunchKernelGGL(k_diff_build, dim3(strips, nsb, kDiffSubs), dim3(256), sizeof(int32_t) * kStripRows * gx, st,
                       rects, counts, n, splat_block, gx, gy, nsb, diff);
    hipLaunchKernelGGL(k_diff_cols, dim3((nsb * gx + 63) / 64), dim3(64), 0, st, diff, nsb, gx, gy);
    hipLaunchKernelGGL(k_tile_cut, dim3((v.num_tiles + 255) / 256), dim3(256), 0, st, (const int32_t*)diff, nsb,
                       gx, gy, v.tiles_x, v.num_tiles, cap, cut, full, over_cap);
    // the diff grids are consumed; their buffer now holds the prefix tables P
    hipLaunchKernelGGL(k_cut_rows, dim3((nsb * gy + 63) / 64), dim3(64), 0, st, (const uint32_t*)cut, nsb, v.tiles_x,
                       v.tiles_y, diff);
    hipLaunchKernelGGL(k_cut_cols, dim3((nsb * gx + 63) / 64), dim3(64), 0, st, nsb, v.tiles_x, v.tiles_y, diff);
    hipLaunchKernelGGL(k_count_kept, dim3((n + 255) / 256), dim3(256), 0, st, n, counts, rects, splat_block,
                       v.tiles_x, v.tiles_y, (const int32_t*)diff, kept_counts);
}

void launch_emit_keys(hipStream_t st, uint32_t n, const uint32_t* counts, const uint64_t* offsets,
                      const uint4* rects, uint32_t tiles_x, uint32_t num_tiles, const uint32_t* cut,
                      uint32_t splat_block, uint64_t* keys) {
    if (n == 0) return;
    const size_t lds = (cut && num_tiles <= kCutLdsMax) ? num_tiles : 0;
    hipLaunchKernelGGL(k_keys_wave<true>, dim3((n + 255) / 256), dim3(256), lds, st, n, counts, rects, tiles_x,
                       num_tiles, cut, splat_block, offsets, (uint32_t*)nullptr, keys);
}

void launch_tile_ranges(hipStream_t st, const uint64_t* keys, uint64_t m, uint32_t* ranges, const uint64_t* m_dev) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, keys, m, m_dev, ranges);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint64_t n) {
    if (n == 0) return;
    if (n >= 4096 && ((uintptr_t)p & 15u) == 0) {
        const uint64_t n4 = n / 4;
        hipLaunchKernelGGL(k_fill_u32x4, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, (uint4*)p, v, n4);
        if (n % 4) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, p + 4 * n4, v, n % 4);
        return;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}

}  // namespace nps
