// k_ext.hip -- EXT path: 3D Gaussian splats with SH colour, depth-ordered
// uncapped tile lists, early termination, L1 loss and FP32 gradients.
//
//   k_ext_preprocess : per splat, the projection of nps_ext.h, the 3-sigma tile
//                      rectangle (binary search with the float overlap predicate
//                      of the REF binning, splat_backward.wgsl:148-153), the
//                      composite record and the depth key (float bits of w > 0,
//                      monotone in w) << 32 | id;
//   depth order      : the per-splat keys sorted on their high 32 bits (stable,
//                      so equal depths stay in id order) give every splat its
//                      rank in (depth, id) order; the last pass gathers the
//                      tile counts in rank order;
//   k_bin_*          : binning in rank order without a key sort (row buckets,
//                      then a stable column pass) into per-tile lists in
//                      (depth, id) order; k_ext_keys + a stable LSD sort on the
//                      tile bits is the fallback for grids above 256 x 256 tiles;
//   k_composite_ext  : per 16x16 tile, front-to-back compositing with early
//                      termination, image + L1/L2 loss, then the reverse pass;
//   k_ext_chain      : per splat, the 2D gradients (screen centre, conic, colour,
//                      opacity) through the projection to the 64 parameters, in
//                      FP64, added to the FP32 gradient buffer;
//   k_adam_ext       : Adam with the per-group learning rates.
#include <algorithm>

#include "nps_common.h"
#include "nps_ext.h"
#include "nps_kernels.h"

namespace nps {

// ---------------------------------------------------------------- preprocess

// One splat; returns the float bits of its depth w when it is drawn (a tile
// rectangle), else 0xFFFFFFFF.
template <int DEG>
__device__ __forceinline__ uint32_t ext_preprocess_splat(uint32_t i, const float4* __restrict__ splats,
                                                         const float4* __restrict__ sh, const ViewParams& v,
                                                         float cam0, float cam1, float cam2,
                                                         ExtRec* __restrict__ rec, uint32_t* __restrict__ counts,
                                                         uint4* __restrict__ rects, uint64_t* __restrict__ dkey) {
    constexpr int kNk = (DEG + 1) * (DEG + 1);
    constexpr int kSh4 = (3 * kNk + 3) / 4;
    float s[16];
    float c[4 * kSh4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float4 t = splats[4 * (size_t)i + q];
        s[4 * q + 0] = t.x; s[4 * q + 1] = t.y; s[4 * q + 2] = t.z; s[4 * q + 3] = t.w;
    }
#pragma unroll
    for (int q = 0; q < kSh4; ++q) {
        float4 t = sh[12 * (size_t)i + q];
        c[4 * q + 0] = t.x; c[4 * q + 1] = t.y; c[4 * q + 2] = t.z; c[4 * q + 3] = t.w;
    }
    counts[i] = 0u;
    dkey[i] = 0xFFFFFFFF00000000ull | i;  // not drawn: after every drawn splat
    const float cam[3] = {cam0, cam1, cam2};
    ExtProj o;
    ext_project(s, c, DEG, v.vp, cam, v.W, v.H, o);
    if (!o.valid) return 0xFFFFFFFFu;
    const float ebx = o.ext[0] * kExtBinScale + kExtBinPad, eby = o.ext[1] * kExtBinScale + kExtBinPad;
    const float sminx = o.scr[0] - ebx, smaxx = o.scr[0] + ebx;
    const float sminy = o.scr[1] - eby, smaxy = o.scr[1] + eby;
    if (!(sminx == sminx && smaxx == smaxx && sminy == sminy && smaxy == smaxy)) return 0xFFFFFFFFu;
    const uint32_t W = v.W, H = v.H, T = v.tile;
    const float aspect = v.aspect;
    const uint4 tr = tile_rect(sminx, smaxx, sminy, smaxy, W, H, T, v.tiles_x, v.tiles_y, aspect);
    const uint32_t tx0 = tr.x, tx_end = tr.y, ty0 = tr.z, ty_end = tr.w;
    if (tx0 >= tx_end || ty0 >= ty_end) return 0xFFFFFFFFu;
    counts[i] = (tx_end - tx0) * (ty_end - ty0);
    rects[i] = make_uint4(tx0, tx_end, ty0, ty_end);
    dkey[i] = ((uint64_t)__float_as_uint(o.w) << 32) | i;
    ExtRec f;
    f.scr_x = o.scr[0];
    f.scr_y = o.scr[1];
    f.A = o.A;
    f.B = o.B;
    f.C = o.C;
    f.op = o.op;
    f.r = o.rgb[0];
    f.g = o.rgb[1];
    f.b = o.rgb[2];
    // alpha = min(op exp(-r^2/2), 0.99) < 1/255 for every r^2 > rcut2:
    // rcut2 >= 2 ln(255 op) + 1e-3 leaves a relative margin of e^(-5e-4) in
    // alpha, far above the rounding of exp and of the product (< 1e-6).
    if (o.op < kExtMinAlpha * 0.999f) {
        f.rcut2 = -INFINITY;
    } else {
        float l = 2.0f * log_det(o.op * 255.0f);
        f.rcut2 = fmax_w(l, 0.0f) * 1.0001f + 1e-3f;
    }
    f.ext_x = o.ext[0];
    f.ext_y = o.ext[1];
    rec[i] = f;
    return __float_as_uint(o.w);
}

// Per splat the projection of ext_preprocess_splat; per workgroup (256 splats)
// the least depth bits of its drawn splats (0xFFFFFFFF: none) into wmin_part,
// the base of the depth order's compressed keys (null: not needed).  The grid
// also clears clr.n[c] words at clr.p[c] (null: none), the view's small
// counters, so they need no launch of their own.
struct ClearPair {
    uint32_t* p[3];
    uint32_t n[3];
};
template <int DEG>
__global__ __launch_bounds__(256) void k_ext_preprocess(const float4* __restrict__ splats,
                                                        const float4* __restrict__ sh, uint32_t n, ViewParams v,
                                                        float cam0, float cam1, float cam2,
                                                        ExtRec* __restrict__ rec, uint32_t* __restrict__ counts,
                                                        uint4* __restrict__ rects, uint64_t* __restrict__ dkey,
                                                        uint32_t* __restrict__ wmin_part, ClearPair clr) {
    __shared__ uint32_t s_min[4];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (clr.p[c])
            for (uint32_t k = i; k < clr.n[c]; k += gridDim.x * blockDim.x) clr.p[c][k] = 0u;
    uint32_t wb = 0xFFFFFFFFu;
    if (i < n) wb = ext_preprocess_splat<DEG>(i, splats, sh, v, cam0, cam1, cam2, rec, counts, rects, dkey);
    if (!wmin_part) return;  // uniform
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) wb = min(wb, (uint32_t)__shfl_xor((int)wb, o, 64));
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = wb;
    __syncthreads();
    if (threadIdx.x == 0) wmin_part[blockIdx.x] = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
}

// One key tile << 32 | id per (splat, covered tile), splats taken in rank
// order, each splat's tiles row by row.  A wave owns 64 consecutive ranks, whose
// keys are contiguous from offsets[first rank]: the wave scans its 64 tile
// counts, then its lanes write consecutive keys of that range, each finding its
// splat by binary search over the inclusive counts (fully coalesced stores).
__global__ __launch_bounds__(256) void k_ext_keys(uint32_t n, const uint64_t* __restrict__ order,
                                                  const uint32_t* __restrict__ rank_counts,
                                                  const uint4* __restrict__ rects, uint32_t tiles_x,
                                                  const uint64_t* __restrict__ offsets, uint64_t* __restrict__ keys,
                                                  uint64_t cap, uint32_t* __restrict__ overflow) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wbase = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
    if (wbase >= n) return;  // whole wave
    const uint32_t mine = wbase + lane;
    const uint32_t my_cnt = mine < n ? rank_counts[mine] : 0u;
    uint32_t my_id = 0u, my_x0 = 0u, my_w = 1u, my_y0 = 0u;
    float my_rcp = 1.0f;
    if (my_cnt) {
        my_id = (uint32_t)order[mine];
        const uint4 r = rects[my_id];
        my_x0 = r.x;
        my_w = r.y - r.x;
        my_y0 = r.z;
        my_rcp = 1.0f / (float)my_w;
    }
    uint32_t incl = my_cnt;  // inclusive scan of the counts over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += y;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    const uint64_t pos0 = offsets[wbase];
    if (pos0 + total > cap && lane == 0) *overflow = 1u;
    // every lane runs every iteration: a cross-lane read from an inactive lane returns 0
    for (uint32_t g0 = 0u; g0 < total; g0 += 64u) {
        const uint32_t g = g0 + lane;
        // s = first lane whose inclusive count exceeds g (63 when g >= total)
        uint32_t s = 0u;
#pragma unroll
        for (uint32_t b = 32u; b > 0u; b >>= 1) {
            const uint32_t v = (uint32_t)__shfl((int)incl, (int)(s + b - 1u), 64);
            if (v <= g) s += b;
        }
        const uint32_t k = g - ((uint32_t)__shfl((int)incl, (int)s, 64) - (uint32_t)__shfl((int)my_cnt, (int)s, 64));
        const uint32_t w = (uint32_t)__shfl((int)my_w, (int)s, 64);
        const float rcp = __shfl(my_rcp, (int)s, 64);
        // k / w from the float reciprocal (within one of the quotient for k < 2^24), corrected
        uint32_t q = (uint32_t)((float)k * rcp);
        int32_t rem = (int32_t)k - (int32_t)(q * w);
        if (rem < 0) {
            --q;
            rem += (int32_t)w;
        } else if (rem >= (int32_t)w) {
            ++q;
            rem -= (int32_t)w;
        }
        const uint32_t ty = (uint32_t)__shfl((int)my_y0, (int)s, 64) + q;
        const uint32_t tx = (uint32_t)__shfl((int)my_x0, (int)s, 64) + (uint32_t)rem;
        const uint64_t id = (uint64_t)(uint32_t)__shfl((int)my_id, (int)s, 64);
        if (g < total && pos0 + g < cap) keys[pos0 + g] = ((uint64_t)(ty * tiles_x + tx) << 32) | id;
    }
}

// ---------------------------------------------------------------- binning
//
// Tile lists in (tile, depth, id) order without sorting 8-byte keys.  A key is
// (tile, splat) for every tile of a splat's rectangle; keys are enumerated in
// depth-rank order (a splat's tiles row by row), and each tile's list must keep
// that order.  Two stable counting passes, the first fused into the
// enumeration:
//   k_bin_row_count  : per workgroup of 1024 ranks, its keys per tile row: a
//                      splat adds its width to rows y0..y1-1 (a difference
//                      array in LDS);
//   exclusive scan of the row-major counts: where each (row, workgroup) starts
//                      in the row buckets;
//   k_bin_row_emit   : the same workgroups enumerate their keys in rank order
//                      and write the 4-byte item tx << id_bits | id at its
//                      place in its row bucket (a running place per wave and
//                      row; the lanes of a round ranked by a match mask).  The
//                      tiles of one row of a splat are consecutive keys in one
//                      bucket: runs of consecutive stores;
//   k_bin_col_count  : per 4096 items, the count per tile column and per tile
//                      (an item's row is the bucket its position lies in);
//   exclusive scans: column positions; tile ranges in column-major order;
//   k_bin_col_scatter: a stable scatter on the column writing splat ids.
// Both passes keep the enumeration order within a bucket, so the output holds
// the tiles column by column, each tile's ids in (depth, id) order.
constexpr int kBinMaxBits = 8;      // rows and columns of the tile grid: at most 256 each
constexpr int kBinThreads = 1024;   // ranks per workgroup of the row kernels
constexpr int kBinWaves = kBinThreads / 64;
constexpr int kBinColItems = 16;
constexpr int kBinColBlock = 256 * kBinColItems;  // items per workgroup of the column kernels

// counts1[y * nblk + blk] = keys of ranks [1024 blk, 1024 blk + 1024) in tile row y.
// A rank's splat is drawn when its depth bits (order's high word) are not
// 0xFFFFFFFF: drawn depths are at least 0.1.
__global__ __launch_bounds__(kBinThreads) void k_bin_row_count(uint32_t n, const uint64_t* __restrict__ order,
                                                               const uint4* __restrict__ rects, uint32_t ny,
                                                               uint32_t* __restrict__ counts1) {
    __shared__ int s_diff[(1 << kBinMaxBits) + 1];
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k <= ny; k += kBinThreads) s_diff[k] = 0;
    __syncthreads();
    const uint32_t r = blockIdx.x * kBinThreads + tid;
    const uint64_t o = r < n ? order[r] : ~0ull;
    if ((uint32_t)(o >> 32) != 0xFFFFFFFFu) {
        const uint4 q = rects[(uint32_t)o];
        atomicAdd(&s_diff[q.z], (int)(q.y - q.x));
        atomicAdd(&s_diff[q.w], -(int)(q.y - q.x));
    }
    __syncthreads();
    if (tid < 64) {  // prefix over ny <= 256 rows, one wave
        int carry = 0;
        for (uint32_t y0 = 0; y0 < ny; y0 += 64) {
            const uint32_t y = y0 + tid;
            int x = y < ny ? s_diff[y] : 0;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(x, o, 64);
                if ((int)tid >= o) x += t;
            }
            if (y < ny) counts1[(size_t)y * gridDim.x + blockIdx.x] = (uint32_t)(carry + x);
            carry += __shfl(x, 63, 64);
        }
    }
}

// The keys of the same workgroups, written as items at their row places.
// offs1: exclusive scan of counts1 (row-major).  Items beyond cap are dropped
// (*overflow = 1).
__global__ __launch_bounds__(kBinThreads) void k_bin_row_emit(uint32_t n, const uint64_t* __restrict__ order,
                                                              const uint4* __restrict__ rects, uint32_t ny,
                                                              int row_bits, int id_bits,
                                                              const uint32_t* __restrict__ offs1,
                                                              uint32_t* __restrict__ items, uint64_t cap,
                                                              uint32_t* __restrict__ overflow) {
    __shared__ int s_row[kBinWaves][(1 << kBinMaxBits) + 1];  // per wave: counts, then running places
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    for (uint32_t k = tid; k < kBinWaves * ((1u << kBinMaxBits) + 1u); k += kBinThreads) (&s_row[0][0])[k] = 0;
    __syncthreads();
    const uint32_t mine = blockIdx.x * kBinThreads + tid;
    const uint64_t my_o = mine < n ? order[mine] : ~0ull;
    uint32_t my_cnt = 0u, my_id = 0u, my_x0 = 0u, my_w = 1u, my_y0 = 0u;
    float my_rcp = 1.0f;
    if ((uint32_t)(my_o >> 32) != 0xFFFFFFFFu) {  // drawn (as k_bin_row_count)
        my_id = (uint32_t)my_o;
        const uint4 q = rects[my_id];
        my_x0 = q.x;
        my_w = q.y - q.x;
        my_y0 = q.z;
        my_cnt = my_w * (q.w - q.z);
        my_rcp = 1.0f / (float)my_w;
        atomicAdd(&s_row[wid][q.z], (int)my_w);
        atomicAdd(&s_row[wid][q.w], -(int)my_w);
    }
    __syncthreads();
    {   // wave w's count per row: prefix of its difference array (lanes over rows)
        int carry = 0;
        for (uint32_t y0 = 0; y0 < ny; y0 += 64) {
            const uint32_t y = y0 + lane;
            int x = y < ny ? s_row[wid][y] : 0;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(x, o, 64);
                if ((int)lane >= o) x += t;
            }
            const int tot = __shfl(x, 63, 64);
            __builtin_amdgcn_wave_barrier();
            if (y < ny) s_row[wid][y] = carry + x;
            __builtin_amdgcn_wave_barrier();
            carry += tot;
        }
    }
    __syncthreads();
    for (uint32_t y = tid; y < ny; y += kBinThreads) {  // places: the waves in order from the workgroup's place
        uint32_t place = offs1[(size_t)y * gridDim.x + blockIdx.x];
        for (uint32_t w = 0; w < (uint32_t)kBinWaves; ++w) {
            const uint32_t c = (uint32_t)s_row[w][y];
            s_row[w][y] = (int)place;
            place += c;
        }
    }
    __syncthreads();
    // enumerate the wave's keys in rank order, 64 per round
    uint32_t incl = my_cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += y;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t g0 = 0u; g0 < total; g0 += 64u) {
        const uint32_t g = g0 + lane;
        const bool valid = g < total;
        // s = first lane whose inclusive count exceeds g (63 when g >= total)
        uint32_t sl = 0u;
#pragma unroll
        for (uint32_t b = 32u; b > 0u; b >>= 1) {
            const uint32_t v = (uint32_t)__shfl((int)incl, (int)(sl + b - 1u), 64);
            if (v <= g) sl += b;
        }
        const uint32_t k = g - ((uint32_t)__shfl((int)incl, (int)sl, 64) - (uint32_t)__shfl((int)my_cnt, (int)sl, 64));
        const uint32_t w = (uint32_t)__shfl((int)my_w, (int)sl, 64);
        const float rcp = __shfl(my_rcp, (int)sl, 64);
        // k / w from the float reciprocal (within one of the quotient for k < 2^24), corrected
        uint32_t qq = (uint32_t)((float)k * rcp);
        int32_t rem = (int32_t)k - (int32_t)(qq * w);
        if (rem < 0) {
            --qq;
            rem += (int32_t)w;
        } else if (rem >= (int32_t)w) {
            ++qq;
            rem -= (int32_t)w;
        }
        // every lane runs the cross-lane reads (a read from an inactive lane returns 0)
        const uint32_t ty_all = (uint32_t)__shfl((int)my_y0, (int)sl, 64) + qq;
        const uint32_t tx = (uint32_t)__shfl((int)my_x0, (int)sl, 64) + (uint32_t)rem;
        const uint32_t id = (uint32_t)__shfl((int)my_id, (int)sl, 64);
        const uint32_t ty = valid ? ty_all : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int q = 0; q < kBinMaxBits; ++q) {  // bits above row_bits are 0 in every lane: no-ops
            const uint64_t bal = __ballot((ty >> q) & 1u);
            peers &= ((ty >> q) & 1u) ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t place = valid ? (uint32_t)s_row[wid][ty] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0u) s_row[wid][ty] = (int)(place + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            const uint32_t pos = place + before;
            if ((uint64_t)pos < cap) items[pos] = (tx << id_bits) | id;
            else *overflow = 1u;
        }
    }
}

// Per 4096 items: counts2[x * nblk + blk] (tile column x) and tile_cnt[tile] +=.
// The row buckets start at offs1[y * nblk1] (y < ny; the item count after the
// last); a workgroup's items lie in the rows y_first..y_last of its first and
// last item.  Counts need no order: every item adds 1 to its column's and its
// tile's LDS counter (LDS atomics; lanes that hit one address serialise).
__global__ __launch_bounds__(256) void k_bin_col_count(const uint32_t* __restrict__ items, uint64_t m,
                                                       const uint32_t* __restrict__ m_dev,
                                                       const uint32_t* __restrict__ offs1, uint32_t nblk1,
                                                       uint32_t ny, uint32_t nx, int id_bits,
                                                       uint32_t* __restrict__ counts2,
                                                       uint32_t* __restrict__ tile_cnt) {
    constexpr int kLocal = 2048;  // per-workgroup tile counters (rows in the workgroup x nx)
    __shared__ uint32_t s_start[(1 << kBinMaxBits) + 1];
    __shared__ uint32_t s_c[256];
    __shared__ uint32_t s_t[kLocal];
    if (m_dev) m = min((uint64_t)*m_dev, m);
    const uint32_t tid = threadIdx.x;
    for (uint32_t y = tid; y < ny; y += 256) s_start[y] = offs1[(size_t)y * nblk1];
    s_c[tid] = 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kBinColBlock;
    if (base >= m) {
        if (tid < nx) counts2[(size_t)tid * gridDim.x + blockIdx.x] = 0u;
        return;
    }
    const uint32_t last = (uint32_t)min<uint64_t>(base + kBinColBlock, m) - 1u;
    auto row_of = [&](uint32_t p) -> uint32_t {  // largest y with start[y] <= p
        uint32_t y = 0u;
        for (uint32_t b = 1u << kBinMaxBits; b > 0u; b >>= 1)
            if (y + b < ny && s_start[y + b] <= p) y += b;
        return y;
    };
    const uint32_t y_first = row_of((uint32_t)base), y_last = row_of(last);
    const bool local = (y_last - y_first + 1u) * nx <= (uint32_t)kLocal;
    if (local)
        for (uint32_t k = tid; k < (y_last - y_first + 1u) * nx; k += 256) s_t[k] = 0u;
    __syncthreads();
    uint32_t it[kBinColItems];
#pragma unroll
    for (int r = 0; r < kBinColItems; ++r) {
        const uint64_t pos = base + (uint64_t)r * 256 + tid;
        it[r] = pos < m ? items[pos] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kBinColItems; ++r) {
        const uint64_t pos = base + (uint64_t)r * 256 + tid;
        if (pos < m) {
            const uint32_t x = it[r] >> id_bits;
            const uint32_t y = (y_first == y_last) ? y_first : row_of((uint32_t)pos);
            atomicAdd(&s_c[x], 1u);
            if (local) atomicAdd(&s_t[(y - y_first) * nx + x], 1u);
            else atomicAdd(&tile_cnt[y * nx + x], 1u);
        }
    }
    __syncthreads();
    if (tid < nx) counts2[(size_t)tid * gridDim.x + blockIdx.x] = s_c[tid];
    if (local)
        for (uint32_t k = tid; k < (y_last - y_first + 1u) * nx; k += 256) {
            const uint32_t c = s_t[k];
            if (c) atomicAdd(&tile_cnt[(y_first + k / nx) * nx + k % nx], c);
        }
}

// Stable scatter of the items on their column (offs2: exclusive scan of
// counts2), writing the splat ids: 4 waves x 1024 items, ranked per wave as in
// k_radix_scatter, staged in LDS in output order and written as runs of
// consecutive ids per column.
__global__ __launch_bounds__(256) void k_bin_col_scatter(const uint32_t* __restrict__ items, uint64_t m,
                                                         const uint32_t* __restrict__ m_dev, uint32_t nx, int col_bits,
                                                         int id_bits, const uint32_t* __restrict__ offs2,
                                                         uint32_t* __restrict__ ids) {
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t s_ids[kBinColBlock];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_w4[4];
    if (m_dev) m = min((uint64_t)*m_dev, m);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    for (uint32_t k = tid; k < 4 * 256; k += 256) (&wcnt[0][0])[k] = 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kBinColBlock;
    if (base >= m) return;
    const uint32_t nloc = (uint32_t)min<uint64_t>((uint64_t)kBinColBlock, m - base);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wbase = base + (uint64_t)wid * (kBinColBlock / 4);
    const uint32_t idmask = id_bits >= 32 ? 0xFFFFFFFFu : ((1u << id_bits) - 1u);
    uint32_t it[kBinColItems], rank[kBinColItems];
#pragma unroll
    for (int r = 0; r < kBinColItems; ++r) {
        const uint64_t pos = wbase + (uint64_t)r * 64 + lane;
        it[r] = pos < m ? items[pos] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kBinColItems; ++r) {
        const uint64_t pos = wbase + (uint64_t)r * 64 + lane;
        const bool valid = pos < m;
        const uint32_t d = it[r] >> id_bits;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int q = 0; q < kBinMaxBits; ++q) {  // bits above col_bits are 0 in every lane: no-ops
            const uint64_t bal = __ballot((d >> q) & 1u);
            peers &= ((d >> q) & 1u) ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t cnt = valid ? wcnt[wid][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0u) wcnt[wid][d] = cnt + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[r] = cnt + before;
    }
    __syncthreads();
    {   // one thread per column: the workgroup's local start of column d, the waves' starts
        const uint32_t d = tid;
        const bool col = d < nx;
        const uint32_t c0 = col ? wcnt[0][d] : 0u, c1 = col ? wcnt[1][d] : 0u, c2 = col ? wcnt[2][d] : 0u,
                       c3 = col ? wcnt[3][d] : 0u;
        const uint32_t c = c0 + c1 + c2 + c3;
        uint32_t x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if ((int)lane >= o) x += y;
        }
        if (lane == 63u) s_w4[wid] = x;
        __syncthreads();
        uint32_t wb = 0u;
        for (uint32_t w = 0; w < wid; ++w) wb += s_w4[w];
        const uint32_t lstart = wb + x - c;
        if (col) {
            s_gbase[d] = offs2[(size_t)d * gridDim.x + blockIdx.x] - lstart;  // modulo 2^32
            wcnt[0][d] = lstart;
            wcnt[1][d] = lstart + c0;
            wcnt[2][d] = lstart + c0 + c1;
            wcnt[3][d] = lstart + c0 + c1 + c2;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kBinColItems; ++r) {
        const uint64_t pos = wbase + (uint64_t)r * 64 + lane;
        if (pos < m) s_ids[wcnt[wid][it[r] >> id_bits] + rank[r]] = it[r];
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kBinColItems; ++k) {
        const uint32_t p = (uint32_t)k * 256u + tid;
        if (p < nloc) {
            const uint32_t v = s_ids[p];
            ids[s_gbase[v >> id_bits] + p] = v & idmask;
        }
    }
}

// Splat ids of sorted tile << 32 | id keys (the fallback path).
__global__ void k_keys_to_ids(const uint64_t* __restrict__ keys, uint64_t m, const uint64_t* __restrict__ m_dev,
                              uint32_t* __restrict__ ids) {
    if (m_dev) m = min(*m_dev, m);
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) ids[j] = (uint32_t)keys[j];
}

// ---------------------------------------------------------------- composite

__device__ __forceinline__ float dpp_f(float x, int ctrl_sel) {
    int r = 0;
    switch (ctrl_sel) {
        case 0: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false); break;
        case 1: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false); break;
        case 2: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xF, 0xF, false); break;
        default: r = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xF, 0xF, false); break;
    }
    return __int_as_float(r);
}

__device__ __forceinline__ float swap32_sum(float a, float b) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap16_sum(float a, float b) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Wave-wide sums of 8 floats by a transpose reduction: v_permlane32_swap (keep
// lane bit 5), v_permlane16_swap (bit 4), DPP row_mirror (bit 3), then sums
// over lane bits 0-2 (row_half_mirror, quad_perm); lanes with (lane & 7) == 0
// return the total of value ((lane>>5)&1)*4 + ((lane>>4)&1)*2 + ((lane>>3)&1).
__device__ __forceinline__ float wave_reduce8(float (&v)[8], int lane) {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = swap32_sum(v[k], v[k + 4]);
#pragma unroll
    for (int k = 0; k < 2; ++k) v[k] = swap16_sum(v[k], v[k + 2]);
    {
        const bool hi = lane & 8;
        float send = hi ? v[0] : v[1];
        float keep = hi ? v[1] : v[0];
        v[0] = keep + dpp_f(send, 3);
    }
    float x = v[0];
    x = x + dpp_f(x, 2);
    x = x + dpp_f(x, 0);
    x = x + dpp_f(x, 1);
    return x;
}

// Workgroup order of the composite.  Super-tiles of 2x2 tiles are ranked by
// their total list length, longest first, and dealt round-robin to the 8 XCDs
// (workgroup b runs on XCD b % 8); the 4 tiles of a super-tile go to the same
// XCD, where they share splat records in its L2.  order[8 * (4 q + j) + x] =
// tile j of the super-tile of rank 8 q + x, or ~0 (no tile).  The ranking is a
// counting sort on the length quantised to 1024 levels (shifted so the longest
// fits): super-tiles of one level take their places in arrival order, which
// changes only which workgroup runs which tile, never a result.  Above
// kOrderMax super-tiles the ranking is skipped.
constexpr uint32_t kOrderMax = 8192;
constexpr uint32_t kOrderLevels = 1024;
constexpr uint32_t kOrderTiles = 8192;  // tile lengths kept in LDS up to this many tiles

// kFromCounts: src = per-tile list lengths, and the kernel first writes the
// tile ranges with the tiles taken column by column (the binning's output
// order: [start, end) of a tile = the lengths of the tiles before it in
// column-major order); else src = the ranges.
// n / d and n % d for n < 2^32 with m = floor((2^32 - 1) / d): umulhi(n, m)
// is the quotient or one less, fixed by one comparison.
__device__ __forceinline__ void divmod_u32(uint32_t n, uint32_t d, uint32_t m, uint32_t& q, uint32_t& r) {
    q = __umulhi(n, m);
    r = n - q * d;
    if (r >= d) {
        ++q;
        r -= d;
    }
}

template <bool kFromCounts>
__global__ __launch_bounds__(1024) void k_tile_order(const uint32_t* __restrict__ src, uint32_t tiles_x,
                                                     uint32_t tiles_y, uint32_t* __restrict__ order,
                                                     uint32_t* __restrict__ ranges_out) {
    __shared__ uint32_t s_rank[kOrderMax];  // super-tile of each rank
    __shared__ uint32_t s_cnt[kOrderLevels];
    __shared__ uint32_t s_wsum[16];
    __shared__ uint32_t s_max;
    __shared__ uint32_t s_len[kFromCounts ? kOrderTiles : 1];  // tile lengths (up to kOrderTiles tiles)
    const uint32_t sx = (tiles_x + 1u) / 2u, sy = (tiles_y + 1u) / 2u, S = sx * sy;
    const uint32_t msx = 0xFFFFFFFFu / sx;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wid = tid >> 6;
    const uint32_t ntiles = tiles_x * tiles_y;
    const bool in_lds = kFromCounts && ntiles <= kOrderTiles;
    if (kFromCounts) {
        // list position c = tx * tiles_y + ty; thread tid owns positions [per tid, per tid + per)
        const uint32_t per = (ntiles + 1023u) / 1024u;
        const uint32_t c0 = tid * per;
        const uint32_t tx0 = c0 / tiles_y, ty0 = c0 - tx0 * tiles_y;
        constexpr uint32_t kFew = 8;  // up to 8192 tiles: all loads issued at once
        uint32_t few[kFew];
        uint32_t own = 0u;
        if (per <= kFew) {
            uint32_t tx = tx0, ty = ty0;
#pragma unroll
            for (uint32_t k = 0; k < kFew; ++k) {
                const uint32_t c = c0 + k;
                few[k] = (k < per && c < ntiles) ? src[ty * tiles_x + tx] : 0u;
                own += few[k];
                if (++ty == tiles_y) {
                    ty = 0u;
                    ++tx;
                }
            }
        } else {
            uint32_t tx = tx0, ty = ty0;
            for (uint32_t k = 0; k < per; ++k) {
                if (c0 + k < ntiles) own += src[ty * tiles_x + tx];
                if (++ty == tiles_y) {
                    ty = 0u;
                    ++tx;
                }
            }
        }
        uint32_t x = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if ((int)lane >= o) x += y;
        }
        if (lane == 63u) s_wsum[wid] = x;
        __syncthreads();
        uint32_t run = x - own;
        for (uint32_t w = 0; w < wid; ++w) run += s_wsum[w];
        uint32_t tx = tx0, ty = ty0;
        for (uint32_t k = 0; k < per; ++k) {
            if (c0 + k < ntiles) {
                const uint32_t t = ty * tiles_x + tx;
                const uint32_t len = per <= kFew ? few[k < kFew ? k : 0] : src[t];
                ranges_out[2 * t] = run;
                ranges_out[2 * t + 1] = run + len;
                if (in_lds) s_len[t] = len;
                run += len;
            }
            if (++ty == tiles_y) {
                ty = 0u;
                ++tx;
            }
        }
        __syncthreads();
    }
    auto tile_len = [&](uint32_t t) -> uint32_t {
        if (kFromCounts) return in_lds ? s_len[t] : src[t];
        return src[2 * t + 1] - src[2 * t];
    };
    const bool rank = S <= kOrderMax;
    if (rank) {
        constexpr uint32_t kPer = kOrderMax / 1024u;
        uint32_t cost[kPer];
        uint32_t cmax = 0u;
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t sidx = tid + 1024u * q;
            uint32_t c = 0u;
            if (sidx < S) {
                uint32_t sty, stx;
                divmod_u32(sidx, sx, msx, sty, stx);
                for (uint32_t j = 0; j < 4u; ++j) {
                    const uint32_t tx = 2u * stx + (j & 1u), ty = 2u * sty + (j >> 1);
                    if (tx < tiles_x && ty < tiles_y) c += tile_len(ty * tiles_x + tx);
                }
            }
            cost[q] = c;
            cmax = max(cmax, c);
        }
        s_cnt[tid] = 0u;
        if (tid == 0) s_max = 0u;
        __syncthreads();
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) cmax = max(cmax, (uint32_t)__shfl_xor((int)cmax, o, 64));
        if (lane == 0u) atomicMax(&s_max, cmax);  // one LDS atomic per wave, not per thread
        __syncthreads();
        uint32_t shift = 0u;
        while ((s_max >> shift) >= kOrderLevels) ++shift;
        // Super-tiles of the last level (no work, common at the frame's edges) are
        // counted and placed with one LDS atomic per wave, the others per thread.
        const uint64_t lt = (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
        uint32_t level[kPer];  // 0 = longest
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            level[q] = kOrderLevels - 1u - (cost[q] >> shift);
            const bool in = tid + 1024u * q < S;
            const bool last_lv = in && level[q] == kOrderLevels - 1u;
            const uint64_t bl = __ballot(last_lv);
            if (in && !last_lv) atomicAdd(&s_cnt[level[q]], 1u);
            if (bl && lane == (uint32_t)__builtin_ctzll(bl)) atomicAdd(&s_cnt[kOrderLevels - 1u], (uint32_t)__popcll(bl));
        }
        __syncthreads();
        // exclusive scan of the 1024 level counts (one per thread)
        const uint32_t v = s_cnt[tid];
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if ((int)lane >= o) x += y;
        }
        if (lane == 63u) s_wsum[wid] = x;
        __syncthreads();
        uint32_t base = 0u;
        for (uint32_t w = 0; w < wid; ++w) base += s_wsum[w];
        s_cnt[tid] = base + x - v;
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const bool in = tid + 1024u * q < S;
            const bool last_lv = in && level[q] == kOrderLevels - 1u;
            const uint64_t bl = __ballot(last_lv);
            uint32_t first = 0u;
            if (bl && lane == (uint32_t)__builtin_ctzll(bl))
                first = atomicAdd(&s_cnt[kOrderLevels - 1u], (uint32_t)__popcll(bl));
            first = (uint32_t)__shfl((int)first, bl ? (int)__builtin_ctzll(bl) : 0, 64);
            if (in && !last_lv) s_rank[atomicAdd(&s_cnt[level[q]], 1u)] = tid + 1024u * q;
            if (last_lv) s_rank[first + (uint32_t)__popcll(bl & lt)] = tid + 1024u * q;
        }
        __syncthreads();
    }
    const uint32_t R = (S + 7u) / 8u * 8u;
    for (uint32_t r = tid; r < R; r += blockDim.x) {
        const uint32_t q = r / 8u, x = r % 8u;
        const uint32_t sidx = r < S ? (rank ? s_rank[r] : r) : ~0u;
        uint32_t sty = 0u, stx = 0u;
        if (sidx != ~0u) divmod_u32(sidx, sx, msx, sty, stx);
        for (uint32_t j = 0; j < 4u; ++j) {
            uint32_t t = ~0u;
            if (sidx != ~0u) {
                const uint32_t tx = 2u * stx + (j & 1u), ty = 2u * sty + (j >> 1);
                if (tx < tiles_x && ty < tiles_y) t = ty * tiles_x + tx;
            }
            order[8u * (4u * q + j) + x] = t;
        }
    }
}

// Composite of one TILE x TILE tile (one workgroup, one thread per pixel, each
// wave an 8x8 pixel block), records staged in LDS BATCH at a time.
//   forward : front to back, alphas of kFwdGroup entries at a time (independent), then
//             the transmittance chain with early termination; T checkpointed
//             every 8 entries ([slot][chunk][pixel]); per wave one bit per entry:
//             did any of its pixels composite it;
//   reverse : batches and 8-entry chunks backwards, chunks without a bit
//             skipped, alpha and T recomputed from the checkpoint, 9 FP32
//             per-pair gradients summed over the wave, the 4..16 wave partials
//             of an entry summed in FP64 and added to the splat's 2D gradients.
// Forward: entries whose alphas are computed together (independent LDS reads and
// arithmetic); 4 keeps the kernel at 5 waves per SIMD, where 8 needs 128 VGPRs (4).
constexpr int kFwdGroup = 4;
template <int TILE, int BATCH>
__global__ __launch_bounds__(TILE * TILE) void k_composite_ext(ViewParams v, ExtCompositeArgs A) {
    constexpr int NT = TILE * TILE;            // threads = pixels
    constexpr int NW = NT / 64;                // waves
    constexpr int BW = TILE / 8;               // wave blocks per tile row
    constexpr int MW = (BATCH + 63) / 64;      // 64-bit mask words per batch
    constexpr int NCH = BATCH / 8;             // chunks per batch
    constexpr int NSTAGE = (3 * BATCH + NT - 1) / NT;  // float4 record pieces per thread
    __shared__ ExtRec s_rec[BATCH];
    __shared__ uint32_t s_ids[BATCH];
    const float4* rec4 = reinterpret_cast<const float4*>(s_rec);  // 3 float4 per record
    __shared__ float s_part[NW][BATCH][9];
    __shared__ double s_red[2][NW];
    __shared__ int s_max[NW];
    __shared__ unsigned long long s_cnt[4][NW];

    const uint32_t tile = A.order[blockIdx.x];
    if (tile >= v.num_tiles) return;  // padding slot of the order (uniform per workgroup)
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t start = A.ranges[2 * tile], end = A.ranges[2 * tile + 1];
    const int len = (int)(end - start);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t px = tx * TILE + (wid % BW) * 8 + (lane & 7);
    const uint32_t py = ty * TILE + (wid / BW) * 8 + (lane >> 3);
    const bool valid = px < v.W && py < v.H;
    const float pxf = px_to_x(px, v.W, v.aspect);
    const float pyf = py_to_y(py, v.H);
    // the wave's 8x8 block: pixel-centre coordinates of its extreme columns and rows
    // (px_to_x grows with px, py_to_y falls with py)
    const uint32_t bx0 = tx * TILE + (wid % BW) * 8, by0 = ty * TILE + (wid / BW) * 8;
    const float blk_x0 = px_to_x(bx0, v.W, v.aspect), blk_x1 = px_to_x(bx0 + 7u, v.W, v.aspect);
    const float blk_y0 = py_to_y(by0 + 7u, v.H), blk_y1 = py_to_y(by0, v.H);
    // this tile's batch slots: [start/BATCH + q, ...), q the tile's place in the
    // order the lists are stored in (starts grow with q), never overlap another tile's
    const uint32_t qplace = A.col_major ? tx * v.tiles_y + ty : tile;
    const size_t slot0 = (size_t)(start / BATCH) + qplace;

    // Batch staging: piece q = tid + k*NT (< 3*BATCH) is float4 (q % 3) of record
    // q / 3.  The next batch's pieces are loaded into registers while the current
    // batch is composited, with the ids of the batch after that.
    auto load_id = [&](int b0, int k) -> uint32_t {
        const int q = tid + k * NT, j = q / 3;
        return (q < 3 * BATCH && b0 >= 0 && b0 + j < len) ? A.ids[start + (uint32_t)(b0 + j)] : 0u;
    };
    auto load_rec = [&](uint32_t id, int b0, int k) -> float4 {
        const int q = tid + k * NT, j = q / 3;
        return (q < 3 * BATCH && b0 >= 0 && b0 + j < len) ? reinterpret_cast<const float4*>(&A.rec[id])[q - 3 * j]
                                                          : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    uint32_t id_cur[NSTAGE], id_next[NSTAGE];
    float4 rec_cur[NSTAGE];
    auto put = [&](int nb) {
#pragma unroll
        for (int k = 0; k < NSTAGE; ++k) {
            const int q = tid + k * NT, j = q / 3;
            if (q < 3 * BATCH && j < nb) {
                reinterpret_cast<float4*>(&s_rec[j])[q - 3 * j] = rec_cur[k];
                if (q - 3 * j == 0) s_ids[j] = id_cur[k];
            }
        }
    };
    auto prefetch = [&](int b_next, int b_after) {
#pragma unroll
        for (int k = 0; k < NSTAGE; ++k) {
            id_cur[k] = id_next[k];
            rec_cur[k] = load_rec(id_cur[k], b_next, k);
            id_next[k] = load_id(b_after, k);
        }
    };
    auto prime = [&](int b_first, int b_second) {
#pragma unroll
        for (int k = 0; k < NSTAGE; ++k) {
            id_cur[k] = load_id(b_first, k);
            rec_cur[k] = load_rec(id_cur[k], b_first, k);
            id_next[k] = load_id(b_second, k);
        }
    };

    // ---------------- forward ----------------
    float T = 1.0f, Cr = 0.0f, Cg = 0.0f, Cb = 0.0f;
    int last = 0;
    bool alive = valid;
    uint32_t n_fw = 0u;    // entries the wave evaluated (uniform)
    uint32_t n_take = 0u;  // pairs the wave composited (uniform)
    int stop = 0;        // 1 + index of the entry that stopped this pixel (0: not stopped)
    int fw_extent = 0;   // entries staged by the forward (uniform)
    prime(0, BATCH);
    for (int b0 = 0, bi = 0; b0 < len; b0 += BATCH, ++bi) {
        if (!__syncthreads_or(alive ? 1 : 0)) break;
        const int nb = min(BATCH, len - b0);
        fw_extent = b0 + nb;
        put(nb);
        __syncthreads();
        prefetch(b0 + BATCH, b0 + 2 * BATCH);
        // Entries that cannot reach the wave's block: lane j of the wave tests
        // entry 64w + j against the block's pixel-centre rectangle, first the
        // 3-sigma box (the margin, 1e-5 of the coordinates' magnitude, is far
        // above the rounding of the pixel test |px - centre| <= extent), then the
        // level set r^2 <= rcut2 outside which the pixel test fails.  An entry
        // left out has alpha 0 on every pixel of the block.
        uint64_t hit[MW];
#pragma unroll
        for (int w = 0; w < MW; ++w) {
            const int j = 64 * w + lane;
            bool h = false;
            if (j < nb) {
                const ExtRec& f = s_rec[j];
                const float mx = 1e-5f * (fabsf(f.scr_x) + f.ext_x + 4.0f);
                const float my = 1e-5f * (fabsf(f.scr_y) + f.ext_y + 4.0f);
                h = (f.scr_x - f.ext_x <= blk_x1 + mx) && (f.scr_x + f.ext_x >= blk_x0 - mx) &&
                    (f.scr_y - f.ext_y <= blk_y1 + my) && (f.scr_y + f.ext_y >= blk_y0 - my);
                h = h && conic_reaches_rect(f.A, f.B, f.C, blk_x0 - f.scr_x, blk_x1 - f.scr_x, blk_y0 - f.scr_y,
                                            blk_y1 - f.scr_y, f.rcut2);
            }
            const uint64_t bal = __ballot(h);
            hit[w] = uniform_u64(bal);
        }
        uint64_t mask[MW];
#pragma unroll
        for (int w = 0; w < MW; ++w) mask[w] = 0ull;
        float* ck = A.ckpt + (slot0 + bi) * (size_t)(NCH * NT);
        // Only the entries that may hit are evaluated, kFwdGroup at a time in
        // ascending order (an entry left out changes no lane's state).  The transmittance
        // checkpoint of chunk c (entries 8c..8c+7) is T before the first entry
        // of the chunk, written when the walk first reaches the chunk; chunks
        // the walk passes over without an entry are written with the same T.
        int ck_next = 0;  // first chunk of the batch whose checkpoint is not written (uniform)
        bool wave_alive = __ballot(alive) != 0ull;
#pragma unroll
        for (int w = 0; w < MW; ++w) {
            uint64_t m = hit[w];
            while (m != 0ull && wave_alive) {
                int e[kFwdGroup];
#pragma unroll
                for (int j = 0; j < kFwdGroup; ++j) {
                    e[j] = m != 0ull ? 64 * w + (int)__builtin_ctzll(m) : -1;
                    m &= m - 1ull;
                }
                // alphas of the group's entries are independent: computed first,
                // branch-free (selects), so their LDS reads and arithmetic overlap
                float ca[kFwdGroup];
                bool ctk[kFwdGroup];
#pragma unroll
                for (int j = 0; j < kFwdGroup; ++j) {
                    const int q = e[j] < 0 ? 0 : e[j];
                    float G;
                    ctk[j] = ext_pair_eval(pxf, pyf, rec4[3 * q], rec4[3 * q + 1], rec4[3 * q + 2], e[j] >= 0, ca[j], G);
                }
                // then the front-to-back chain, entry by entry: a pixel that does
                // not composite the entry adds 0 to C, which leaves C unchanged bit
                // for bit (fma(0, x, C) = C)
#pragma unroll
                for (int j = 0; j < kFwdGroup; ++j) {
                    if (e[j] < 0) break;  // uniform
                    ++n_fw;
                    if (A.backward)
                        for (; ck_next <= (e[j] >> 3); ++ck_next) ck[ck_next * NT + tid] = T;
                    const float a = ca[j];
                    const bool take = alive & ctk[j];
                    const float tT = fmaf(-T, a, T);
                    const bool dies = take & (tT < kExtTMin);
                    const bool inc = take & !dies;
                    const float Ta = T * (inc ? a : 0.0f);
                    const float4 p1 = rec4[3 * e[j] + 1];
                    const float cb = rec4[3 * e[j] + 2].x;
                    Cr = fmaf(Ta, p1.z, Cr);
                    Cg = fmaf(Ta, p1.w, Cg);
                    Cb = fmaf(Ta, cb, Cb);
                    T = inc ? tT : T;
                    last = inc ? b0 + e[j] + 1 : last;
                    stop = dies ? b0 + e[j] + 1 : stop;
                    alive = alive & !dies;
                    const uint64_t bal = ballot_b(inc);
                    if (bal != 0ull) {
                        mask[e[j] >> 6] |= 1ull << (e[j] & 63);
                        n_take += (uint32_t)__popcll(bal);
                    }
                }
                wave_alive = __ballot(alive) != 0ull;
            }
        }
        if (A.backward && lane < MW) {
#pragma unroll
            for (int w = 0; w < MW; ++w)
                if (lane == w) A.act[((slot0 + bi) * NW + wid) * MW + w] = mask[w];
        }
    }
    Cr = Cr + T * kExtBg;
    Cg = Cg + T * kExtBg;
    Cb = Cb + T * kExtBg;

    // ---------------- image, loss, seeds ----------------
    float4 tg = make_float4(0.f, 0.f, 0.f, 0.f);
    double l1 = 0.0, l2 = 0.0;
    if (valid) {
        size_t pix = (size_t)py * v.W + px;
        tg = A.tgt_rgba[pix];
        A.c_img[3 * pix + 0] = Cr;
        A.c_img[3 * pix + 1] = Cg;
        A.c_img[3 * pix + 2] = Cb;
        double er = (double)Cr - (double)tg.x, eg = (double)Cg - (double)tg.y, eb = (double)Cb - (double)tg.z;
        l1 = (fabs(er) + fabs(eg)) + fabs(eb);
        l2 = (er * er + eg * eg) + eb * eb;
    }
    const float dCr = Cr > tg.x ? 1.0f : (Cr < tg.x ? -1.0f : 0.0f);
    const float dCg = Cg > tg.y ? 1.0f : (Cg < tg.y ? -1.0f : 0.0f);
    const float dCb = Cb > tg.z ? 1.0f : (Cb < tg.z ? -1.0f : 0.0f);
    int ml = valid ? last : 0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        l2 += __shfl_xor(l2, o, 64);
        l1 += __shfl_xor(l1, o, 64);
        ml = max(ml, __shfl_xor(ml, o, 64));
    }
    if (lane == 0) {
        s_red[0][wid] = l2;
        s_red[1][wid] = l1;
        s_max[wid] = ml;
    }
    __syncthreads();
    int maxlast = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) maxlast = max(maxlast, s_max[w]);

    // ---------------- reverse pass ----------------
    uint32_t n_went = 0u;
    if (A.backward) {
        float dT = (dCr * kExtBg + dCg * kExtBg) + dCb * kExtBg;
        const int nbr = (maxlast + BATCH - 1) / BATCH;
        prime((nbr - 1) * BATCH, (nbr - 2) * BATCH);
        // Global loads the walk waits on are issued ahead: the activity words of a
        // batch while the previous batch is processed, the checkpoint of the
        // batch's first active chunk before its records are staged, and that of
        // each next active chunk while the current one is processed.
        auto load_act = [&](int b, uint64_t (&m)[MW]) {
#pragma unroll
            for (int w = 0; w < MW; ++w) m[w] = b >= 0 ? A.act[((slot0 + b) * NW + wid) * MW + w] : 0ull;
        };
        uint64_t act_next[MW];
        load_act(nbr - 1, act_next);
        for (int bi = nbr - 1; bi >= 0; --bi) {
            const int b0 = bi * BATCH;
            const int nb = min(BATCH, len - b0);
            uint64_t mask[MW];
#pragma unroll
            for (int w = 0; w < MW; ++w)  // the same word in every lane
                mask[w] = uniform_u64(act_next[w]);
            load_act(bi - 1, act_next);
            uint32_t chunks = 0u;  // bit c: chunk c has an active entry (uniform)
#pragma unroll
            for (int c = 0; c < NCH; ++c)
                chunks |= (((uint32_t)(mask[c >> 3] >> (8 * (c & 7))) & 0xFFu) != 0u ? 1u : 0u) << c;
            const float* ck = A.ckpt + (slot0 + bi) * (size_t)(NCH * NT);
            int c = chunks ? 31 - __builtin_clz(chunks) : -1;
            float T_next = c >= 0 ? ck[c * NT + tid] : 1.0f;
            __syncthreads();
            put(nb);
            for (int e = tid; e < NW * BATCH * 9; e += NT) (&s_part[0][0][0])[e] = 0.0f;
            __syncthreads();
            prefetch(b0 - BATCH, b0 - 2 * BATCH);
            while (c >= 0) {  // active chunks, last first (uniform)
                chunks &= ~(1u << c);
                const int c_next = chunks ? 31 - __builtin_clz(chunks) : -1;
                float Tc = T_next;
                if (c_next >= 0) T_next = ck[c_next * NT + tid];
                const uint32_t bits = (uint32_t)(mask[c >> 3] >> (8 * (c & 7))) & 0xFFu;
                float ra[8], rT[8], rG[8];
                // an entry whose bit is 0 changed no lane's transmittance before
                // that lane's last composited entry (it would have been composited):
                // skipped.  Alpha and T are the forward's bit for bit (same
                // operations); a lane that did not composite the entry gets alpha
                // and G zero, which zeroes all of its terms below and leaves T
                // unchanged (fma(-T, 0, T) = T).
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    ra[j] = 0.0f;
                    rG[j] = 0.0f;
                    rT[j] = Tc;
                    if ((bits >> j) & 1u) {
                        const int e = b0 + 8 * c + j;
                        const int q = 8 * c + j;
                        float a, G;
                        const bool ok =
                            ext_pair_eval(pxf, pyf, rec4[3 * q], rec4[3 * q + 1], rec4[3 * q + 2], e < last, a, G);
                        ra[j] = ok ? a : 0.0f;
                        rG[j] = ok ? G : 0.0f;
                        Tc = fmaf(-Tc, ra[j], Tc);
                    }
                }
                // the opacity terms of the chunk's entries, reduced together after it
                float o8[8];
#pragma unroll
                for (int j = 7; j >= 0; --j) {
                    o8[j] = 0.0f;
                    if ((bits >> j) & 1u) {
                        const int jj = 8 * c + j;
                        n_went += 1u;
                        const float4 p0 = rec4[3 * jj];
                        const float4 p1 = rec4[3 * jj + 1];
                        const float4 p2 = rec4[3 * jj + 2];
                        const float fr = p1.z, fg = p1.w, fb = p2.x, fop = p1.y;
                        // per-pair gradient terms (only FP32 accuracy needed): with
                        // s = dC . rgb, dL/dalpha = T (s - dT), and the seed of the
                        // transmittance before the entry is dT + alpha (s - dT).
                        // The geometry enters through r^2 = d^T Q d only: with
                        // w = dL/dalpha G (= dL/dop) and dL/dr^2 = -op w / 2, the wave
                        // sums of w d and w d d^T give the centre and conic gradients
                        // once multiplied by the entry's op and conic (after the sum).
                        const float dx = pxf - p0.x, dy = pyf - p0.y;
                        const float Tk = rT[j], a = ra[j];
                        // no gradient through a clamped alpha
                        const float G = fop * rG[j] > kExtMaxAlpha ? 0.0f : rG[j];
                        const float u = fmaf(dCb, fb, fmaf(dCg, fg, dCr * fr)) - dT;
                        const float da = Tk * u;
                        dT = fmaf(a, u, dT);
                        const float w = da * G;
                        const float wx = w * dx, wy = w * dy;
                        const float Tka = Tk * a;
                        float gv[8];
                        gv[5] = dCr * Tka;
                        gv[6] = dCg * Tka;
                        gv[7] = dCb * Tka;
                        gv[0] = wx;
                        gv[1] = wy;
                        gv[2] = wx * dx;
                        gv[3] = wx * dy;
                        gv[4] = wy * dy;
                        o8[j] = w;
                        const float tot = wave_reduce8(gv, lane);
                        if ((lane & 7) == 0) {
                            int idx = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
                            s_part[wid][jj][idx] = tot;
                        }
                    }
                }
                {
                    // lanes with (lane & 7) == 0 hold the opacity total of entry
                    // 8c + k, k from lane bits 5, 4, 3 (zero for skipped entries)
                    const float tot = wave_reduce8(o8, lane);
                    const int k = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
                    if ((lane & 7) == 0 && ((bits >> k) & 1u)) s_part[wid][8 * c + k][8] = tot;
                }
                c = c_next;
            }
            __syncthreads();
            // per entry and term: the wave partials summed in FP64 in wave order;
            // slots 0-4 hold the sums of w dx, w dy, w dx^2, w dx dy, w dy^2, taken
            // to d scr (= -2 Q (dL/dr^2) d) and d conic (A, B, C) here
            for (int e = tid; e < nb * 9; e += NT) {
                const int j = e / 9, k = e - 9 * j;
                auto part_sum = [&](int kk) {
                    double s = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) s += (double)s_part[w][j][kk];
                    return s;
                };
                const double op = (double)s_rec[j].op;
                double sum;
                if (k == 0)
                    sum = op * ((double)s_rec[j].A * part_sum(0) + (double)s_rec[j].B * part_sum(1));
                else if (k == 1)
                    sum = op * ((double)s_rec[j].B * part_sum(0) + (double)s_rec[j].C * part_sum(1));
                else if (k == 2 || k == 4)
                    sum = -0.5 * op * part_sum(k);
                else if (k == 3)
                    sum = -op * part_sum(3);
                else
                    sum = part_sum(k);
                if (sum != 0.0) unsafeAtomicAdd(&A.g2[9 * (size_t)s_ids[j] + k], sum);
            }
        }
    }
    // counters of the tile, in its slot of A.part: pairs evaluated by the forward
    // (those before the pixel stopped, the stopping entry included), forward
    // wave-entries, pairs taken by the reverse pass, reverse wave-entries
    uint32_t ev = valid ? (uint32_t)(stop > 0 ? stop : len) : 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) ev += __shfl_xor(ev, o, 64);
    if (lane == 0) {
        s_cnt[0][wid] = ev;
        s_cnt[1][wid] = n_fw;
        s_cnt[2][wid] = A.backward ? n_take : 0u;  // = the pairs the reverse pass takes
        s_cnt[3][wid] = n_went;
    }
    __syncthreads();
    if (tid == 0) {
        double a2 = 0.0, a1 = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            a2 += s_red[0][w];
            a1 += s_red[1][w];
        }
        A.tile_loss[2 * tile] = a2;
        A.tile_loss[2 * tile + 1] = a1;
        A.tile_work[2 * tile] = (uint32_t)fw_extent;
        A.tile_work[2 * tile + 1] = (uint32_t)maxlast;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            unsigned long long c = 0ull;
#pragma unroll
            for (int w = 0; w < NW; ++w) c += s_cnt[k][w];
            A.part[4 * (size_t)tile + k] = c;
        }
    }
}

// ---------------------------------------------------------------- chain rule

// kStore: the gradient buffer is known to be all zero (the last Adam step
// cleared it): the splat's gradients are stored, not read and added.
// flags (nflags words, 16 apart; null = none): when any is set a view of the
// step dropped keys, the host redoes the step view by view, and this launch
// leaves gradients and densification statistics untouched.
// g2_clear (null = keep): the 2D gradients read are cleared, so g2 is all zero
// again for the next view (only splats with a tile count can have any).
template <int DEG, bool kStore>
__global__ __launch_bounds__(256) void k_ext_chain(const float4* __restrict__ splats, const float4* __restrict__ sh,
                                                   uint32_t n, ViewParams v, float cam0, float cam1, float cam2,
                                                   const double* __restrict__ g2, const uint32_t* __restrict__ counts,
                                                   float* __restrict__ grads, float2* __restrict__ accum,
                                                   const uint32_t* __restrict__ flags, uint32_t nflags,
                                                   double* __restrict__ g2_clear) {
    constexpr int kNk = (DEG + 1) * (DEG + 1);
    constexpr int kSh4 = (3 * kNk + 3) / 4;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (counts[i] == 0u) return;  // culled or off-screen: no pair, no gradient
    double gg[9];
    bool any = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        gg[k] = g2[9 * (size_t)i + k];
        any = any || gg[k] != 0.0;
    }
    if (g2_clear && any) {
#pragma unroll
        for (int k = 0; k < 9; ++k) g2_clear[9 * (size_t)i + k] = 0.0;
    }
    if (flags) {
        uint32_t bad = 0u;
        for (uint32_t k = 0; k < nflags; ++k) bad |= flags[16 * k];
        if (bad) return;
    }
    float s[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float4 t = splats[4 * (size_t)i + q];
        s[4 * q + 0] = t.x; s[4 * q + 1] = t.y; s[4 * q + 2] = t.z; s[4 * q + 3] = t.w;
    }
    const float cam[3] = {cam0, cam1, cam2};
    ExtProj o;
    ext_project(s, nullptr, DEG, v.vp, cam, v.W, v.H, o);  // geometry only: the colour follows below
    if (!o.valid) return;
    // densification statistics: |d L / d ndc| of the centre and the view count
    {
        double gx = gg[0] * (double)v.aspect, gy = gg[1];
        float2 acc = accum[i];
        acc.x = acc.x + (float)sqrt(gx * gx + gy * gy);
        acc.y = acc.y + 1.0f;
        accum[i] = acc;
    }
    if (!any) return;
    float* out = grads + 64 * (size_t)i;
    const float* vp = v.vp;
    // geometry first (opacity, scales, quaternion, and the position terms of
    // the projection); the SH part then needs only the view direction, so the
    // projection's values and the SH coefficients are never live together
    double G[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) G[k] = 0.0;
    G[3] = gg[8] * (double)(o.op * (1.0f - o.op));
    // conic -> covariance: dL/dSigma = -Q (dL/dQ) Q with dL/dQ symmetric
    const double A = o.A, B = o.B, C = o.C;
    const double gA = gg[2], gB = gg[3] * 0.5, gC = gg[4];
    const double t00 = A * gA + B * gB, t01 = A * gB + B * gC, t10 = B * gA + C * gB, t11 = B * gB + C * gC;
    const double s00 = -(t00 * A + t01 * B), s01 = -(t00 * B + t01 * C), s11 = -(t10 * B + t11 * C);
    const double gcov00 = s00, gcov01 = 2.0 * s01, gcov11 = s11;
    const double aspect = (double)v.aspect, w = o.w, cx = o.cx, cy = o.cy;
    // powers of 1/w once (FP64 divisions are long instruction sequences)
    const double iw = 1.0 / w, iw2 = iw * iw, iw3 = iw2 * iw;
    double gcx = gg[0] * aspect * iw;
    double gcy = gg[1] * iw;
    double gw = -gg[0] * cx * aspect * iw2 - gg[1] * cy * iw2;
    double gqn[4] = {0.0, 0.0, 0.0, 0.0};
    const double qw = o.qn[0], qx = o.qn[1], qy = o.qn[2], qz = o.qn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double mx = o.m[k][0], my = o.m[k][1];
        const double gmx = 2.0 * gcov00 * mx + gcov01 * my;
        const double gmy = 2.0 * gcov11 * my + gcov01 * mx;
        const double gs = gmx * (double)o.axs[k][0] + gmy * (double)o.axs[k][1];
        G[4 + k] += gs * (double)o.s[k];
        const double gax = gmx * (double)o.s[k], gay = gmy * (double)o.s[k];
        const float* a = o.axw[k];
        const double acx = (double)((vp[0] * a[0] + vp[4] * a[1]) + vp[8] * a[2]);
        const double acy = (double)((vp[1] * a[0] + vp[5] * a[1]) + vp[9] * a[2]);
        const double acw = (double)((vp[3] * a[0] + vp[7] * a[1]) + vp[11] * a[2]);
        const double gacx = gax * aspect * iw, gacy = gay * iw;
        const double gacw = -gax * aspect * cx * iw2 - gay * cy * iw2;
        gcx += -gax * aspect * acw * iw2;
        gcy += -gay * acw * iw2;
        gw += gax * aspect * (-acx * iw2 + 2.0 * cx * acw * iw3) + gay * (-acy * iw2 + 2.0 * cy * acw * iw3);
        double gaw[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
            gaw[j] = gacx * (double)vp[4 * j + 0] + gacy * (double)vp[4 * j + 1] + gacw * (double)vp[4 * j + 3];
        // partial derivatives of column k of R(q) (polynomial form)
        double dRw[3], dRx[3], dRy[3], dRz[3];
        if (k == 0) {
            dRw[0] = 0; dRw[1] = 2 * qz; dRw[2] = -2 * qy;
            dRx[0] = 0; dRx[1] = 2 * qy; dRx[2] = 2 * qz;
            dRy[0] = -4 * qy; dRy[1] = 2 * qx; dRy[2] = -2 * qw;
            dRz[0] = -4 * qz; dRz[1] = 2 * qw; dRz[2] = 2 * qx;
        } else if (k == 1) {
            dRw[0] = -2 * qz; dRw[1] = 0; dRw[2] = 2 * qx;
            dRx[0] = 2 * qy; dRx[1] = -4 * qx; dRx[2] = 2 * qw;
            dRy[0] = 2 * qx; dRy[1] = 0; dRy[2] = 2 * qz;
            dRz[0] = -2 * qw; dRz[1] = -4 * qz; dRz[2] = 2 * qy;
        } else {
            dRw[0] = 2 * qy; dRw[1] = -2 * qx; dRw[2] = 0;
            dRx[0] = 2 * qz; dRx[1] = -2 * qw; dRx[2] = -4 * qx;
            dRy[0] = 2 * qw; dRy[1] = 2 * qz; dRy[2] = -4 * qy;
            dRz[0] = 2 * qx; dRz[1] = 2 * qy; dRz[2] = 0;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            gqn[0] += gaw[j] * dRw[j];
            gqn[1] += gaw[j] * dRx[j];
            gqn[2] += gaw[j] * dRy[j];
            gqn[3] += gaw[j] * dRz[j];
        }
    }
    const double dq = gqn[0] * qw + gqn[1] * qx + gqn[2] * qy + gqn[3] * qz;
    const double iql = 1.0 / (double)o.qlen;
#pragma unroll
    for (int j = 0; j < 4; ++j) G[8 + j] += (gqn[j] - dq * (double)o.qn[j]) * iql;
#pragma unroll
    for (int j = 0; j < 3; ++j)
        G[j] += gcx * (double)vp[4 * j + 0] + gcy * (double)vp[4 * j + 1] + gw * (double)vp[4 * j + 3];
#pragma unroll
    for (int j = 3; j < 12; ++j)
        if (j != 7) out[j] = kStore ? (float)G[j] : out[j] + (float)G[j];
    // SH colour: forward value (its sign gates the channel, as in the
    // projection), coefficient gradients, and the view direction's gradient
    const float dir0 = o.dir[0], dir1 = o.dir[1], dir2 = o.dir[2], dlen = o.dlen;
    float c[4 * kSh4];
#pragma unroll
    for (int q = 0; q < kSh4; ++q) {
        float4 t = sh[12 * (size_t)i + q];
        c[4 * q + 0] = t.x; c[4 * q + 1] = t.y; c[4 * q + 2] = t.z; c[4 * q + 3] = t.w;
    }
    float Y[16];
    sh_basis(dir0, dir1, dir2, DEG, Y);
    // per coefficient k: sum over the drawn channels of dL/drgb * coefficient,
    // which the direction gradient then weights with dY_k/d dir
    double wk[kNk];
#pragma unroll
    for (int k = 0; k < kNk; ++k) wk[k] = 0.0;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float raw = 0.0f;
#pragma unroll
        for (int k = 0; k < kNk; ++k) raw = raw + Y[k] * c[3 * k + ch];
        raw = raw + 0.5f;
        if (!(raw > 0.0f)) continue;
        const double drgb = gg[5 + ch];
#pragma unroll
        for (int k = 0; k < kNk; ++k) {
            // one term per SH coefficient: written directly
            float& gk = out[16 + 3 * k + ch];
            if (kStore) gk = (float)(drgb * (double)Y[k]);
            else gk += (float)(drgb * (double)Y[k]);
            wk[k] += drgb * (double)c[3 * k + ch];
        }
    }
    double gdir[3] = {0.0, 0.0, 0.0};
    sh_grad_weighted(dir0, dir1, dir2, DEG, wk, gdir);
    const double dd = gdir[0] * dir0 + gdir[1] * dir1 + gdir[2] * dir2;
    const double dirv[3] = {dir0, dir1, dir2};
    const double idl = 1.0 / (double)dlen;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double gpos = (gdir[j] - dd * dirv[j]) * idl;
        out[j] = kStore ? (float)(G[j] + gpos) : out[j] + (float)(G[j] + gpos);
    }
}

// ---------------------------------------------------------------- Adam

// Learning rate of parameter p of the 64 (record then SH): means 1.6e-4,
// opacity logit 5e-2, log-scales 5e-3, quaternion 1e-3, SH DC 2.5e-3, higher
// SH bands 2.5e-3 / 20; 0 (left alone) for the pads and SH above the degree.
__device__ __forceinline__ float ext_lr(int p, int deg) {
    if (p < 3) return 1.6e-4f;
    if (p == 3) return 5e-2f;
    if (p >= 4 && p < 7) return 5e-3f;
    if (p >= 8 && p < 12) return 1e-3f;
    if (p >= 16) {
        int k = (p - 16) / 3;
        if (k >= (deg + 1) * (deg + 1)) return 0.0f;
        return k == 0 ? 2.5e-3f : 2.5e-3f / 20.0f;
    }
    return 0.0f;
}

// Adam's step counter lives on the device: ctl = {t, 1 - b1^t, 1 - b2^t, skip}.
// One thread: unless a flag (one per 16 words) reports a dropped key, t += 1
// and the bias corrections are recomputed; otherwise the step is skipped.
__global__ void k_adam_ext_prep(float* __restrict__ ctl, const uint32_t* __restrict__ flags, uint32_t nflags) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t bad = 0u;
    for (uint32_t k = 0; k < nflags; ++k) bad |= flags[16 * k];
    if (bad) {
        ctl[3] = 1.0f;
        return;
    }
    const float t = ctl[0] + 1.0f;
    ctl[0] = t;
    ctl[1] = 1.0f - pow_det(0.9f, t);
    ctl[2] = 1.0f - pow_det(0.999f, t);
    ctl[3] = 0.0f;
}

// One thread per float4 of the 64 parameters: reads parameter + grad + m + v
// (64 B), writes parameter + m + v and a zeroed grad (64 B).
__global__ __launch_bounds__(256) void k_adam_ext(float4* __restrict__ splats, float4* __restrict__ sh,
                                                  float4* __restrict__ grads, float4* __restrict__ m,
                                                  float4* __restrict__ v, uint32_t n, int deg,
                                                  const float* __restrict__ ctl) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 16 * (size_t)n) return;
    if (ctl[3] != 0.0f) return;
    const float bc1 = ctl[1], bc2 = ctl[2];
    const size_t i = idx >> 4;
    const int part = (int)(idx & 15);
    const float lr0 = ext_lr(4 * part, deg);
    const float lr3 = ext_lr(4 * part + 3, deg);
    if (lr0 == 0.0f && lr3 == 0.0f && ext_lr(4 * part + 1, deg) == 0.0f && ext_lr(4 * part + 2, deg) == 0.0f) {
        grads[idx] = make_float4(0.f, 0.f, 0.f, 0.f);  // nothing to update; the step still clears
        return;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-15f;
    float4* xp = part < 4 ? &splats[4 * i + part] : &sh[12 * i + (part - 4)];
    float4 x4 = *xp, g4 = grads[idx], m4 = m[idx], v4 = v[idx];
    float x[4] = {x4.x, x4.y, x4.z, x4.w}, g[4] = {g4.x, g4.y, g4.z, g4.w};
    float mm[4] = {m4.x, m4.y, m4.z, m4.w}, vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const float lr = ext_lr(4 * part + l, deg);
        if (lr == 0.0f) continue;
        float m1 = b1 * mm[l] + (1.0f - b1) * g[l];
        float v1 = b2 * vv[l] + ((1.0f - b2) * g[l]) * g[l];
        mm[l] = m1;
        vv[l] = v1;
        float upd = (lr * (m1 / bc1)) / (sqrtf(v1 / bc2) + eps);
        x[l] = x[l] - upd;
    }
    *xp = make_float4(x[0], x[1], x[2], x[3]);
    m[idx] = make_float4(mm[0], mm[1], mm[2], mm[3]);
    v[idx] = make_float4(vv[0], vv[1], vv[2], vv[3]);
    grads[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------- densification

namespace {
constexpr float kExtLn16 = 0.47000363f;     // ln 1.6: log-scale reduction of a split
constexpr float kExtDeadLogit = -10000.0f;  // parked slot: opacity 0

// Wellons' lowbias32 integer hash
__device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// standard normal sample k of child c of a split landing in `slot`: Box-Muller
// on two hashed uniforms, deterministic functions only
__device__ __forceinline__ float adc_normal(uint32_t seed, uint32_t slot, uint32_t child, uint32_t k) {
    uint32_t h1 = lowbias32(seed ^ lowbias32(slot * 8u + child * 4u + k));
    uint32_t h2 = lowbias32(h1 ^ 0x9e3779b9u);
    float u1 = (float)((h1 >> 8) + 1u) * 0x1p-24f;
    float u2 = (float)(h2 >> 8) * 0x1p-24f;
    float r = sqrtf(-2.0f * log_det(u1));
    return r * sin_det(6.2831855f * u2 + 1.5707964f);
}
}  // namespace

// dead[i] = sigmoid(logit) < prune; qual[i] = live, seen (count > 0) and
// summed |dL/d ndc| / count >= threshold
__global__ __launch_bounds__(256) void k_adc_ext_mark(const float* __restrict__ splats,
                                                      const float2* __restrict__ accum, uint32_t n, float threshold,
                                                      float prune, uint32_t* __restrict__ dead,
                                                      uint32_t* __restrict__ qual) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float op = 1.0f / (1.0f + exp_det(-splats[16 * (size_t)i + 3]));
    uint32_t d = op < prune ? 1u : 0u;
    uint32_t q = 0u;
    if (!d) {
        float2 a = accum[i];
        q = (a.y > 0.0f && a.x / a.y >= threshold) ? 1u : 0u;
    }
    dead[i] = d;
    qual[i] = q;
}

// dead_list[rank] = slot, qual_list[rank] = slot; cnt[0] = qualifiers, cnt[1] = dead slots
__global__ __launch_bounds__(256) void k_adc_ext_lists(uint32_t n, const uint32_t* __restrict__ dead,
                                                       const uint32_t* __restrict__ dead_rank,
                                                       const uint32_t* __restrict__ qual,
                                                       const uint32_t* __restrict__ qual_rank,
                                                       uint32_t* __restrict__ dead_list,
                                                       uint32_t* __restrict__ qual_list, uint32_t* __restrict__ cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dead[i]) dead_list[dead_rank[i]] = i;
    if (qual[i]) qual_list[qual_rank[i]] = i;
    if (i == n - 1) {
        cnt[0] = qual_rank[i] + qual[i];
        cnt[1] = dead_rank[i] + dead[i];
    }
}

// Thread k < min(qualifiers, dead): the k-th placement; min <= k < dead: parks
// the k-th dead slot.  Parents are read whole before any write; a parent slot is
// written only by its own thread and dead slots are never parents.  The Adam
// state zeroing and the SH copy are left to k_adc_ext_state (coalesced);
// split[k] records whether placement k split its parent.
__global__ __launch_bounds__(256) void k_adc_ext_apply(uint32_t n, float4* __restrict__ splats,
                                                       const uint32_t* __restrict__ dead_list,
                                                       const uint32_t* __restrict__ qual_list,
                                                       uint32_t* __restrict__ cnt, uint32_t* __restrict__ split,
                                                       float split_scale, uint32_t seed) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nq = cnt[0], nd = cnt[1];
    const uint32_t P = nq < nd ? nq : nd;
    if (k >= nd) return;
    const uint32_t d = dead_list[k];
    if (k >= P) {
        float4 r0 = splats[4 * (size_t)d];
        r0.w = kExtDeadLogit;
        splats[4 * (size_t)d] = r0;
        return;
    }
    const uint32_t q = qual_list[k];
    float p[16];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float4 t = splats[4 * (size_t)q + j];
        p[4 * j] = t.x; p[4 * j + 1] = t.y; p[4 * j + 2] = t.z; p[4 * j + 3] = t.w;
    }
    float sc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) sc[a] = exp_det(p[4 + a]);
    float smax = sc[0] > sc[1] ? sc[0] : sc[1];
    smax = smax > sc[2] ? smax : sc[2];
    if (smax > split_scale) {
        float qq = ((p[8] * p[8] + p[9] * p[9]) + p[10] * p[10]) + p[11] * p[11];
        float ql = sqrtf(qq > 1e-24f ? qq : 1e-24f);
        float qw = p[8] / ql, qx = p[9] / ql, qy = p[10] / ql, qz = p[11] / ql;
        const float ax[3][3] = {
            {1.0f - 2.0f * (qy * qy + qz * qz), 2.0f * (qx * qy + qw * qz), 2.0f * (qx * qz - qw * qy)},
            {2.0f * (qx * qy - qw * qz), 1.0f - 2.0f * (qx * qx + qz * qz), 2.0f * (qy * qz + qw * qx)},
            {2.0f * (qx * qz + qw * qy), 2.0f * (qy * qz - qw * qx), 1.0f - 2.0f * (qx * qx + qy * qy)}};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const uint32_t slot = c == 0 ? q : d;
            float o[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) o[a] = adc_normal(seed, slot, (uint32_t)c, (uint32_t)a) * sc[a];
            float out[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) out[j] = p[j];
#pragma unroll
            for (int j = 0; j < 3; ++j) out[j] = p[j] + ((ax[0][j] * o[0] + ax[1][j] * o[1]) + ax[2][j] * o[2]);
#pragma unroll
            for (int a = 0; a < 3; ++a) out[4 + a] = p[4 + a] - kExtLn16;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                splats[4 * (size_t)slot + j] = make_float4(out[4 * j], out[4 * j + 1], out[4 * j + 2], out[4 * j + 3]);
        }
        split[k] = 1u;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            splats[4 * (size_t)d + j] = make_float4(p[4 * j], p[4 * j + 1], p[4 * j + 2], p[4 * j + 3]);
        split[k] = 0u;
    }
    wave_count2(smax > split_scale, &cnt[2], &cnt[3]);
}

// Adam state and SH of the densification, one thread per float4 (grid-stride,
// t = 16 k + j): every dead-list slot's m and v are zeroed (children and parked
// slots), a split parent's too, and placement k's child takes its parent's SH.
__global__ __launch_bounds__(256) void k_adc_ext_state(float4* __restrict__ sh, float4* __restrict__ m,
                                                       float4* __restrict__ v, const uint32_t* __restrict__ dead_list,
                                                       const uint32_t* __restrict__ qual_list,
                                                       const uint32_t* __restrict__ cnt,
                                                       const uint32_t* __restrict__ split) {
    const uint32_t nq = cnt[0], nd = cnt[1];
    const uint32_t P = nq < nd ? nq : nd;
    const uint64_t total = 16ull * nd;
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(t >> 4), j = (uint32_t)(t & 15u);
        const uint32_t d = dead_list[k];
        m[16 * (size_t)d + j] = z4;
        v[16 * (size_t)d + j] = z4;
        if (k < P) {
            const uint32_t q = qual_list[k];
            if (j < 12u) sh[12 * (size_t)d + j] = sh[12 * (size_t)q + j];
            if (split[k]) {
                m[16 * (size_t)q + j] = z4;
                v[16 * (size_t)q + j] = z4;
            }
        }
    }
}

void run_adc_ext(hipStream_t st, float* splats, float* sh, float* m, float* v, float* accum, uint32_t n,
                 const AdcExtParams& prm, uint32_t seed, const AdcExtBuffers& b) {
    launch_fill_u32(st, b.cnt, 0u, 4);
    if (n == 0) return;
    const dim3 g((n + 255) / 256), blk(256);
    hipLaunchKernelGGL(k_adc_ext_mark, g, blk, 0, st, (const float*)splats, (const float2*)accum, n, prm.threshold,
                       prm.prune_opacity, b.dead, b.qual);
    exclusive_scan_u32(st, b.dead, b.dead_rank, n, b.scan_tmp);
    exclusive_scan_u32(st, b.qual, b.qual_rank, n, b.scan_tmp);
    hipLaunchKernelGGL(k_adc_ext_lists, g, blk, 0, st, n, (const uint32_t*)b.dead, (const uint32_t*)b.dead_rank,
                       (const uint32_t*)b.qual, (const uint32_t*)b.qual_rank, b.dead_list, b.qual_list, b.cnt);
    // the qualifier flags are not read after the lists: they hold the split flags
    hipLaunchKernelGGL(k_adc_ext_apply, g, blk, 0, st, n, (float4*)splats, (const uint32_t*)b.dead_list,
                       (const uint32_t*)b.qual_list, b.cnt, b.qual, prm.percent_dense * prm.extent, seed);
    hipLaunchKernelGGL(k_adc_ext_state, dim3(2048), blk, 0, st, (float4*)sh, (float4*)m, (float4*)v,
                       (const uint32_t*)b.dead_list, (const uint32_t*)b.qual_list, (const uint32_t*)b.cnt,
                       (const uint32_t*)b.qual);
    launch_fill_u32(st, (uint32_t*)accum, 0u, 2 * (uint64_t)n);
}

// Features of the qualifying parents, row qual_rank[i] (the densification's
// placement order): mean (0-2), world covariance xx xy xz yy yz zz (3-8),
// opacity (9), DC colour (10-12), mean |dL/d ndc| (13), view count (14),
// log-scales (15-17), largest scale (18), largest scale / split scale (19),
// split flag (20), opacity logit (21), summed |dL/d ndc| (22), distance to the
// current camera and its inverse (23, 24), and the direction of the parent's
// next Adam steps, m / (sqrt(v) + 1e-12), for the mean (25-27), the log-scales
// (28-30) and the opacity logit (31).
__global__ __launch_bounds__(256) void k_adc_ext_features(const float* __restrict__ splats,
                                                          const float* __restrict__ sh,
                                                          const float2* __restrict__ accum,
                                                          const float* __restrict__ m, const float* __restrict__ v,
                                                          uint32_t n, const uint32_t* __restrict__ qual,
                                                          const uint32_t* __restrict__ qual_rank, float split_scale,
                                                          float cx, float cy, float cz, float* __restrict__ feat) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !qual[i]) return;
    const float* p = splats + 16 * (size_t)i;
    float f[32];
    f[0] = p[0];
    f[1] = p[1];
    f[2] = p[2];
    float qq = ((p[8] * p[8] + p[9] * p[9]) + p[10] * p[10]) + p[11] * p[11];
    float ql = sqrtf(qq > 1e-24f ? qq : 1e-24f);
    const float qw = p[8] / ql, qx = p[9] / ql, qy = p[10] / ql, qz = p[11] / ql;
    const float R[3][3] = {{1.0f - 2.0f * (qy * qy + qz * qz), 2.0f * (qx * qy - qw * qz), 2.0f * (qx * qz + qw * qy)},
                           {2.0f * (qx * qy + qw * qz), 1.0f - 2.0f * (qx * qx + qz * qz), 2.0f * (qy * qz - qw * qx)},
                           {2.0f * (qx * qz - qw * qy), 2.0f * (qy * qz + qw * qx), 1.0f - 2.0f * (qx * qx + qy * qy)}};
    float s2[3], sc[3];
    for (int a = 0; a < 3; ++a) {
        sc[a] = __expf(p[4 + a]);
        s2[a] = sc[a] * sc[a];
    }
    int idx = 3;
    for (int r = 0; r < 3; ++r)
        for (int c = r; c < 3; ++c) f[idx++] = (R[r][0] * R[c][0] * s2[0] + R[r][1] * R[c][1] * s2[1]) + R[r][2] * R[c][2] * s2[2];
    f[9] = 1.0f / (1.0f + __expf(-p[3]));
    for (int c = 0; c < 3; ++c) f[10 + c] = sh[48 * (size_t)i + c] * 0.28209479177387814f + 0.5f;
    const float2 a = accum[i];
    f[13] = a.y > 0.0f ? a.x / a.y : 0.0f;
    f[14] = a.y;
    for (int c = 0; c < 3; ++c) f[15 + c] = p[4 + c];
    const float smax = fmaxf(fmaxf(sc[0], sc[1]), sc[2]);
    f[18] = smax;
    f[19] = smax / split_scale;
    f[20] = smax > split_scale ? 1.0f : 0.0f;
    f[21] = p[3];
    f[22] = a.x;
    const float dx = p[0] - cx, dy = p[1] - cy, dz = p[2] - cz;
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    f[23] = dist;
    f[24] = 1.0f / fmaxf(dist, 1e-6f);
    const int cols[7] = {0, 1, 2, 4, 5, 6, 3};
    for (int c = 0; c < 7; ++c) {
        const size_t k = 64 * (size_t)i + cols[c];
        f[25 + c] = m[k] / (sqrtf(v[k]) + 1e-12f);
    }
    float* o = feat + 32 * (size_t)qual_rank[i];
    for (int c = 0; c < 32; ++c) o[c] = f[c];
}

void launch_adc_ext_features(hipStream_t st, const float* splats, const float* sh, const float* accum,
                             const float* m, const float* v, uint32_t n, const AdcExtParams& prm, const float cam[3],
                             const AdcExtBuffers& b, float* feat) {
    launch_fill_u32(st, b.cnt, 0u, 4);
    if (n == 0) return;
    const dim3 g((n + 255) / 256), blk(256);
    hipLaunchKernelGGL(k_adc_ext_mark, g, blk, 0, st, splats, (const float2*)accum, n, prm.threshold,
                       prm.prune_opacity, b.dead, b.qual);
    exclusive_scan_u32(st, b.qual, b.qual_rank, n, b.scan_tmp);
    exclusive_scan_u32(st, b.dead, b.dead_rank, n, b.scan_tmp);
    hipLaunchKernelGGL(k_adc_ext_lists, g, blk, 0, st, n, (const uint32_t*)b.dead, (const uint32_t*)b.dead_rank,
                       (const uint32_t*)b.qual, (const uint32_t*)b.qual_rank, b.dead_list, b.qual_list, b.cnt);
    if (feat)
        hipLaunchKernelGGL(k_adc_ext_features, g, blk, 0, st, splats, sh, (const float2*)accum, m, v, n,
                           (const uint32_t*)b.qual, (const uint32_t*)b.qual_rank, prm.percent_dense * prm.extent,
                           cam[0], cam[1], cam[2], feat);
}

// Child k < min(qualifiers, dead) of the last densification: child 0 (parent
// slot) and child 1 (dead slot) get position, log-scale and opacity-logit deltas.
__global__ __launch_bounds__(256) void k_adc_ext_deltas(float* __restrict__ splats,
                                                        const uint32_t* __restrict__ dead_list,
                                                        const uint32_t* __restrict__ qual_list,
                                                        const uint32_t* __restrict__ cnt,
                                                        const float* __restrict__ delta, uint32_t nrows) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t P = cnt[0] < cnt[1] ? cnt[0] : cnt[1];
    if (k >= P || k >= nrows) return;
    const uint32_t slots[2] = {qual_list[k], dead_list[k]};
    const float* dk = delta + 14 * (size_t)k;
    for (int c = 0; c < 2; ++c) {
        float* s = splats + 16 * (size_t)slots[c];
        for (int j = 0; j < 3; ++j) s[j] += dk[7 * c + j];
        for (int j = 0; j < 3; ++j) s[4 + j] += dk[7 * c + 3 + j];
        s[3] += dk[7 * c + 6];
    }
}

void launch_adc_ext_deltas(hipStream_t st, float* splats, const AdcExtBuffers& b, const float* delta, uint32_t nrows) {
    if (nrows == 0) return;
    hipLaunchKernelGGL(k_adc_ext_deltas, dim3((nrows + 255) / 256), dim3(256), 0, st, splats,
                       (const uint32_t*)b.dead_list, (const uint32_t*)b.qual_list, (const uint32_t*)b.cnt, delta,
                       nrows);
}

// out[0..1] = sums of the per-tile (L2, L1) pairs, in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void k_sum_tile_loss(const double* __restrict__ tl, uint32_t ntiles,
                                                       double* __restrict__ out) {
    __shared__ double s[2][256];
    double a = 0.0, b = 0.0;
    for (uint32_t t = threadIdx.x; t < ntiles; t += 256u) {
        a += tl[2 * t];
        b += tl[2 * t + 1];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s[0][0];
        out[1] = s[1][0];
    }
}

void launch_sum_tile_loss(hipStream_t st, const double* tile_loss, uint32_t ntiles, double* out2) {
    hipLaunchKernelGGL(k_sum_tile_loss, dim3(1), dim3(256), 0, st, tile_loss, ntiles, out2);
}

// ---------------------------------------------------------------- launchers

void launch_ext_preprocess(hipStream_t st, const float* splats, const float* sh, uint32_t n, int deg,
                           const ViewParams& v, const float cam[3], ExtRec* rec, uint32_t* counts, uint4* rects,
                           uint64_t* dkey, uint32_t* wmin_part, uint32_t* clr0, uint32_t nclr0, uint32_t* clr1,
                           uint32_t nclr1, uint32_t* clr2, uint32_t nclr2) {
    if (n == 0) {
        if (nclr0) launch_fill_u32(st, clr0, 0u, nclr0);
        if (nclr1) launch_fill_u32(st, clr1, 0u, nclr1);
        if (nclr2) launch_fill_u32(st, clr2, 0u, nclr2);
        return;
    }
    const ClearPair clr{{nclr0 ? clr0 : nullptr, nclr1 ? clr1 : nullptr, nclr2 ? clr2 : nullptr},
                        {nclr0, nclr1, nclr2}};
    dim3 g((n + 255) / 256), b(256);
    const float4* s4 = (const float4*)splats;
    const float4* h4 = (const float4*)sh;
    switch (deg) {
        case 0: hipLaunchKernelGGL(k_ext_preprocess<0>, g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], rec, counts, rects, dkey, wmin_part, clr); break;
        case 1: hipLaunchKernelGGL(k_ext_preprocess<1>, g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], rec, counts, rects, dkey, wmin_part, clr); break;
        case 2: hipLaunchKernelGGL(k_ext_preprocess<2>, g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], rec, counts, rects, dkey, wmin_part, clr); break;
        default: hipLaunchKernelGGL(k_ext_preprocess<3>, g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], rec, counts, rects, dkey, wmin_part, clr); break;
    }
}

void launch_ext_keys(hipStream_t st, uint32_t n, const uint64_t* order, const uint32_t* rank_counts,
                     const uint4* rects, uint32_t tiles_x, const uint64_t* offsets, uint64_t* keys, uint64_t cap,
                     uint32_t* overflow) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_ext_keys, dim3((n + 255) / 256), dim3(256), 0, st, n, order, rank_counts, rects, tiles_x,
                       offsets, keys, cap, overflow);
}

// bits needed to write x (0 for x = 0)
static int bits_for_u32(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }

bool ext_bin_fits(uint32_t n, uint32_t tiles_x, uint32_t tiles_y, uint64_t cap) {
    const int xb = bits_for_u32(tiles_x > 0 ? tiles_x - 1 : 0), yb = bits_for_u32(tiles_y > 0 ? tiles_y - 1 : 0);
    const int idb = bits_for_u32(n > 1 ? n - 1 : 0);
    return xb <= kBinMaxBits && yb <= kBinMaxBits && xb + idb <= 32 && cap < 0xFFFFFFFFull;
}

size_t ext_bin_row_words(uint32_t n, uint32_t tiles_y) {
    const size_t nb1 = ((size_t)n + kBinThreads - 1) / kBinThreads;
    return (size_t)tiles_y * std::max<size_t>(nb1, 1) + 1;  // + the total
}

size_t ext_bin_col_words(uint32_t tiles_x, uint64_t cap) {
    const size_t nb2 = (size_t)((cap + kBinColBlock - 1) / kBinColBlock);
    return (size_t)tiles_x * std::max<size_t>(nb2, 1) + 1;
}

void ext_bin_rows(hipStream_t st, uint32_t n, const uint64_t* order, const uint4* rects, uint32_t tiles_y,
                  uint32_t* rcnt, uint32_t* roff, void* rscan) {
    if (n == 0 || tiles_y == 0) {  // no keys: the count word alone
        launch_fill_u32(st, roff + (ext_bin_row_words(n, tiles_y) - 1), 0u, 1);
        return;
    }
    const uint32_t nb1 = (n + kBinThreads - 1) / kBinThreads;
    hipLaunchKernelGGL(k_bin_row_count, dim3(nb1), dim3(kBinThreads), 0, st, n, order, rects, tiles_y, rcnt);
    exclusive_scan_u32(st, rcnt, roff, (uint64_t)tiles_y * nb1, rscan);
}

void ext_bin(hipStream_t st, uint32_t n, const uint64_t* order, const uint4* rects,
             uint32_t tiles_x, uint32_t tiles_y, uint64_t cap, const uint32_t* m_dev, const uint32_t* roff,
             uint32_t* items, uint32_t* ids, uint32_t* ranges, uint32_t* tile_order, uint32_t* ccnt, uint32_t* coff,
             uint32_t* tile_cnt, void* cscan, uint32_t* overflow, bool tile_cnt_clear) {
    const int xb = bits_for_u32(tiles_x > 0 ? tiles_x - 1 : 0), yb = bits_for_u32(tiles_y > 0 ? tiles_y - 1 : 0);
    const int idb = bits_for_u32(n > 1 ? n - 1 : 0);
    const uint32_t ntiles = tiles_x * tiles_y;
    if (!tile_cnt_clear) launch_fill_u32(st, tile_cnt, 0u, ntiles);
    if (n > 0 && cap > 0 && ntiles > 0) {
        const uint32_t nb1 = (n + kBinThreads - 1) / kBinThreads;
        hipLaunchKernelGGL(k_bin_row_emit, dim3(nb1), dim3(kBinThreads), 0, st, n, order, rects, tiles_y, yb, idb,
                           roff, items, cap, overflow);
        const uint32_t nb2 = (uint32_t)((cap + kBinColBlock - 1) / kBinColBlock);
        hipLaunchKernelGGL(k_bin_col_count, dim3(nb2), dim3(256), 0, st, (const uint32_t*)items, cap, m_dev, roff,
                           nb1, tiles_y, tiles_x, idb, ccnt, tile_cnt);
        exclusive_scan_u32(st, ccnt, coff, (uint64_t)tiles_x * nb2, cscan);
        hipLaunchKernelGGL(k_bin_col_scatter, dim3(nb2), dim3(256), 0, st, (const uint32_t*)items, cap, m_dev,
                           tiles_x, xb, idb, (const uint32_t*)coff, ids);
    }
    if (ntiles > 0)  // ranges (column-major list order) and the composite's workgroup order
        hipLaunchKernelGGL(k_tile_order<true>, dim3(1), dim3(1024), 0, st, (const uint32_t*)tile_cnt, tiles_x,
                           tiles_y, tile_order, ranges);
}

void launch_keys_to_ids(hipStream_t st, const uint64_t* keys, uint64_t m, const uint64_t* m_dev, uint32_t* ids) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_keys_to_ids, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, keys, m, m_dev, ids);
}

uint32_t ext_order_slots(uint32_t tiles_x, uint32_t tiles_y) {
    const uint32_t S = ((tiles_x + 1u) / 2u) * ((tiles_y + 1u) / 2u);
    return 4u * ((S + 7u) / 8u * 8u);
}

void launch_tile_order(hipStream_t st, const uint32_t* ranges, uint32_t tiles_x, uint32_t tiles_y, uint32_t* order) {
    if (tiles_x == 0 || tiles_y == 0) return;
    hipLaunchKernelGGL(k_tile_order<false>, dim3(1), dim3(1024), 0, st, ranges, tiles_x, tiles_y, order, nullptr);
}

template <int TILE, int BATCH>
static void launch_composite_t(hipStream_t st, const ViewParams& v, const ExtCompositeArgs& a) {
    hipLaunchKernelGGL((k_composite_ext<TILE, BATCH>), dim3(ext_order_slots(v.tiles_x, v.tiles_y)),
                       dim3(TILE * TILE), 0, st, v, a);
}

bool ext_tune_valid(int tile, int batch) {
    return (tile == 8 || tile == 16 || tile == 32) && (batch == 32 || batch == 64 || batch == 128);
}

void launch_composite_ext(hipStream_t st, const ViewParams& v, const ExtCompositeArgs& a, int batch) {
    if (v.num_tiles == 0) return;
    const int tile = (int)v.tile;
#define NPS_COMPOSITE_CASE(T, B) \
    if (tile == T && batch == B) return launch_composite_t<T, B>(st, v, a);
    NPS_COMPOSITE_CASE(8, 32) NPS_COMPOSITE_CASE(8, 64) NPS_COMPOSITE_CASE(8, 128)
    NPS_COMPOSITE_CASE(16, 32) NPS_COMPOSITE_CASE(16, 64) NPS_COMPOSITE_CASE(16, 128)
    NPS_COMPOSITE_CASE(32, 32) NPS_COMPOSITE_CASE(32, 64) NPS_COMPOSITE_CASE(32, 128)
#undef NPS_COMPOSITE_CASE
}

size_t ext_ckpt_floats(size_t slots, int tile, int batch) { return slots * (size_t)(batch / 8) * tile * tile; }
size_t ext_act_words(size_t slots, int tile, int batch) {
    return slots * (size_t)(tile * tile / 64) * (size_t)((batch + 63) / 64);
}

template <bool kStore>
static void launch_ext_chain_t(hipStream_t st, const float4* s4, const float4* h4, uint32_t n, int deg,
                               const ViewParams& v, const float cam[3], const double* g2, const uint32_t* counts,
                               float* grads, float2* a2, const uint32_t* fl, uint32_t nfl, double* g2c) {
    dim3 g((n + 255) / 256), b(256);
    switch (deg) {
        case 0: hipLaunchKernelGGL((k_ext_chain<0, kStore>), g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], g2, counts, grads, a2, fl, nfl, g2c); break;
        case 1: hipLaunchKernelGGL((k_ext_chain<1, kStore>), g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], g2, counts, grads, a2, fl, nfl, g2c); break;
        case 2: hipLaunchKernelGGL((k_ext_chain<2, kStore>), g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], g2, counts, grads, a2, fl, nfl, g2c); break;
        default: hipLaunchKernelGGL((k_ext_chain<3, kStore>), g, b, 0, st, s4, h4, n, v, cam[0], cam[1], cam[2], g2, counts, grads, a2, fl, nfl, g2c); break;
    }
}

void launch_ext_chain(hipStream_t st, const float* splats, const float* sh, uint32_t n, int deg, const ViewParams& v,
                      const float cam[3], const double* g2, const uint32_t* counts, float* grads, float* accum,
                      bool grads_zero, const uint32_t* flags, uint32_t nflags, double* g2_clear) {
    if (n == 0) return;
    const float4* s4 = (const float4*)splats;
    const float4* h4 = (const float4*)sh;
    float2* a2 = (float2*)accum;
    if (grads_zero) launch_ext_chain_t<true>(st, s4, h4, n, deg, v, cam, g2, counts, grads, a2, flags, nflags, g2_clear);
    else launch_ext_chain_t<false>(st, s4, h4, n, deg, v, cam, g2, counts, grads, a2, flags, nflags, g2_clear);
}

void launch_adam_ext(hipStream_t st, float* splats, float* sh, float* grads, float* m, float* v, uint32_t n, int deg,
                     float* ctl, const uint32_t* flags, uint32_t nflags) {
    hipLaunchKernelGGL(k_adam_ext_prep, dim3(1), dim3(64), 0, st, ctl, flags, nflags);
    if (n == 0) return;
    const size_t threads = 16 * (size_t)n;
    hipLaunchKernelGGL(k_adam_ext, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, (float4*)splats,
                       (float4*)sh, (float4*)grads, (float4*)m, (float4*)v, n, deg, (const float*)ctl);
}

}  // namespace nps
