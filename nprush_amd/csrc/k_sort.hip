// k_sort.hip -- device prefix sums and an LSD radix sort for 64-bit keys
// (optional 32-bit values), written for 64-wide wavefronts.
//
// Scan: reduce-then-scan over blocks of 2048 items (256 threads x 8).  Up to
// 1024 blocks, each block of the second kernel sums the partials before it
// (two launches); above that one workgroup scans the partials in between.
//
// Radix pass (8-bit or narrower digit at `shift`):
//   k_radix_hist    : 256 KI keys per workgroup (KI = 16; striped: item r of
//                     thread t is position r*256+t), LDS histogram, digit-major output
//                     hist[d * nblk + blk] so one exclusive scan yields every
//                     (digit, block) output offset;
//   k_radix_scatter : each wave ranks its own contiguous 64 KI keys in KI rounds
//                     of 64; keys with equal digits are found with one
//                     ballot per digit bit (match mask), their order inside a
//                     round is popcount(mask & lanes_below), and a per-wave LDS
//                     counter carries the rank across rounds.  Wave offsets per
//                     digit are then prefix-summed across the 4 waves.  Equal
//                     digits keep their input order, so passes compose (LSD).
#include <algorithm>

#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortBlock = kSortThreads * kSortItems;  // 4096
constexpr int kRadix = 256;

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = (T)__shfl_up((unsigned long long)x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread over a 256-thread block.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds4, T* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T x = wave_inclusive_scan<T>(v);
    if (lane == 63) lds4[wid] = x;
    __syncthreads();
    T wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        T s = lds4[w];
        if (w < wid) wbase += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return wbase + x - v;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const TI* __restrict__ in, uint64_t n,
                                                              TO* __restrict__ partials) {
    __shared__ TO lds4[4];
    uint64_t base = (uint64_t)blockIdx.x * kScanBlock;
    TO s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t idx = base + (uint64_t)k * kScanThreads + threadIdx.x;
        if (idx < n) s += (TO)in[idx];
    }
    TO total;
    (void)block_exclusive_scan<TO>(s, lds4, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// One workgroup: in-place exclusive scan of `n` partials, any n.
template <typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(TO* __restrict__ data, uint64_t n) {
    __shared__ TO lds4[4];
    TO carry = 0;
    for (uint64_t base = 0; base < n; base += kScanBlock) {
        TO v[kScanItems];
        TO s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + k;
            v[k] = idx < n ? data[idx] : (TO)0;
            s += v[k];
        }
        TO total;
        TO ex = block_exclusive_scan<TO>(s, lds4, &total);
        TO run = carry + ex;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + k;
            if (idx < n) data[idx] = run;
            run += v[k];
        }
        carry += total;
    }
}

// out[i] = sum(in[0..i)) for i < n and out[n] = total.
template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const TI* __restrict__ in, uint64_t n,
                                                             const TO* __restrict__ partials, TO* __restrict__ out) {
    __shared__ TO lds4[4];
    uint64_t base = (uint64_t)blockIdx.x * kScanBlock;
    TO v[kScanItems];
    TO s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + k;
        v[k] = idx < n ? (TO)in[idx] : (TO)0;
        s += v[k];
    }
    TO total;
    TO ex = block_exclusive_scan<TO>(s, lds4, &total);
    TO run = partials[blockIdx.x] + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) out[n] = run;
}

// As k_scan_apply, for at most kScanDirectBlocks blocks: each block sums the
// partials of the blocks before it itself (at most kScanDirectBlocks - 1
// loads spread over its threads), so no separate scan of the partials runs.
constexpr uint64_t kScanDirectBlocks = 1024;

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply_direct(const TI* __restrict__ in, uint64_t n,
                                                                    const TO* __restrict__ partials,
                                                                    TO* __restrict__ out) {
    __shared__ TO lds4[4];
    TO pre = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kScanThreads) pre += partials[j];
    TO before;
    (void)block_exclusive_scan<TO>(pre, lds4, &before);
    uint64_t base = (uint64_t)blockIdx.x * kScanBlock;
    TO v[kScanItems];
    TO s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + k;
        v[k] = idx < n ? (TO)in[idx] : (TO)0;
        s += v[k];
    }
    TO total;
    TO ex = block_exclusive_scan<TO>(s, lds4, &total);
    TO run = before + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) out[n] = run;
}

size_t scan_temp_bytes(uint64_t n) { return sizeof(uint64_t) * ((n + kScanBlock - 1) / kScanBlock + 1); }

template <typename TI, typename TO>
static void exclusive_scan(hipStream_t st, const TI* in, TO* out, uint64_t n, void* temp) {
    uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb == 0) nb = 1;
    TO* partials = (TO*)temp;
    hipLaunchKernelGGL((k_scan_reduce<TI, TO>), dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n, partials);
    if (nb <= kScanDirectBlocks) {
        hipLaunchKernelGGL((k_scan_apply_direct<TI, TO>), dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n,
                           (const TO*)partials, out);
        return;
    }
    hipLaunchKernelGGL((k_scan_partials<TO>), dim3(1), dim3(kScanThreads), 0, st, partials, nb);
    hipLaunchKernelGGL((k_scan_apply<TI, TO>), dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, n,
                       (const TO*)partials, out);
}

void exclusive_scan_u32_to_u64(hipStream_t st, const uint32_t* in, uint64_t* out, uint64_t n, void* temp) {
    exclusive_scan<uint32_t, uint64_t>(st, in, out, n, temp);
}
void exclusive_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, void* temp) {
    exclusive_scan<uint32_t, uint32_t>(st, in, out, n, temp);
}

// m_dev (optional): the key count lives in device memory (capacity m); the
// grid covers the capacity and blocks past the count contribute nothing.
// RADIX: 256 (8-bit digits) or 2048 (11-bit digits).  KI: keys per thread;
// a workgroup takes 256 KI keys, each of its 4 waves ranks 64 KI of them.
template <int RADIX, int KI>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t* __restrict__ keys, uint64_t m,
                                                             const uint64_t* __restrict__ m_dev, int shift,
                                                             uint32_t mask, uint32_t nblk,
                                                             uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[RADIX];
    if (m_dev) m = min(*m_dev, m);
    for (int d = threadIdx.x; d < RADIX; d += kSortThreads) h[d] = 0u;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * (kSortThreads * KI);
#pragma unroll
    for (int r = 0; r < KI; ++r) {
        uint64_t pos = base + (uint64_t)r * kSortThreads + threadIdx.x;
        if (pos < m) {
            uint32_t d = (uint32_t)(keys[pos] >> shift) & mask;
            atomicAdd(&h[d], 1u);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += kSortThreads) hist[(uint64_t)d * nblk + blockIdx.x] = h[d];
}

template <bool kHasVals, int RADIX, int KI>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t* __restrict__ keys_in,
                                                                const uint32_t* __restrict__ vals_in, uint64_t m,
                                                                const uint64_t* __restrict__ m_dev,
                                                                int shift, uint32_t mask, int bits, uint32_t nblk,
                                                                const uint32_t* __restrict__ offsets,
                                                                uint64_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ vals_out,
                                                                const uint32_t* __restrict__ gsrc,
                                                                uint32_t* __restrict__ gdst) {
    __shared__ uint32_t wcnt[4][RADIX];
    if (m_dev) m = min(*m_dev, m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < 4 * RADIX; k += kSortThreads) (&wcnt[0][0])[k] = 0u;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t base = (uint64_t)blockIdx.x * (kSortThreads * KI) + (uint64_t)wid * (64 * KI);
    uint64_t key[KI];
    uint32_t val[KI];
    uint32_t rank[KI];
#pragma unroll
    for (int r = 0; r < KI; ++r) {
        uint64_t pos = base + (uint64_t)r * 64 + lane;
        bool valid = pos < m;
        key[r] = valid ? keys_in[pos] : 0ull;
        if (kHasVals) val[r] = valid ? vals_in[pos] : 0u;
    }
#pragma unroll
    for (int r = 0; r < KI; ++r) {
        uint64_t pos = base + (uint64_t)r * 64 + lane;
        bool valid = pos < m;
        uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        uint64_t peers = __ballot(valid);
        constexpr int kBits = RADIX == 2048 ? 11 : 8;
#pragma unroll
        for (int b = 0; b < kBits; ++b) {  // bits at and above `bits` are 0 in every lane: no-ops
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        uint32_t cnt = valid ? wcnt[wid][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0u) wcnt[wid][d] = cnt + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[r] = cnt + before;
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += kSortThreads) {
        uint32_t g = offsets[(uint64_t)d * nblk + blockIdx.x];
        uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d];
        wcnt[0][d] = g;
        wcnt[1][d] = g + c0;
        wcnt[2][d] = g + c0 + c1;
        wcnt[3][d] = g + c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < KI; ++r) {
        uint64_t pos = base + (uint64_t)r * 64 + lane;
        if (pos < m) {
            uint32_t d = (uint32_t)(key[r] >> shift) & mask;
            uint32_t dst = wcnt[wid][d] + rank[r];
            keys_out[dst] = key[r];
            if (kHasVals) vals_out[dst] = val[r];
            if (gsrc) gdst[dst] = gsrc[(uint32_t)key[r]];
        }
    }
}

constexpr int kRadixWide = 2048;  // 11-bit digits (radix_sort_u64 with wide = true)
// Keys per thread of the wide passes.  Smaller workgroups (4: 1024 keys, 4x
// the waves) measured slower at 1M keys: the 2048-digit histogram written
// and scanned per workgroup then outweighs the occupancy gained.
constexpr int kWideItems = 16;

size_t radix_temp_bytes(uint64_t m, bool wide) {
    const uint64_t block = wide ? kSortThreads * kWideItems : kSortBlock;
    uint64_t nblk = (m + block - 1) / block;
    if (nblk == 0) nblk = 1;
    uint64_t hist = (uint64_t)(wide ? kRadixWide : kRadix) * nblk;
    return sizeof(uint32_t) * 2 * (hist + 1) + scan_temp_bytes(hist) + 256;
}

// Sorts keys[0..m) (and vals when non-null) by bits [bit_lo, bit_hi).  Returns
// true when the result is in the *_alt buffers.  With m_dev, the count is
// min(*m_dev, m) read on the device (m is then the capacity): no host sync,
// and the launches are the same for every count (graph capture).  wide:
// 11-bit digits, else 8-bit digits.
bool radix_sort_u64(hipStream_t st, uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                    uint64_t m, int bit_lo, int bit_hi, void* temp, const uint64_t* m_dev, bool wide,
                    const uint32_t* gather_src, uint32_t* gather_dst) {
    if (m == 0 || bit_hi <= bit_lo) return false;
    const uint64_t block = wide ? kSortThreads * kWideItems : kSortBlock;
    uint32_t nblk = (uint32_t)((m + block - 1) / block);
    const int max_bits = wide ? 11 : 8;
    const int radix = wide ? kRadixWide : kRadix;
    uint64_t hist_n = (uint64_t)radix * nblk;
    uint32_t* hist = (uint32_t*)temp;
    uint32_t* offs = hist + hist_n + 1;
    void* scan_tmp = (void*)(((uintptr_t)(offs + hist_n + 1) + 255) & ~(uintptr_t)255);
    int total_bits = bit_hi - bit_lo;
    int passes = (total_bits + max_bits - 1) / max_bits;
    int per = (total_bits + passes - 1) / passes;
    bool alt = false;
    uint64_t* kin = keys;
    uint64_t* kout = keys_alt;
    uint32_t* vin = vals;
    uint32_t* vout = vals_alt;
    for (int p = 0; p < passes; ++p) {
        int shift = bit_lo + p * per;
        int bits = (shift + per <= bit_hi) ? per : (bit_hi - shift);
        uint32_t mask = (1u << bits) - 1u;
        const uint64_t hn = (uint64_t)radix * nblk;
        if (wide)
            hipLaunchKernelGGL((k_radix_hist<kRadixWide, kWideItems>), dim3(nblk), dim3(kSortThreads), 0, st,
                               (const uint64_t*)kin, m, m_dev, shift, mask, nblk, hist);
        else
            hipLaunchKernelGGL((k_radix_hist<kRadix, kSortItems>), dim3(nblk), dim3(kSortThreads), 0, st,
                               (const uint64_t*)kin, m, m_dev, shift, mask, nblk, hist);
        exclusive_scan_u32(st, hist, offs, hn, scan_tmp);
        const uint32_t* gs = p == passes - 1 ? gather_src : nullptr;
#define NPS_SCATTER(V, RX, KI)                                                                                      \
    hipLaunchKernelGGL((k_radix_scatter<V, RX, KI>), dim3(nblk), dim3(kSortThreads), 0, st, (const uint64_t*)kin,     \
                       (const uint32_t*)vin, m, m_dev, shift, mask, bits, nblk, (const uint32_t*)offs, kout, vout,    \
                       gs, gather_dst)
        if (vals) {
            if (wide) NPS_SCATTER(true, kRadixWide, kWideItems);
            else NPS_SCATTER(true, kRadix, kSortItems);
        } else {
            if (wide) NPS_SCATTER(false, kRadixWide, kWideItems);
            else NPS_SCATTER(false, kRadix, kSortItems);
        }
#undef NPS_SCATTER
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
        alt = !alt;
    }
    return alt;
}

// ---------------------------------------------------------------- depth order
// Keys depth_bits << 32 | id (depth_bits: the float bits of a positive depth,
// 0xFFFFFFFF for a splat that is not drawn) sorted on the depth bits in two
// 12-bit passes over a range-compressed code: code = min(hi - base, 2^24 - 1)
// with base the least drawn depth bits, and 2^24 - 1 for a splat not drawn.
// The code keeps the order of the depth bits as long as every drawn depth lies
// below base + 2^24 - 1 (a depth ratio of about 4 and more: two binades and
// more of float spacing); the first histogram pass checks that and otherwise
// stores 1 to flags[0] and flags[1] (the order is then a permutation, but not
// the depth order, and the caller sorts again on all 32 bits).
//
// Per pass: k_depth_hist (4096 keys per workgroup) writes its digit counts
// block-major, hist[blk][digit] (one contiguous 16 KB row per workgroup), and
// adds the counts of each group of 16 digits to gsum[group];
// k_depth_offsets (a workgroup per digit group, i.e. per contiguous range of
// the digit-major scan order) turns its 16 columns into the exclusive
// offsets of every (digit, block) in place; k_depth_scatter reads its block's
// row of offsets (contiguous) and scatters stably.

constexpr int kDepthRadix = 4096;
constexpr int kDepthItems = 16;
constexpr int kDepthBlock = kSortThreads * kDepthItems;
constexpr int kDepthGroups = kDepthRadix / 16;     // 256 groups of 16 digits
constexpr int kDepthTileRows = 256;                // rows (blocks) per LDS tile of k_depth_offsets
constexpr uint32_t kDepthCap = (1u << 24) - 1u;

__device__ __forceinline__ uint32_t depth_code(uint64_t key, uint32_t base) {
    const uint32_t hi = (uint32_t)(key >> 32);
    return hi == 0xFFFFFFFFu ? kDepthCap : min(hi - base, kDepthCap);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o, 64));
    return x;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
    return x;
}

// kFirst: reduces the per-workgroup minima part[0..npart) to base (block 0
// stores it to *base_slot) and checks the range; else reads *base_slot.
template <bool kFirst>
__global__ __launch_bounds__(kSortThreads) void k_depth_hist(const uint64_t* __restrict__ keys, uint32_t m, int shift,
                                                             const uint32_t* __restrict__ part, uint32_t npart,
                                                             uint32_t* __restrict__ base_slot,
                                                             uint32_t* __restrict__ flags,
                                                             uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ gsum) {
    __shared__ uint32_t h[kDepthRadix];
    __shared__ uint32_t s_min[4];
    __shared__ uint32_t s_over;
    const uint32_t tid = threadIdx.x;
    for (int d = tid; d < kDepthRadix; d += kSortThreads) h[d] = 0u;
    uint32_t base;
    if (kFirst) {
        uint32_t b = 0xFFFFFFFFu;
        for (uint32_t k = tid; k < npart; k += kSortThreads) b = min(b, part[k]);
        b = wave_min_u32(b);
        if ((tid & 63) == 0) s_min[tid >> 6] = b;
        if (tid == 0) s_over = 0u;
        __syncthreads();
        base = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
        if (blockIdx.x == 0 && tid == 0) *base_slot = base;
    } else {
        __syncthreads();
        base = *base_slot;
    }
    bool over = false;
    const uint32_t b0 = blockIdx.x * (uint32_t)kDepthBlock;
#pragma unroll
    for (int r = 0; r < kDepthItems; ++r) {
        const uint32_t pos = b0 + (uint32_t)r * kSortThreads + tid;
        if (pos < m) {
            const uint64_t key = keys[pos];
            if (kFirst) {
                const uint32_t hi = (uint32_t)(key >> 32);
                over |= hi != 0xFFFFFFFFu && hi - base >= kDepthCap;
            }
            atomicAdd(&h[(depth_code(key, base) >> shift) & (kDepthRadix - 1)], 1u);
        }
    }
    if (kFirst && over) s_over = 1u;
    __syncthreads();
    uint32_t* row = hist + (size_t)blockIdx.x * kDepthRadix;
    for (int d = tid; d < kDepthRadix; d += kSortThreads) row[d] = h[d];
    // digit group tid: 16 counts (read rotated: no LDS bank conflicts)
    uint32_t g = 0u;
#pragma unroll
    for (int j = 0; j < 16; ++j) g += h[16 * tid + ((j + tid) & 15)];
    if (g) atomicAdd(&gsum[tid], g);
    if (kFirst && tid == 0 && s_over) {
        flags[0] = 1u;
        flags[1] = 1u;
    }
}

// One workgroup per group of 16 digits: hist[b][d] becomes the exclusive
// offset of (digit d, block b) in digit-major order, in place.
// Thread t: digit column j = t & 15 of the group, row lane r = t >> 4.  Up to
// kDepthTileRows blocks the group's columns are read once into LDS; beyond,
// a first sweep sums the columns and a second one scans them tile by tile.
__global__ __launch_bounds__(kSortThreads) void k_depth_offsets(uint32_t* __restrict__ hist, uint32_t nblk,
                                                                const uint32_t* __restrict__ gsum) {
    __shared__ uint32_t tile[kDepthTileRows][17];
    __shared__ uint32_t s_w[4];
    __shared__ uint32_t s_col[16][17];
    __shared__ uint32_t s_tot[16];
    __shared__ uint32_t s_base[16];
    const uint32_t tid = threadIdx.x, j = tid & 15u, r = tid >> 4, g = blockIdx.x;
    uint32_t* col = hist + 16u * g + j;
    const bool one = nblk <= (uint32_t)kDepthTileRows;
    // keys with a digit below the group
    uint32_t gb = tid < g ? gsum[tid] : 0u;
    gb = wave_sum_u32(gb);
    if ((tid & 63) == 0) s_w[tid >> 6] = gb;
    // tile rows t0 .. t0 + rows: thread (j, r) scans rows 16 r .. 16 r + 15 of
    // column j serially (exclusive, in the tile), segment sum to s_col[r][j]
    auto load_scan = [&](uint32_t t0) {
        const uint32_t rows = min((uint32_t)kDepthTileRows, nblk - t0);
#pragma unroll 4
        for (uint32_t q = r; q < kDepthTileRows; q += 16)
            tile[q][j] = q < rows ? col[(size_t)(t0 + q) * kDepthRadix] : 0u;
        __syncthreads();
        uint32_t seg = 0u;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t v = tile[16 * r + q][j];
            tile[16 * r + q][j] = seg;
            seg += v;
        }
        s_col[r][j] = seg;
    };
    if (one) {
        load_scan(0u);
    } else {
        uint32_t tot = 0u;
        for (uint32_t row = r; row < nblk; row += 16) tot += col[(size_t)row * kDepthRadix];
        s_col[r][j] = tot;
    }
    __syncthreads();
    if (tid < 16) {
        uint32_t t = 0u;
        for (int q = 0; q < 16; ++q) t += s_col[q][tid];
        s_tot[tid] = t;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        for (int q = 0; q < 16; ++q) {
            s_base[q] = acc;
            acc += s_tot[q];
        }
    }
    __syncthreads();
    uint32_t carry = s_base[j];
    for (uint32_t t0 = 0; t0 < nblk; t0 += kDepthTileRows) {
        if (!one) {
            load_scan(t0);
            __syncthreads();
        }
        const uint32_t rows = min((uint32_t)kDepthTileRows, nblk - t0);
        uint32_t pre = carry, all = carry;
        for (uint32_t q = 0; q < 16; ++q) {
            const uint32_t v = s_col[q][j];
            pre += q < r ? v : 0u;
            all += v;
        }
#pragma unroll 4
        for (uint32_t q = 0; q < 16; ++q) {
            const uint32_t row = 16 * r + q;
            if (row < rows) col[(size_t)(t0 + row) * kDepthRadix] = pre + tile[row][j];
        }
        carry = all;
        __syncthreads();
    }
}

// Stable scatter on the 12-bit digit of the code at `shift` (ranking as in
// k_radix_scatter); per-wave digit counts are 16-bit (a wave ranks 1024 keys).
__global__ __launch_bounds__(kSortThreads) void k_depth_scatter(const uint64_t* __restrict__ keys_in, uint32_t m,
                                                                int shift, const uint32_t* __restrict__ base_slot,
                                                                const uint32_t* __restrict__ offsets,
                                                                uint64_t* __restrict__ keys_out,
                                                                const uint32_t* __restrict__ gsrc,
                                                                uint32_t* __restrict__ gdst) {
    __shared__ uint16_t wcnt[4][kDepthRadix];
    __shared__ uint32_t goff[kDepthRadix];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < 2 * kDepthRadix; k += kSortThreads) ((uint32_t*)&wcnt[0][0])[k] = 0u;
    const uint32_t base = *base_slot;
    const uint32_t* orow = offsets + (size_t)blockIdx.x * kDepthRadix;
    for (int d = threadIdx.x; d < kDepthRadix; d += kSortThreads) goff[d] = orow[d];
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t b0 = blockIdx.x * (uint32_t)kDepthBlock + (uint32_t)wid * (64 * kDepthItems);
    uint64_t key[kDepthItems];
    uint32_t rank[kDepthItems];
#pragma unroll
    for (int r = 0; r < kDepthItems; ++r) {
        const uint32_t pos = b0 + (uint32_t)r * 64 + lane;
        key[r] = pos < m ? keys_in[pos] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < kDepthItems; ++r) {
        const uint32_t pos = b0 + (uint32_t)r * 64 + lane;
        const bool valid = pos < m;
        const uint32_t d = (depth_code(key[r], base) >> shift) & (kDepthRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 12; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t cnt = valid ? (uint32_t)wcnt[wid][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0u) wcnt[wid][d] = (uint16_t)(cnt + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        rank[r] = cnt + before;
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kDepthRadix; d += kSortThreads) {
        const uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d];
        wcnt[0][d] = 0;
        wcnt[1][d] = (uint16_t)c0;
        wcnt[2][d] = (uint16_t)(c0 + c1);
        wcnt[3][d] = (uint16_t)(c0 + c1 + c2);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kDepthItems; ++r) {
        const uint32_t pos = b0 + (uint32_t)r * 64 + lane;
        if (pos < m) {
            const uint32_t d = (depth_code(key[r], base) >> shift) & (kDepthRadix - 1);
            const uint32_t dst = goff[d] + (uint32_t)wcnt[wid][d] + rank[r];
            keys_out[dst] = key[r];
            if (gsrc) gdst[dst] = gsrc[(uint32_t)key[r]];
        }
    }
}

// temp: base slot and two passes' group sums (1024 words), then the block-major
// histogram / offsets (4096 words per block of 4096 keys)
size_t depth_sort_temp_bytes(uint64_t m) {
    uint64_t nblk = (m + kDepthBlock - 1) / kDepthBlock;
    if (nblk == 0) nblk = 1;
    return sizeof(uint32_t) * (1024 + (uint64_t)kDepthRadix * nblk);
}

uint32_t* depth_sort_zero_region(void* temp) { return (uint32_t*)temp + 256; }

void depth_sort_narrow(hipStream_t st, uint64_t* keys, uint64_t* keys_alt, uint32_t m, void* temp,
                       const uint32_t* wmin_part, uint32_t npart, uint32_t* flags, const uint32_t* gather_src,
                       uint32_t* gather_dst, bool counters_clear) {
    static_assert(2 * kDepthGroups == depth_sort_zero_words, "zero region: the two passes' group sums");
    if (m == 0) return;
    const uint32_t nblk = (m + kDepthBlock - 1) / kDepthBlock;
    uint32_t* base_slot = (uint32_t*)temp;
    uint32_t* gsum = depth_sort_zero_region(temp);  // [2][kDepthGroups]
    uint32_t* hist = base_slot + 1024;
    if (!counters_clear) launch_fill_u32(st, gsum, 0u, 2 * kDepthGroups);
    hipLaunchKernelGGL(k_depth_hist<true>, dim3(nblk), dim3(kSortThreads), 0, st, (const uint64_t*)keys, m, 0,
                       wmin_part, npart, base_slot, flags, hist, gsum);
    hipLaunchKernelGGL(k_depth_offsets, dim3(kDepthGroups), dim3(kSortThreads), 0, st, hist, nblk,
                       (const uint32_t*)gsum);
    hipLaunchKernelGGL(k_depth_scatter, dim3(nblk), dim3(kSortThreads), 0, st, (const uint64_t*)keys, m, 0,
                       (const uint32_t*)base_slot, (const uint32_t*)hist, keys_alt, (const uint32_t*)nullptr,
                       (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_depth_hist<false>, dim3(nblk), dim3(kSortThreads), 0, st, (const uint64_t*)keys_alt, m, 12,
                       wmin_part, npart, base_slot, flags, hist, gsum + kDepthGroups);
    hipLaunchKernelGGL(k_depth_offsets, dim3(kDepthGroups), dim3(kSortThreads), 0, st, hist, nblk,
                       (const uint32_t*)(gsum + kDepthGroups));
    hipLaunchKernelGGL(k_depth_scatter, dim3(nblk), dim3(kSortThreads), 0, st, (const uint64_t*)keys_alt, m, 12,
                       (const uint32_t*)base_slot, (const uint32_t*)hist, keys, gather_src, gather_dst);
}

// Sums per-workgroup counter slots (PartMap): one workgroup, stats[dst[k]] +=
// sum over slots of part[K * slot + k].  Ordered after the producing kernel on
// the stream, so plain read-modify-writes suffice.
__global__ __launch_bounds__(1024) void k_sum_parts(const unsigned long long* __restrict__ part, uint32_t nslot,
                                                    PartMap m, unsigned long long* __restrict__ stats) {
    __shared__ unsigned long long s[8][16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (uint32_t k = 0; k < m.nk; ++k) {
        unsigned long long acc = 0ull;
        for (uint32_t i = tid; i < nslot; i += 1024u) acc += part[(size_t)m.K * i + k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) s[k][wid] = acc;
    }
    __syncthreads();
    if ((uint32_t)tid < m.nk) {
        unsigned long long t = 0ull;
#pragma unroll
        for (int w = 0; w < 16; ++w) t += s[tid][w];
        stats[m.dst[tid]] += t;
    }
}

void launch_sum_parts(hipStream_t st, const unsigned long long* part, uint32_t nslot, const PartMap& m,
                      unsigned long long* stats) {
    if (nslot == 0 || m.nk == 0) return;
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, st, part, nslot, m, stats);
}

}  // namespace nps
