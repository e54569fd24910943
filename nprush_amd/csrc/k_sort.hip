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
