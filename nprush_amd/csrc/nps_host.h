// nps_host.h -- host-side helpers shared by the C-ABI translation units:
// growable device buffers, the tile-list scratch of one view and its build
// (count scan -> key emission -> LSD radix sort on the tile bits -> ranges).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "../../include/nprush_hip.h"
#include "nps_kernels.h"

namespace nps {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Scratch {
    DevBuf counts, rects, offsets, keys, keys_alt, ranges, scan_tmp, sort_tmp;
    DevBuf diff, cut, full, kept;  // per-tile id-block cut of capped (REF) lists
    uint64_t m = 0;
    bool in_alt = false;
    bool have_full = false;
    uint32_t splat_block = 0;
};

// Grows b to at least `bytes` (25 % headroom); a live buffer is freed only
// after `st` has drained.
inline int ensure_dev(DevBuf& b, size_t bytes, hipStream_t st, std::string& err) {
    if (bytes <= b.bytes && b.p) return NPS_OK;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            err = std::string("sync: ") + hipGetErrorString(e);
            return NPS_ERR_HIP;
        }
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t alloc = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&b.p, alloc);
    if (e != hipSuccess) {
        b.p = nullptr;
        err = "hipMalloc(" + std::to_string(alloc) + "): " + hipGetErrorString(e);
        return NPS_ERR_HIP;
    }
    b.bytes = alloc;
    return NPS_OK;
}

inline void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

inline void free_scratch(Scratch& s) {
    free_buf(s.counts); free_buf(s.rects); free_buf(s.offsets); free_buf(s.keys);
    free_buf(s.keys_alt); free_buf(s.ranges); free_buf(s.scan_tmp); free_buf(s.sort_tmp);
    free_buf(s.diff); free_buf(s.cut); free_buf(s.full); free_buf(s.kept);
}

// Ids per block of the per-tile cut: about 32 blocks, at least 4096 splats each.
inline uint32_t splat_block_of(uint32_t n) {
    uint32_t b = (n + 31u) / 32u;
    return b < 4096u ? 4096u : b;
}

inline ViewParams make_view(const float* vp, const float* vp_inv, uint32_t W, uint32_t H, uint32_t tile = kTile) {
    ViewParams v;
    memset(&v, 0, sizeof(v));
    memcpy(v.vp, vp, sizeof(v.vp));
    if (vp_inv) memcpy(v.vp_inv, vp_inv, sizeof(v.vp_inv));
    v.W = W;
    v.H = H;
    v.aspect = (float)W / (float)H;
    v.tile = tile;
    v.tiles_x = (W + tile - 1) / tile;
    v.tiles_y = (H + tile - 1) / tile;
    v.num_tiles = v.tiles_x * v.tiles_y;
    return v;
}

inline int bits_for(uint32_t x) {
    int b = 1;
    while (b < 32 && (x >> b) != 0u) ++b;
    return b;
}

// keys for the counts already in sc.counts/sc.rects: scan, emit, sort, ranges.
// ev_sort / ev_ranges (may be null) are recorded before the sort and the ranges.
inline int build_tile_lists_on(hipStream_t st, std::string& err, Scratch& sc, const ViewParams& v, uint32_t n,
                               hipEvent_t ev_sort, hipEvent_t ev_ranges) {
    int rc;
    if ((rc = ensure_dev(sc.offsets, sizeof(uint64_t) * ((size_t)n + 1), st, err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(sc.scan_tmp, scan_temp_bytes(n) + 64, st, err)) != NPS_OK) return rc;
    // capped lists: key counts under the per-tile cut; uncapped: the rectangle sizes
    const uint32_t* key_counts = (const uint32_t*)(sc.have_full ? sc.kept.p : sc.counts.p);
    exclusive_scan_u32_to_u64(st, key_counts, (uint64_t*)sc.offsets.p, n, sc.scan_tmp.p);
    uint64_t m = 0;
    hipError_t e = hipMemcpyAsync(&m, (uint64_t*)sc.offsets.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        err = std::string("key count readback: ") + hipGetErrorString(e);
        return NPS_ERR_HIP;
    }
    if (m >= 0xFFFFFFFFull) {
        err = "tile key count exceeds 2^32";
        return NPS_ERR_CAPACITY;
    }
    sc.m = m;
    if ((rc = ensure_dev(sc.keys, sizeof(uint64_t) * std::max<uint64_t>(m, 1), st, err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(sc.keys_alt, sizeof(uint64_t) * std::max<uint64_t>(m, 1), st, err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(sc.sort_tmp, radix_temp_bytes(m), st, err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(sc.ranges, sizeof(uint32_t) * 2 * (size_t)v.num_tiles, st, err)) != NPS_OK) return rc;
    // splats that keep no key skip emission entirely
    launch_emit_keys(st, n, key_counts, (const uint64_t*)sc.offsets.p, (const uint4*)sc.rects.p, v.tiles_x,
                     v.num_tiles, sc.have_full ? (const uint32_t*)sc.cut.p : nullptr,
                     sc.splat_block ? sc.splat_block : 1u, (uint64_t*)sc.keys.p);
    if (ev_sort) (void)hipEventRecord(ev_sort, st);
    int tile_bits = bits_for(v.num_tiles > 0 ? v.num_tiles - 1 : 0);
    sc.in_alt = radix_sort_u64(st, (uint64_t*)sc.keys.p, (uint64_t*)sc.keys_alt.p, nullptr, nullptr, m, 32,
                               32 + tile_bits, sc.sort_tmp.p);
    if (ev_ranges) (void)hipEventRecord(ev_ranges, st);
    launch_fill_u32(st, (uint32_t*)sc.ranges.p, 0u, 2 * (uint64_t)v.num_tiles);
    const uint64_t* sorted = (const uint64_t*)(sc.in_alt ? sc.keys_alt.p : sc.keys.p);
    launch_tile_ranges(st, sorted, m, (uint32_t*)sc.ranges.p);
    e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("tile lists: ") + hipGetErrorString(e);
        return NPS_ERR_HIP;
    }
    return NPS_OK;
}

}  // namespace nps
