// nps_api.cpp -- the C ABI of libnprush_hip.so (declared in include/nprush_hip.h).
//
// Host orchestration of one training view, in the order of
// GpuSplatOptimizerManager.dispatch (src/lib/gpu/GpuSplatOptimizerManager.ts:365-388):
//   backward = preprocess -> scan of per-splat tile counts -> key emission ->
//              LSD radix sort on the tile bits -> tile ranges -> fused composite
//   step     = fused Adam
//   adc      = every adc_period-th dispatch
// EXT mode: preprocess (SH colour, 3-sigma rectangles, depth keys) -> scan ->
// (tile | depth, id) pairs -> LSD radix sort with values -> tile ranges ->
// composite with early termination -> per-splat chain rule; step = EXT Adam.
// All work is enqueued on the handle's stream (default priority).  Per-view
// scratch grows on demand and is kept for the next call.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nprush_hip.h"
#include "nps_host.h"
#include "nps_kernels.h"

using namespace nps;

namespace {

// Buffers of one EXT view pipeline: the handle's own (nps_backward,
// nps_forward) and one set per lane of a multi-view step.
struct ExtView {
    DevBuf rec, counts, rects, dkey, dkey_alt, rank_counts, offsets, scan_tmp, keys, keys_alt, sort_tmp, ranges;
    DevBuf order, tile_work, ckpt, act, g2, c_img, tile_loss, stats, flags;  // flags: [0] key overflow
    DevBuf part;  // per-tile counter slots of the composite
    DevBuf bin_rcnt, bin_roff, bin_rscan;         // binning, row pass: counts, offsets, scan temp
    DevBuf wmin;                                  // per preprocess workgroup: least drawn depth bits
    DevBuf bin_cnt, bin_off, bin_tile, bin_scan;  // binning, column pass: counts, offsets, per-tile counts, temp
    uint32_t* ids = nullptr;  // splat ids in (tile, depth, id) order: keys or keys_alt read as 32-bit words
    bool col_major = false;   // ids stored tile column by tile column (ext_bin)
    uint64_t m = 0;    // key count of the last eager run
    uint64_t cap = 0;  // key capacity of a captured run
    // Shape of the last eager run that completed (splats, view, tile, batch):
    // the next run of the same shape sizes its key buffers from that run's
    // count instead of waiting for its own.
    uint64_t shape = 0;
    // Tiles of the last composite whose counter slots (part) are not summed into
    // stats yet: the statistics are summed when they are read (nps_read_stats).
    uint32_t part_tiles = 0;
    // Speculative views sort the depths in two narrow passes (depth_sort_narrow)
    // unless a recent one found its depth range too wide: then the next
    // narrow_skip speculative views use the full 32-bit sort.
    uint32_t narrow_skip = 0;
    // entries [0, g2_zero_n) of g2 are zero: the chain rule of the last view that
    // wrote it cleared them (eager views with their own g2 and the chain right
    // after the composite)
    uint32_t g2_zero_n = 0;
    bool in_alt = false;
    uint32_t num_tiles = 0;
    uint32_t* flags_at = nullptr;  // flags in a shared buffer instead of `flags` (16 words)
    hipEvent_t ev_count = nullptr;  // eager: recorded after the key-count copy
    uint64_t* count_host = nullptr; // eager: pinned landing slot of the key count
};

void free_ext_view(ExtView& x);

// Tile edge and records per LDS stage of the EXT composite.
struct ExtTune {
    int tile = kTile;
    int batch = kBatch;
};

// The splats an EXT view pipeline reads and the gradients it adds into.
struct SceneRef {
    const float* splats;
    const float* sh;
    uint32_t n;
    int deg;
    float* gradx;
    float* accum;
};

}  // namespace

struct nps_ctx {
    std::vector<struct npf_ctx*> renders;  // npf handles drawing this handle's splats
    std::vector<struct npt_ctx*> steps;    // multi-view steps over this scene
    nps_config cfg;
    int device = 0;
    bool ext = false;  // NPS_MODE_EXT
    int deg = 0;       // EXT SH degree
    ExtTune tune;      // EXT composite tile size and batch
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    float* splats = nullptr;
    int32_t* grads = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* accum = nullptr;
    float* sh = nullptr;     // EXT: 48 per splat
    float* gradx = nullptr;  // EXT: 64 FP32 gradients per splat
    bool gradx_zero = true;  // EXT: gradx is all zero (created cleared, or cleared by the last Adam step)
    float adam_t = 0.0f;
    int64_t step_count = 0;
    AdcExtParams adcx = {2e-4f, 1.0f, 0.005f, 0.01f};  // EXT densification
    uint32_t adc_calls = 0;                             // seed of the EXT split samples
    float vp[16] = {0}, vp_inv[16] = {0};
    int blur = 0;
    bool has_camera = false;
    uint32_t W = 0, H = 0;
    bool has_target = false;
    DevBuf tgt_rgba, tgt_depth;
    DevBuf fwd, bwd, ckpt, c_img, d_img, tile_loss, stats, part;
    ExtView xv;      // EXT: buffers of nps_backward / nps_forward
    DevBuf adam_dev;  // EXT: Adam step counter and bias corrections on the device (t, 1-b1^t, 1-b2^t, skip)
    Scratch sc;
    uint32_t last_tiles = 0;
    bool have_backward = false;
    // ADC buffers
    DevBuf adc_u32, adc_tmp, adc_scan;
    DevBuf adc_feat, adc_delta, snap;  // learned densification: features, deltas, state snapshot
    bool have_snap = false;
    float snap_adam_t = 0.0f;     // host step counters at the snapshot
    int64_t snap_step_count = 0;
    int64_t stats_host[12] = {0};
    bool timing = false;
    hipEvent_t ev[8] = {nullptr};
    float timing_ms[8] = {0};
    std::string err;
};

struct npf_ctx {
    nps_ctx* owner = nullptr;
    float vp[16] = {0};
    bool has_vp = false;
    uint32_t W = 0, H = 0;
    DevBuf color, depth, rec;
    Scratch sc;
};

struct npb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf src, dst, temp;
    std::string err;
};

namespace {

int set_err(nps_ctx* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define NPS_HIP(h, expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return set_err((h), NPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure(nps_ctx* h, DevBuf& b, size_t bytes) { return ensure_dev(b, bytes, h->stream, h->err); }

void free_ext_view(ExtView& x) {
    DevBuf* all[] = {&x.rec, &x.counts, &x.rects, &x.dkey, &x.dkey_alt, &x.rank_counts, &x.offsets, &x.scan_tmp,
                     &x.keys, &x.keys_alt, &x.sort_tmp, &x.ranges, &x.order, &x.tile_work, &x.ckpt, &x.act,
                     &x.g2, &x.c_img, &x.tile_loss, &x.stats, &x.flags, &x.part, &x.bin_cnt, &x.bin_off,
                     &x.bin_tile, &x.bin_scan, &x.bin_rcnt, &x.bin_roff, &x.bin_rscan, &x.wmin};
    for (DevBuf* b : all) free_buf(*b);
    if (x.ev_count) {
        (void)hipEventDestroy(x.ev_count);
        x.ev_count = nullptr;
    }
    if (x.count_host) {
        (void)hipHostFree(x.count_host);
        x.count_host = nullptr;
    }
}

int build_tile_lists(nps_ctx* h, Scratch& sc, const ViewParams& v, uint32_t n, int ev_sort, int ev_ranges) {
    hipEvent_t es = (h->timing && ev_sort >= 0) ? h->ev[ev_sort] : nullptr;
    hipEvent_t er = (h->timing && ev_ranges >= 0) ? h->ev[ev_ranges] : nullptr;
    return build_tile_lists_on(h->stream, h->err, sc, v, n, es, er);
}

// Where a view's per-splat tile counts and 2D gradients go, and whether its
// chain-rule kernel runs right away.  A captured multi-view step keeps them
// per view and runs every scene's chain kernels after the composites, in the
// step order of its views, on one stream per scene.
struct ViewOut {
    uint32_t* counts = nullptr;  // n (default: the view buffers')
    double* g2 = nullptr;        // 9n (default: the view buffers')
    bool defer_chain = false;
    bool grads_zero = false;     // the scene's FP32 gradients are known to be zero: the chain stores
};

// Key-buffer capacity of a captured view with `m` keys measured: 25 % + 64K headroom.
uint64_t ext_key_capacity(uint64_t m) { return m + m / 4 + 65536; }

// Allocates every buffer of view x for `n` splats, a W x H view and `cap` keys
// (capture mode needs them all in place before the capture starts).
int ext_view_reserve(ExtView& x, uint32_t n, uint32_t W, uint32_t H, uint64_t cap, bool grads, const ExtTune& tu,
                     hipStream_t st, std::string& err) {
    const size_t n1 = std::max<uint32_t>(n, 1);
    const uint32_t T = (uint32_t)tu.tile;
    const uint32_t tx = (W + T - 1) / T, ty = (H + T - 1) / T, nt = std::max<uint32_t>(tx * ty, 1);
    const size_t c1 = std::max<uint64_t>(cap, 1);
    const size_t slots = (size_t)(cap / (uint64_t)tu.batch) + nt + 1;
    struct {
        DevBuf* b;
        size_t bytes;
    } need[] = {
        {&x.rec, sizeof(ExtRec) * n1}, {&x.counts, 4 * n1}, {&x.rects, 16 * n1}, {&x.dkey, 8 * n1},
        {&x.dkey_alt, 8 * n1}, {&x.rank_counts, 4 * n1}, {&x.offsets, 8 * (n1 + 1)},
        {&x.scan_tmp, scan_temp_bytes(n1) + 64}, {&x.stats, 64}, {&x.flags, 64},
        {&x.sort_tmp, std::max({radix_temp_bytes(c1), radix_temp_bytes(n1, true), depth_sort_temp_bytes(n1)})},
        {&x.wmin, 4 * ((n1 + 255) / 256)},
        {&x.keys, 8 * c1}, {&x.keys_alt, 8 * c1},
        {&x.ranges, 8 * (size_t)nt}, {&x.order, 4 * (size_t)ext_order_slots(tx, ty)}, {&x.tile_work, 8 * (size_t)nt},
        {&x.c_img, 12 * (size_t)W * H}, {&x.tile_loss, 16 * (size_t)nt}, {&x.part, 32 * (size_t)nt},
        {&x.bin_rcnt, 4 * ext_bin_row_words(n, ty)}, {&x.bin_roff, 4 * ext_bin_row_words(n, ty)},
        {&x.bin_rscan, scan_temp_bytes(ext_bin_row_words(n, ty)) + 64},
        {&x.bin_cnt, 4 * ext_bin_col_words(tx, c1)}, {&x.bin_off, 4 * ext_bin_col_words(tx, c1)},
        {&x.bin_tile, 4 * (size_t)nt}, {&x.bin_scan, scan_temp_bytes(ext_bin_col_words(tx, c1)) + 64},
    };
    int rc;
    for (auto& q : need)
        if ((rc = ensure_dev(*q.b, q.bytes, st, err)) != NPS_OK) return rc;
    if (grads) {
        if ((rc = ensure_dev(x.ckpt, sizeof(float) * ext_ckpt_floats(slots, tu.tile, tu.batch), st, err)) != NPS_OK)
            return rc;
        if ((rc = ensure_dev(x.act, sizeof(uint64_t) * ext_act_words(slots, tu.tile, tu.batch), st, err)) != NPS_OK)
            return rc;
        void* g2_old = x.g2.p;
        if ((rc = ensure_dev(x.g2, sizeof(double) * 9 * n1, st, err)) != NPS_OK) return rc;
        if (x.g2.p != g2_old) x.g2_zero_n = 0;
    }
    return NPS_OK;
}

// One EXT view: preprocess -> depth order of the splats -> keys in that order
// -> stable sort on the tile bits -> ranges -> composite (-> chain rule).
// Eager: the key count is read back (one host sync) and buffers grow to fit.
// Capture: nothing is allocated or read back; the count stays on the device
// and keys beyond x.cap are dropped with flags[0] set.
// ev: optional 8 timing events (eager only).
int run_ext_view(const SceneRef& sc, const ViewParams& v, int batch, const float cam[3], const float4* tgt,
                 ExtView& x, hipStream_t st, bool grads, bool capture, std::string& err, const ViewOut* out,
                 hipEvent_t* ev, bool allow_spec = true) {
    const ExtTune tu{(int)v.tile, batch};
    const uint32_t n = sc.n;
    int rc;
    auto hip = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return NPS_OK;
        err = std::string(what) + ": " + hipGetErrorString(e);
        return NPS_ERR_HIP;
    };
    // Speculative eager run: the last run of the same shape gave a key count, so
    // the buffers are sized from it (ext_key_capacity: +25 %) and the count stays
    // on the device as in a captured run; the host checks the overflow flag once
    // the view is done (no wait in the middle of the view), and redoes the view
    // the waiting way if a key was dropped (the chain rule then changed nothing).
    const uint64_t shape = ((uint64_t)n * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)v.W << 40) ^ ((uint64_t)v.H << 20) ^
                           ((uint64_t)v.tile << 8) ^ (uint64_t)batch;
    const bool spec = !capture && allow_spec && x.m > 0 && x.shape == shape;
    const uint64_t spec_cap = spec ? ext_key_capacity(x.m) : 0;
    const bool dev_count = capture || spec;  // the key count stays on the device
    if (!capture && (rc = ext_view_reserve(x, n, v.W, v.H, spec_cap, grads, tu, st, err)) != NPS_OK) return rc;
    if (ev && (rc = hip(hipEventRecord(ev[0], st), "event")) != NPS_OK) return rc;
    x.part_tiles = 0;
    // flags: cleared per view when eager (by the preprocess); a captured step clears them once, so a
    // dropped key anywhere in a lane survives to the Adam update
    uint32_t* flags = x.flags_at ? x.flags_at : (uint32_t*)x.flags.p;
    uint32_t* counts = (out && out->counts) ? out->counts : (uint32_t*)x.counts.p;
    double* g2 = (out && out->g2) ? out->g2 : (double*)x.g2.p;
    // narrow depth sort: only where a too-wide depth range can be caught after
    // the view (speculative runs redo the view then)
    const bool narrow = spec && x.narrow_skip == 0 && n <= 0x7FFFFFFFu;
    if (spec && x.narrow_skip) --x.narrow_skip;
    // The binning's row pass needs no key count, and its last row offset is the
    // count (u32).  The key path (sorted tile << 32 | id keys) scans per-rank key
    // offsets instead and takes the count from them (u64).
    const bool rows_early = ext_bin_fits(n, v.tiles_x, v.tiles_y, 0);
    const bool key_path = !rows_early || (capture && x.cap >= 0xFFFFFFFFull) || (spec && spec_cap >= 0xFFFFFFFFull);
    // the preprocess also clears the flags (eager), the narrow sort's counters
    // and the binning's per-tile counts
    launch_ext_preprocess(st, sc.splats, sc.sh, n, sc.deg, v, cam, (ExtRec*)x.rec.p, counts, (uint4*)x.rects.p,
                          (uint64_t*)x.dkey.p, narrow ? (uint32_t*)x.wmin.p : nullptr, flags, capture ? 0u : 16u,
                          depth_sort_zero_region(x.sort_tmp.p), narrow ? depth_sort_zero_words : 0u,
                          (uint32_t*)x.bin_tile.p, rows_early ? v.num_tiles : 0u);
    // rank of every splat in (depth, id) order: stable sort of depth << 32 | id on
    // the depth bits; for the key path its last pass also gathers the tile
    // counts in rank order (the binning reads the rectangles instead)
    const uint32_t* gsrc = key_path ? (const uint32_t*)counts : nullptr;
    uint32_t* gdst = key_path ? (uint32_t*)x.rank_counts.p : nullptr;
    bool dalt = false;
    if (narrow)
        depth_sort_narrow(st, (uint64_t*)x.dkey.p, (uint64_t*)x.dkey_alt.p, n, x.sort_tmp.p, (const uint32_t*)x.wmin.p,
                          (n + 255) / 256, flags, gsrc, gdst, true);
    else
        dalt = radix_sort_u64(st, (uint64_t*)x.dkey.p, (uint64_t*)x.dkey_alt.p, nullptr, nullptr, n, 32, 64,
                              x.sort_tmp.p, nullptr, true, gsrc, gdst);
    const uint64_t* order = (const uint64_t*)(dalt ? x.dkey_alt.p : x.dkey.p);
    const uint32_t* count32 = nullptr;
    const uint64_t* count64 = nullptr;
    if (rows_early) {
        ext_bin_rows(st, n, order, (const uint4*)x.rects.p, v.tiles_y,
                     (uint32_t*)x.bin_rcnt.p, (uint32_t*)x.bin_roff.p, x.bin_rscan.p);
        count32 = (const uint32_t*)x.bin_roff.p + (ext_bin_row_words(n, v.tiles_y) - 1);
    }
    if (key_path) {
        exclusive_scan_u32_to_u64(st, (const uint32_t*)x.rank_counts.p, (uint64_t*)x.offsets.p, n, x.scan_tmp.p);
        count64 = (const uint64_t*)x.offsets.p + n;
    }
    uint64_t cap;
    const uint64_t* m_dev = nullptr;
    if (!dev_count) {
        // The host waits for the key count only; the fills queued behind the copy
        // (they do not depend on it) keep the GPU busy meanwhile.
        if (!x.ev_count && (rc = hip(hipEventCreateWithFlags(&x.ev_count, hipEventDisableTiming), "event")) != NPS_OK)
            return rc;
        if (!x.count_host && (rc = hip(hipHostMalloc((void**)&x.count_host, 64, hipHostMallocDefault), "pinned")) != NPS_OK)
            return rc;
        uint32_t* count_host32 = (uint32_t*)(x.count_host + 1);
        if ((rc = hip(count64 ? hipMemcpyAsync(x.count_host, count64, sizeof(uint64_t), hipMemcpyDeviceToHost, st)
                              : hipMemcpyAsync(count_host32, count32, sizeof(uint32_t), hipMemcpyDeviceToHost, st),
                      "count readback")) != NPS_OK)
            return rc;
        if ((rc = hip(hipEventRecord(x.ev_count, st), "event")) != NPS_OK) return rc;
        if (grads) launch_fill_u32(st, (uint32_t*)g2, 0u, 18 * (uint64_t)n);
        if ((rc = hip(hipEventSynchronize(x.ev_count), "sync")) != NPS_OK) return rc;
        const uint64_t m = count64 ? x.count_host[0] : (uint64_t)*count_host32;
        if (m >= 0xFFFFFFFFull) {
            err = "tile key count exceeds 2^32";
            return NPS_ERR_CAPACITY;
        }
        x.m = m;
        cap = m;
        if ((rc = ext_view_reserve(x, n, v.W, v.H, m, grads, tu, st, err)) != NPS_OK) return rc;
    } else {
        cap = capture ? x.cap : spec_cap;
        m_dev = count64;
    }
    x.num_tiles = v.num_tiles;
    x.col_major = rows_early && ext_bin_fits(n, v.tiles_x, v.tiles_y, cap);
    if (x.col_major) {
        // keys bucketed by tile row as they are enumerated, then one stable
        // pass on the tile column: ids and ranges
        x.ids = (uint32_t*)x.keys_alt.p;
        if (ev && (rc = hip(hipEventRecord(ev[1], st), "event")) != NPS_OK) return rc;
        ext_bin(st, n, order, (const uint4*)x.rects.p, v.tiles_x, v.tiles_y, cap,
                dev_count ? count32 : nullptr, (const uint32_t*)x.bin_roff.p, (uint32_t*)x.keys.p, x.ids, (uint32_t*)x.ranges.p,
                (uint32_t*)x.order.p, (uint32_t*)x.bin_cnt.p, (uint32_t*)x.bin_off.p, (uint32_t*)x.bin_tile.p,
                x.bin_scan.p, flags, true);
        if (ev && (rc = hip(hipEventRecord(ev[2], st), "event")) != NPS_OK) return rc;
    } else {
        // 8-byte keys tile << 32 | id in rank order, stable sort on the tile bits
        // (key_path holds here: col_major fails only for a capacity of 2^32 - 1 and more)
        if (!key_path) {
            err = "EXT view: key path without rank counts";
            return NPS_ERR_ARG;
        }
        launch_ext_keys(st, n, order, (const uint32_t*)x.rank_counts.p, (const uint4*)x.rects.p, v.tiles_x,
                        (const uint64_t*)x.offsets.p, (uint64_t*)x.keys.p, cap, flags);
        if (ev && (rc = hip(hipEventRecord(ev[1], st), "event")) != NPS_OK) return rc;
        const int tile_bits = bits_for(v.num_tiles > 0 ? v.num_tiles - 1 : 0);
        x.in_alt = radix_sort_u64(st, (uint64_t*)x.keys.p, (uint64_t*)x.keys_alt.p, nullptr, nullptr, cap, 32,
                                  32 + tile_bits, x.sort_tmp.p, m_dev);
        if (ev && (rc = hip(hipEventRecord(ev[2], st), "event")) != NPS_OK) return rc;
        const uint64_t* sorted = (const uint64_t*)(x.in_alt ? x.keys_alt.p : x.keys.p);
        x.ids = (uint32_t*)(x.in_alt ? x.keys.p : x.keys_alt.p);
        launch_fill_u32(st, (uint32_t*)x.ranges.p, 0u, 2 * (uint64_t)v.num_tiles);
        launch_tile_ranges(st, sorted, cap, (uint32_t*)x.ranges.p, m_dev);
        launch_keys_to_ids(st, sorted, cap, m_dev, x.ids);
    }
    // the chain rule clears what it reads when it runs right after the
    // composite on the view's own g2; then no fill is needed while that holds
    const bool own_g2 = !(out && out->g2);
    const bool clear_chain = grads && own_g2 && !capture && !(out && out->defer_chain);
    const uint32_t zero_n = x.g2_zero_n;
    if (dev_count && grads && !(clear_chain && zero_n >= n)) launch_fill_u32(st, (uint32_t*)g2, 0u, 18 * (uint64_t)n);
    if (grads && own_g2) x.g2_zero_n = 0;
    if (!x.col_major) launch_tile_order(st, (const uint32_t*)x.ranges.p, v.tiles_x, v.tiles_y, (uint32_t*)x.order.p);
    if (ev && (rc = hip(hipEventRecord(ev[3], st), "event")) != NPS_OK) return rc;
    ExtCompositeArgs a;
    a.rec = (const ExtRec*)x.rec.p;
    a.ids = x.ids;
    a.col_major = x.col_major ? 1 : 0;
    a.ranges = (const uint32_t*)x.ranges.p;
    a.tgt_rgba = tgt;
    a.g2 = grads ? g2 : nullptr;
    a.c_img = (float*)x.c_img.p;
    a.tile_loss = (double*)x.tile_loss.p;
    a.ckpt = grads ? (float*)x.ckpt.p : nullptr;
    a.act = grads ? (uint64_t*)x.act.p : nullptr;
    a.stats = (unsigned long long*)x.stats.p;
    a.order = (const uint32_t*)x.order.p;
    a.tile_work = (uint32_t*)x.tile_work.p;
    a.part = (unsigned long long*)x.part.p;
    a.backward = grads ? 1 : 0;
    launch_composite_ext(st, v, a, batch);
    x.part_tiles = v.num_tiles;
    if (ev && (rc = hip(hipEventRecord(ev[4], st), "event")) != NPS_OK) return rc;
    if (grads && !(out && out->defer_chain))
        launch_ext_chain(st, sc.splats, sc.sh, n, sc.deg, v, cam, (const double*)g2, (const uint32_t*)counts, sc.gradx,
                         sc.accum, out && out->grads_zero, spec ? flags : nullptr, spec ? 1u : 0u,
                         clear_chain ? g2 : nullptr);
    if (clear_chain) x.g2_zero_n = std::max(n, zero_n);
    if (ev && (rc = hip(hipEventRecord(ev[7], st), "event")) != NPS_OK) return rc;
    if ((rc = hip(hipGetLastError(), "launch")) != NPS_OK) return rc;
    if (capture) return NPS_OK;
    if (spec) {
        // the overflow flags (a dropped key or a too-wide depth range: flags[0];
        // the latter alone: flags[1]) and the key count, read once the view is done
        if (!x.count_host && (rc = hip(hipHostMalloc((void**)&x.count_host, 64, hipHostMallocDefault), "pinned")) != NPS_OK)
            return rc;
        uint32_t* h32 = (uint32_t*)(x.count_host + 1);
        if ((rc = hip(hipMemcpyAsync(h32 + 1, flags, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st),
                      "flag readback")) != NPS_OK)
            return rc;
        if ((rc = hip(count64 ? hipMemcpyAsync(x.count_host, count64, sizeof(uint64_t), hipMemcpyDeviceToHost, st)
                              : hipMemcpyAsync(h32, count32, sizeof(uint32_t), hipMemcpyDeviceToHost, st),
                      "count readback")) != NPS_OK)
            return rc;
        if ((rc = hip(hipStreamSynchronize(st), "sync")) != NPS_OK) return rc;
        const uint64_t m = count64 ? x.count_host[0] : (uint64_t)h32[0];
        if (h32[1] != 0u) {
            // a key was dropped or the depth order was not exact: size for the
            // count just seen and run the view again (32-bit depth sort)
            if (h32[2] != 0u) x.narrow_skip = 64;
            x.m = m;
            x.shape = 0;
            return run_ext_view(sc, v, batch, cam, tgt, x, st, grads, false, err, out, ev, false);
        }
        x.m = m;
    }
    x.shape = shape;
    return NPS_OK;
}

// Camera centre of a view: VP^-1 (0, 0, 1, 0) dehomogenised, in double, then rounded.
void ext_camera_centre(const float* vp_inv, float cam[3]) {
    for (int k = 0; k < 3; ++k) cam[k] = (float)((double)vp_inv[8 + k] / (double)vp_inv[11]);
}

SceneRef scene_of(nps_ctx* h) { return SceneRef{h->splats, h->sh, h->n, h->deg, h->gradx, h->accum}; }

const int kTuneTiles[3] = {8, 16, 32};
const int kTuneBatches[3] = {32, 64, 128};

// Times the 9 (tile, batch) variants of one view: sort + ranges + work order +
// composite fwd+bwd (HIP events, best of 3 after a warm-up; the chain rule is
// not run, so gradients and accumulators are untouched).  ms9[3 * ti + bi].
int ext_time_variants(const SceneRef& sc, const float* vp, const float* vp_inv, uint32_t W, uint32_t H,
                      const float4* tgt, ExtView& x, hipStream_t st, hipEvent_t* ev, std::string& err, float ms9[9],
                      ExtTune* best) {
    float cam[3];
    ext_camera_centre(vp_inv, cam);
    float best_ms = 1e30f;
    for (int ti = 0; ti < 3; ++ti)
        for (int bi = 0; bi < 3; ++bi) {
            ViewParams v = make_view(vp, vp_inv, W, H, (uint32_t)kTuneTiles[ti]);
            ViewOut out;
            out.defer_chain = true;
            float tmin = 1e30f;
            for (int rep = 0; rep < 4; ++rep) {
                int rc = run_ext_view(sc, v, kTuneBatches[bi], cam, tgt, x, st, true, false, err, &out, ev);
                if (rc != NPS_OK) return rc;
                if (hipEventSynchronize(ev[4]) != hipSuccess) {
                    err = "event sync";
                    return NPS_ERR_HIP;
                }
                float t = 0.0f;
                hipEventElapsedTime(&t, ev[1], ev[4]);
                if (rep > 0) tmin = std::min(tmin, t);
            }
            ms9[3 * ti + bi] = tmin;
            if (tmin < best_ms) {
                best_ms = tmin;
                *best = ExtTune{kTuneTiles[ti], kTuneBatches[bi]};
            }
        }
    return NPS_OK;
}

// nps_backward / nps_forward in EXT mode: the handle's camera and target.
int run_ext(nps_ctx* h, bool grads) {
    if (hipSetDevice(h->device) != hipSuccess) return set_err(h, NPS_ERR_HIP, "hipSetDevice");
    ViewParams v = make_view(h->vp, h->vp_inv, h->W, h->H, (uint32_t)h->tune.tile);
    float cam[3];
    ext_camera_centre(h->vp_inv, cam);
    ViewOut out;
    out.grads_zero = grads && h->gradx_zero;
    int rc = run_ext_view(scene_of(h), v, h->tune.batch, cam, (const float4*)h->tgt_rgba.p, h->xv, h->stream, grads,
                          false, h->err, &out, h->timing ? h->ev : nullptr);
    if (grads) h->gradx_zero = false;
    if (rc != NPS_OK) return rc;
    h->last_tiles = v.num_tiles;
    h->have_backward = true;
    h->stats_host[0] = (int64_t)h->xv.m;
    if (h->timing) {
        NPS_HIP(h, hipEventSynchronize(h->ev[7]));
        float t;
        hipEventElapsedTime(&t, h->ev[0], h->ev[1]); h->timing_ms[0] = t;
        hipEventElapsedTime(&t, h->ev[1], h->ev[2]); h->timing_ms[1] = t;
        hipEventElapsedTime(&t, h->ev[2], h->ev[3]); h->timing_ms[2] = t;
        hipEventElapsedTime(&t, h->ev[3], h->ev[4]); h->timing_ms[3] = t;
        hipEventElapsedTime(&t, h->ev[4], h->ev[7]); h->timing_ms[7] = t;
        hipEventElapsedTime(&t, h->ev[0], h->ev[7]); h->timing_ms[6] = t;
    }
    return NPS_OK;
}

}  // namespace

static void npf_release(npf_ctx* f);
static void npt_detach(npt_ctx* t);

extern "C" {

int nps_config_default(nps_config* cfg) {
    if (!cfg) return NPS_ERR_ARG;
    memset(cfg, 0, sizeof(*cfg));
    cfg->max_tile_splats = 1024;
    cfg->adc_period = 25;
    cfg->adc_enabled = 1;
    cfg->device = 0;
    cfg->ext_tile = 16;
    cfg->ext_batch = 64;
    return NPS_OK;
}

int nps_create(const nps_config* cfg, int32_t num_splats, nps_handle* out) {
    if (!out || num_splats < 0) return NPS_ERR_ARG;
    *out = nullptr;
    nps_ctx* h = new nps_ctx();
    if (cfg) h->cfg = *cfg;
    else nps_config_default(&h->cfg);
    if (h->cfg.max_tile_splats <= 0 || h->cfg.adc_period <= 0 ||
        (h->cfg.mode != NPS_MODE_REF && h->cfg.mode != NPS_MODE_EXT) || h->cfg.sh_degree < 0 ||
        h->cfg.sh_degree > 3) {
        delete h;
        return NPS_ERR_ARG;
    }
    h->ext = h->cfg.mode == NPS_MODE_EXT;
    h->deg = h->cfg.sh_degree;
    if (h->cfg.ext_tile || h->cfg.ext_batch) {
        if (!ext_tune_valid(h->cfg.ext_tile, h->cfg.ext_batch)) {
            delete h;
            return NPS_ERR_ARG;
        }
        h->tune = ExtTune{h->cfg.ext_tile, h->cfg.ext_batch};
    }
    h->device = h->cfg.device;
    h->n = (uint32_t)num_splats;
    auto fail = [&](hipError_t e) {
        (void)e;
        if (h->splats) (void)hipFree(h->splats);
        if (h->grads) (void)hipFree(h->grads);
        if (h->m) (void)hipFree(h->m);
        if (h->v) (void)hipFree(h->v);
        if (h->accum) (void)hipFree(h->accum);
        if (h->sh) (void)hipFree(h->sh);
        if (h->gradx) (void)hipFree(h->gradx);
        if (h->adam_dev.p) (void)hipFree(h->adam_dev.p);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return NPS_ERR_HIP;
    };
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return fail(e);
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    const size_t n1 = (size_t)std::max<uint32_t>(h->n, 1);
    const size_t n16 = 16 * n1;
    const size_t nopt = (h->ext ? 64 : 16) * n1;  // Adam state per splat
    const size_t nacc = (h->ext ? 2 : 1) * n1;
    if ((e = hipMalloc((void**)&h->splats, n16 * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&h->m, nopt * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&h->v, nopt * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void**)&h->accum, nacc * 4)) != hipSuccess) return fail(e);
    hipMemsetAsync(h->splats, 0, n16 * 4, h->stream);
    hipMemsetAsync(h->m, 0, nopt * 4, h->stream);
    hipMemsetAsync(h->v, 0, nopt * 4, h->stream);
    hipMemsetAsync(h->accum, 0, nacc * 4, h->stream);
    if (h->ext) {
        if ((e = hipMalloc(&h->adam_dev.p, 16)) != hipSuccess) return fail(e);
        h->adam_dev.bytes = 16;
        hipMemsetAsync(h->adam_dev.p, 0, 16, h->stream);
        if ((e = hipMalloc((void**)&h->sh, 48 * n1 * 4)) != hipSuccess) return fail(e);
        if ((e = hipMalloc((void**)&h->gradx, 64 * n1 * 4)) != hipSuccess) return fail(e);
        hipMemsetAsync(h->sh, 0, 48 * n1 * 4, h->stream);
        hipMemsetAsync(h->gradx, 0, 64 * n1 * 4, h->stream);
        h->gradx_zero = true;
    } else {
        if ((e = hipMalloc((void**)&h->grads, n16 * 4)) != hipSuccess) return fail(e);
        hipMemsetAsync(h->grads, 0, n16 * 4, h->stream);
    }
    for (int k = 0; k < 8; ++k)
        if ((e = hipEventCreate(&h->ev[k])) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail(e);
    *out = h;
    return NPS_OK;
}

void nps_destroy(nps_handle h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    // handles that use this one: their device memory goes now, and every later
    // call on them returns NPS_ERR_STATE (their destroy calls still free them)
    for (npf_ctx* f : h->renders) npf_release(f);
    h->renders.clear();
    for (npt_ctx* t : std::vector<npt_ctx*>(h->steps)) npt_detach(t);
    h->steps.clear();
    (void)hipFree(h->splats); (void)hipFree(h->m); (void)hipFree(h->v); (void)hipFree(h->accum);
    if (h->grads) (void)hipFree(h->grads);
    if (h->sh) (void)hipFree(h->sh);
    if (h->gradx) (void)hipFree(h->gradx);
    free_ext_view(h->xv);
    free_buf(h->adam_dev);
    free_buf(h->tgt_rgba); free_buf(h->tgt_depth); free_buf(h->fwd); free_buf(h->bwd); free_buf(h->ckpt);
    free_buf(h->c_img); free_buf(h->d_img); free_buf(h->tile_loss); free_buf(h->stats); free_buf(h->part);
    free_buf(h->adc_u32); free_buf(h->adc_tmp); free_buf(h->adc_scan);
    free_buf(h->adc_feat); free_buf(h->adc_delta); free_buf(h->snap);
    free_scratch(h->sc);
    for (int k = 0; k < 8; ++k)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* nps_last_error(nps_handle h) {
    if (!h) return "invalid handle";
    return h->err.c_str();
}

int32_t nps_num_splats(nps_handle h) { return h ? (int32_t)h->n : -1; }

static int check_n(nps_handle h, int32_t n) {
    if (!h) return NPS_ERR_ARG;
    if (n != (int32_t)h->n) return set_err(h, NPS_ERR_ARG, "n must equal num_splats");
    return NPS_OK;
}

#define NPS_COPY(h, dst, src, bytes, kind)                                                       \
    do {                                                                                        \
        NPS_HIP(h, hipSetDevice(h->device));                                                    \
        NPS_HIP(h, hipMemcpyAsync(dst, src, bytes, kind, h->stream));                           \
        NPS_HIP(h, hipStreamSynchronize(h->stream));                                            \
    } while (0)

int nps_write_splats(nps_handle h, const float* s, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (!s && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, h->splats, s, 64 * (size_t)n, hipMemcpyHostToDevice);
    return NPS_OK;
}
int nps_read_splats(nps_handle h, float* s, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (!s && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, s, h->splats, 64 * (size_t)n, hipMemcpyDeviceToHost);
    return NPS_OK;
}
int nps_write_grads(nps_handle h, const int32_t* g, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode: gradients are FP32 (nps_write_grads_f32)");
    if (!g && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, h->grads, g, 64 * (size_t)n, hipMemcpyHostToDevice);
    return NPS_OK;
}
int nps_read_grads(nps_handle h, int32_t* g, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode: gradients are FP32 (nps_read_grads_f32)");
    if (!g && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, g, h->grads, 64 * (size_t)n, hipMemcpyDeviceToHost);
    return NPS_OK;
}
int nps_write_adam(nps_handle h, const float* m, const float* v, float t, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if ((!m || !v) && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    const size_t per = h->ext ? 256 : 64;
    NPS_COPY(h, h->m, m, per * (size_t)n, hipMemcpyHostToDevice);
    NPS_COPY(h, h->v, v, per * (size_t)n, hipMemcpyHostToDevice);
    h->adam_t = t;
    if (h->ext) {
        const float ctl[4] = {t, 0.0f, 0.0f, 0.0f};
        NPS_COPY(h, h->adam_dev.p, ctl, sizeof(ctl), hipMemcpyHostToDevice);
    }
    return NPS_OK;
}
int nps_read_adam(nps_handle h, float* m, float* v, float* t, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    const size_t per = h->ext ? 256 : 64;
    if (m) NPS_COPY(h, m, h->m, per * (size_t)n, hipMemcpyDeviceToHost);
    if (v) NPS_COPY(h, v, h->v, per * (size_t)n, hipMemcpyDeviceToHost);
    if (t) *t = h->adam_t;
    return NPS_OK;
}
int nps_write_accum(nps_handle h, const float* a, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (!a && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, h->accum, a, (h->ext ? 8 : 4) * (size_t)n, hipMemcpyHostToDevice);
    return NPS_OK;
}
int nps_read_accum(nps_handle h, float* a, int32_t n) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (!a && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, a, h->accum, (h->ext ? 8 : 4) * (size_t)n, hipMemcpyDeviceToHost);
    return NPS_OK;
}

static int ext_copy(nps_handle h, int32_t n, const void* src, void* dst, size_t per, hipMemcpyKind kind) {
    int rc = check_n(h, n);
    if (rc) return rc;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    if ((!src || !dst) && n) return set_err(h, NPS_ERR_ARG, "null pointer");
    NPS_COPY(h, dst, src, per * (size_t)n, kind);
    return NPS_OK;
}
int nps_write_sh(nps_handle h, const float* sh, int32_t n) {
    return ext_copy(h, n, sh, h ? h->sh : nullptr, 192, hipMemcpyHostToDevice);
}
int nps_read_sh(nps_handle h, float* sh, int32_t n) {
    return ext_copy(h, n, h ? h->sh : nullptr, sh, 192, hipMemcpyDeviceToHost);
}
int nps_write_grads_f32(nps_handle h, const float* g, int32_t n) {
    if (h) h->gradx_zero = false;
    return ext_copy(h, n, g, h ? h->gradx : nullptr, 256, hipMemcpyHostToDevice);
}
int nps_read_grads_f32(nps_handle h, float* g, int32_t n) {
    return ext_copy(h, n, h ? h->gradx : nullptr, g, 256, hipMemcpyDeviceToHost);
}

int nps_set_camera(nps_handle h, const float vp[16], const float vp_inv[16], int32_t blur_enabled) {
    if (!h || !vp || !vp_inv) return h ? set_err(h, NPS_ERR_ARG, "null matrix") : NPS_ERR_ARG;
    memcpy(h->vp, vp, sizeof(h->vp));
    memcpy(h->vp_inv, vp_inv, sizeof(h->vp_inv));
    h->blur = blur_enabled;
    h->has_camera = true;
    return NPS_OK;
}

static int set_target_impl(nps_handle h, const float* rgba, const float* depth, int32_t w, int32_t hh,
                           hipMemcpyKind kind) {
    if (!h) return NPS_ERR_ARG;
    if (w <= 0 || hh <= 0 || !rgba || !depth) return set_err(h, NPS_ERR_ARG, "bad target");
    size_t px = (size_t)w * (size_t)hh;
    int rc;
    NPS_HIP(h, hipSetDevice(h->device));
    if ((rc = ensure(h, h->tgt_rgba, 16 * px)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->tgt_depth, 4 * px)) != NPS_OK) return rc;
    NPS_HIP(h, hipMemcpyAsync(h->tgt_rgba.p, rgba, 16 * px, kind, h->stream));
    NPS_HIP(h, hipMemcpyAsync(h->tgt_depth.p, depth, 4 * px, kind, h->stream));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    h->W = (uint32_t)w;
    h->H = (uint32_t)hh;
    h->has_target = true;
    return NPS_OK;
}

int nps_set_target(nps_handle h, const float* rgba, const float* depth, int32_t w, int32_t hh) {
    return set_target_impl(h, rgba, depth, w, hh, hipMemcpyHostToDevice);
}
int nps_set_target_dev(nps_handle h, const float* rgba, const float* depth, int32_t w, int32_t hh) {
    return set_target_impl(h, rgba, depth, w, hh, hipMemcpyDeviceToDevice);
}

int nps_backward(nps_handle h) {
    if (!h) return NPS_ERR_ARG;
    if (!h->has_target) return set_err(h, NPS_ERR_STATE, "no backward target (nps_set_target)");
    if (!h->has_camera) return set_err(h, NPS_ERR_STATE, "no camera (nps_set_camera)");
    if (h->ext) return run_ext(h, true);
    NPS_HIP(h, hipSetDevice(h->device));
    const uint32_t n = h->n;
    ViewParams v = make_view(h->vp, h->vp_inv, h->W, h->H);
    Scratch& sc = h->sc;
    int rc;
    if ((rc = ensure(h, h->fwd, sizeof(FwdRec) * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, h->bwd, sizeof(BwdRec) * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, sc.counts, 4 * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, sc.rects, 16 * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, h->stats, 64)) != NPS_OK) return rc;
    const uint32_t cap = (uint32_t)h->cfg.max_tile_splats;
    if (cap > (uint32_t)kMaxCap) return set_err(h, NPS_ERR_ARG, "max_tile_splats above 1024 is not supported");
    sc.splat_block = splat_block_of(n);
    const uint32_t nsb = (std::max<uint32_t>(n, 1) + sc.splat_block - 1) / sc.splat_block;
    const size_t diff_n = 4 * (size_t)nsb * (v.tiles_x + 1) * (v.tiles_y + 1) + 16 * (v.tiles_x + 1);
    if ((rc = ensure(h, sc.diff, 4 * diff_n)) != NPS_OK) return rc;
    if ((rc = ensure(h, sc.cut, 4 * (size_t)std::max<uint32_t>(v.num_tiles, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, sc.full, 4 * (size_t)std::max<uint32_t>(v.num_tiles, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, sc.kept, 4 * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[0], h->stream));
    launch_fill_u32(h->stream, (uint32_t*)h->stats.p, 0u, 16);
    launch_preprocess(h->stream, h->splats, n, v, (FwdRec*)h->fwd.p, (BwdRec*)h->bwd.p, (uint32_t*)sc.counts.p,
                      (uint4*)sc.rects.p);
    launch_tile_cut(h->stream, v, n, (int32_t*)sc.diff.p, sc.splat_block, cap, (uint32_t*)sc.cut.p,
                    (uint32_t*)sc.full.p, (unsigned long long*)h->stats.p + 1, (const uint32_t*)sc.counts.p,
                    (const uint4*)sc.rects.p, (uint32_t*)sc.kept.p);
    sc.have_full = true;
    if ((rc = build_tile_lists(h, sc, v, n, 1, 2)) != NPS_OK) return rc;
    const uint32_t nchunk = (cap + kChunk - 1) / kChunk;
    size_t px = (size_t)h->W * h->H;
    if ((rc = ensure(h, h->ckpt, sizeof(float) * (size_t)v.num_tiles * nchunk * kTilePx)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->c_img, 12 * px)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->d_img, 4 * px)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->tile_loss, 16 * (size_t)std::max<uint32_t>(v.num_tiles, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, h->part, 32 * (size_t)std::max<uint32_t>(v.num_tiles, 1))) != NPS_OK) return rc;
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[3], h->stream));
    CompositeArgs a;
    a.fwd = (const FwdRec*)h->fwd.p;
    a.bwd = (const BwdRec*)h->bwd.p;
    a.keys = (const uint64_t*)(sc.in_alt ? sc.keys_alt.p : sc.keys.p);
    a.ranges = (const uint32_t*)sc.ranges.p;
    a.tgt_rgba = (const float4*)h->tgt_rgba.p;
    a.tgt_depth = (const float*)h->tgt_depth.p;
    a.grads = h->grads;
    a.c_img = (float*)h->c_img.p;
    a.d_img = (float*)h->d_img.p;
    a.tile_loss = (double*)h->tile_loss.p;
    a.ckpt = (float*)h->ckpt.p;
    a.stats = (unsigned long long*)h->stats.p;
    a.cap = cap;
    a.nchunk_max = nchunk;
    a.part = (unsigned long long*)h->part.p;
    launch_composite_ref(h->stream, v, a);
    launch_sum_parts(h->stream, a.part, v.num_tiles, PartMap{4, 4, {0, 2, 3, 4, 0, 0, 0, 0}}, a.stats);
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[4], h->stream));
    NPS_HIP(h, hipGetLastError());
    h->last_tiles = v.num_tiles;
    h->have_backward = true;
    h->stats_host[0] = (int64_t)sc.m;
    if (h->timing) {
        NPS_HIP(h, hipEventSynchronize(h->ev[4]));
        float t;
        hipEventElapsedTime(&t, h->ev[0], h->ev[1]); h->timing_ms[0] = t;
        hipEventElapsedTime(&t, h->ev[1], h->ev[2]); h->timing_ms[1] = t;
        hipEventElapsedTime(&t, h->ev[2], h->ev[3]); h->timing_ms[2] = t;
        hipEventElapsedTime(&t, h->ev[3], h->ev[4]); h->timing_ms[3] = t;
        hipEventElapsedTime(&t, h->ev[0], h->ev[4]); h->timing_ms[6] = t;
    }
    return NPS_OK;
}

int nps_forward(nps_handle h) {
    if (!h) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "nps_forward is EXT mode only (REF renders with npf_*)");
    if (!h->has_target) return set_err(h, NPS_ERR_STATE, "no target size (nps_set_target)");
    if (!h->has_camera) return set_err(h, NPS_ERR_STATE, "no camera (nps_set_camera)");
    return run_ext(h, false);
}

int nps_autotune(nps_handle h, int32_t* tile, int32_t* batch, float* ms9) {
    if (!h) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    if (!h->has_target) return set_err(h, NPS_ERR_STATE, "no target (nps_set_target)");
    if (!h->has_camera) return set_err(h, NPS_ERR_STATE, "no camera (nps_set_camera)");
    NPS_HIP(h, hipSetDevice(h->device));
    float t9[9];
    ExtTune best = h->tune;
    int rc = ext_time_variants(scene_of(h), h->vp, h->vp_inv, h->W, h->H, (const float4*)h->tgt_rgba.p, h->xv,
                               h->stream, h->ev, h->err, t9, &best);
    if (rc != NPS_OK) return rc;
    h->tune = best;
    h->have_backward = false;  // the view buffers hold the last variant timed
    if (tile) *tile = best.tile;
    if (batch) *batch = best.batch;
    if (ms9) memcpy(ms9, t9, sizeof(t9));
    return NPS_OK;
}

int nps_step(nps_handle h) {
    if (!h) return NPS_ERR_ARG;
    NPS_HIP(h, hipSetDevice(h->device));
    float t = h->adam_t + 1.0f;
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[5], h->stream));
    if (h->ext) {
        launch_adam_ext(h->stream, h->splats, h->sh, h->gradx, h->m, h->v, h->n, h->deg, (float*)h->adam_dev.p,
                        nullptr, 0);
        h->gradx_zero = true;  // Adam clears every gradient it reads
    } else launch_adam_ref(h->stream, h->splats, h->grads, h->m, h->v, h->accum, h->n, t);
    if (h->timing) {
        NPS_HIP(h, hipEventRecord(h->ev[6], h->stream));
        NPS_HIP(h, hipEventSynchronize(h->ev[6]));
        hipEventElapsedTime(&h->timing_ms[4], h->ev[5], h->ev[6]);
    }
    NPS_HIP(h, hipGetLastError());
    h->adam_t = t;
    return NPS_OK;
}

// The EXT densification's lists in the handle's ADC buffers.  The rank arrays
// hold n + 1 entries: the exclusive scan also writes the total at [n].
static int adc_ext_buffers(nps_handle h, AdcExtBuffers& b) {
    const uint32_t n = h->n;
    const size_t nn = (size_t)std::max<uint32_t>(n, 1);
    int rc;
    if ((rc = ensure(h, h->adc_u32, 4 * (6 * nn + 10))) != NPS_OK) return rc;
    if ((rc = ensure(h, h->adc_scan, scan_temp_bytes(n) + 64)) != NPS_OK) return rc;
    uint32_t* base = (uint32_t*)h->adc_u32.p;
    b.dead = base;
    b.qual = base + nn;
    b.dead_rank = base + 2 * nn;      // n + 1
    b.qual_rank = base + 3 * nn + 1;  // n + 1
    b.dead_list = base + 4 * nn + 2;
    b.qual_list = base + 5 * nn + 2;
    b.cnt = base + 6 * nn + 2;
    b.scan_tmp = h->adc_scan.p;
    return NPS_OK;
}

static int adc_ext(nps_handle h) {
    NPS_HIP(h, hipSetDevice(h->device));
    const uint32_t n = h->n;
    int rc;
    AdcExtBuffers b;
    if ((rc = adc_ext_buffers(h, b)) != NPS_OK) return rc;
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[5], h->stream));
    run_adc_ext(h->stream, h->splats, h->sh, h->m, h->v, h->accum, n, h->adcx, h->adc_calls, b);
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[6], h->stream));
    NPS_HIP(h, hipGetLastError());
    uint32_t cnt[4] = {0, 0, 0, 0};
    NPS_HIP(h, hipMemcpyAsync(cnt, b.cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    if (h->timing) hipEventElapsedTime(&h->timing_ms[5], h->ev[5], h->ev[6]);
    h->stats_host[4] = cnt[0];
    h->stats_host[5] = std::min(cnt[0], cnt[1]);
    h->stats_host[6] = cnt[2];
    h->stats_host[7] = cnt[3];
    h->stats_host[8] = cnt[1];
    h->adc_calls++;
    return NPS_OK;
}

int nps_adc_features(nps_handle h, float* feats, int32_t max_rows, int32_t* nq) {
    if (!h || !nq || max_rows < 0) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    NPS_HIP(h, hipSetDevice(h->device));
    int rc;
    AdcExtBuffers b;
    if ((rc = adc_ext_buffers(h, b)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->adc_feat, 128 * (size_t)std::max<uint32_t>(h->n, 1))) != NPS_OK) return rc;
    float cam[3];
    ext_camera_centre(h->vp_inv, cam);
    launch_adc_ext_features(h->stream, h->splats, h->sh, h->accum, h->m, h->v, h->n, h->adcx, cam, b,
                            (float*)h->adc_feat.p);
    NPS_HIP(h, hipGetLastError());
    uint32_t cnt[2] = {0, 0};
    NPS_HIP(h, hipMemcpyAsync(cnt, b.cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    *nq = (int32_t)cnt[0];
    const size_t rows = std::min<size_t>(cnt[0], (size_t)max_rows);
    if (feats && rows) {
        NPS_HIP(h, hipMemcpyAsync(feats, h->adc_feat.p, 128 * rows, hipMemcpyDeviceToHost, h->stream));
        NPS_HIP(h, hipStreamSynchronize(h->stream));
    }
    return NPS_OK;
}

int nps_adc_features_dev(nps_handle h, const float** feats_dev, int32_t* nq) {
    if (!h || !feats_dev || !nq) return NPS_ERR_ARG;
    int rc = nps_adc_features(h, nullptr, 0, nq);
    if (rc != NPS_OK) return rc;
    *feats_dev = (const float*)h->adc_feat.p;
    return NPS_OK;
}

int nps_adc_apply_deltas_dev(nps_handle h, const float* deltas_dev, int32_t rows) {
    if (!h || rows < 0 || (rows > 0 && !deltas_dev)) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    if (!h->adc_u32.p) return set_err(h, NPS_ERR_STATE, "no densification yet");
    if (rows == 0) return NPS_OK;
    NPS_HIP(h, hipSetDevice(h->device));
    AdcExtBuffers b;
    int rc;
    if ((rc = adc_ext_buffers(h, b)) != NPS_OK) return rc;
    launch_adc_ext_deltas(h->stream, h->splats, b, deltas_dev, (uint32_t)rows);
    NPS_HIP(h, hipGetLastError());
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nps_adc_children(nps_handle h, int32_t* parents, int32_t* children, int32_t max_rows, int32_t* count) {
    if (!h || !count || max_rows < 0) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    if (!h->adc_u32.p) return set_err(h, NPS_ERR_STATE, "no densification yet");
    NPS_HIP(h, hipSetDevice(h->device));
    AdcExtBuffers b;
    int rc;
    if ((rc = adc_ext_buffers(h, b)) != NPS_OK) return rc;
    uint32_t cnt[2] = {0, 0};
    NPS_HIP(h, hipMemcpyAsync(cnt, b.cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    const uint32_t P = std::min(cnt[0], cnt[1]);
    *count = (int32_t)P;
    const size_t rows = std::min<size_t>(P, (size_t)max_rows);
    if (rows && parents) NPS_HIP(h, hipMemcpyAsync(parents, b.qual_list, 4 * rows, hipMemcpyDeviceToHost, h->stream));
    if (rows && children) NPS_HIP(h, hipMemcpyAsync(children, b.dead_list, 4 * rows, hipMemcpyDeviceToHost, h->stream));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nps_adc_apply_deltas(nps_handle h, const float* deltas, int32_t rows) {
    if (!h || rows < 0 || (rows > 0 && !deltas)) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    if (!h->adc_u32.p) return set_err(h, NPS_ERR_STATE, "no densification yet");
    if (rows == 0) return NPS_OK;
    NPS_HIP(h, hipSetDevice(h->device));
    AdcExtBuffers b;
    int rc;
    if ((rc = adc_ext_buffers(h, b)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->adc_delta, 56 * (size_t)rows)) != NPS_OK) return rc;
    NPS_HIP(h, hipMemcpyAsync(h->adc_delta.p, deltas, 56 * (size_t)rows, hipMemcpyHostToDevice, h->stream));
    launch_adc_ext_deltas(h->stream, h->splats, b, (const float*)h->adc_delta.p, (uint32_t)rows);
    NPS_HIP(h, hipGetLastError());
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nps_snapshot(nps_handle h, int32_t restore) {
    if (!h) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    NPS_HIP(h, hipSetDevice(h->device));
    const size_t n = h->n;
    // splats, SH, Adam m and v, accumulators, Adam's (t, 1 - b1^t, 1 - b2^t, skip)
    const size_t sz[6] = {64 * n, 192 * n, 256 * n, 256 * n, 8 * n, 16};
    void* live[6] = {h->splats, h->sh, h->m, h->v, h->accum, h->adam_dev.p};
    size_t tot = 0;
    for (size_t s : sz) tot += s;
    if (restore) {
        if (h->snap.bytes < tot || !h->have_snap) return set_err(h, NPS_ERR_STATE, "no snapshot");
    } else {
        int rc;
        if ((rc = ensure(h, h->snap, tot)) != NPS_OK) return rc;
    }
    size_t off = 0;
    for (int k = 0; k < 6; ++k) {
        char* s = (char*)h->snap.p + off;
        if (restore) NPS_HIP(h, hipMemcpyAsync(live[k], s, sz[k], hipMemcpyDeviceToDevice, h->stream));
        else NPS_HIP(h, hipMemcpyAsync(s, live[k], sz[k], hipMemcpyDeviceToDevice, h->stream));
        off += sz[k];
    }
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    if (restore) {
        h->adam_t = h->snap_adam_t;
        h->step_count = h->snap_step_count;
    } else {
        h->have_snap = true;
        h->snap_adam_t = h->adam_t;
        h->snap_step_count = h->step_count;
    }
    return NPS_OK;
}

int nps_set_adc_ext(nps_handle h, float grad_threshold, float scene_extent, float prune_opacity,
                    float percent_dense) {
    if (!h) return NPS_ERR_ARG;
    if (!h->ext) return set_err(h, NPS_ERR_STATE, "EXT mode only");
    if (!(grad_threshold >= 0.0f) || !(scene_extent > 0.0f) || !(prune_opacity >= 0.0f) || !(percent_dense > 0.0f))
        return set_err(h, NPS_ERR_ARG, "bad densification parameters");
    h->adcx = AdcExtParams{grad_threshold, scene_extent, prune_opacity, percent_dense};
    return NPS_OK;
}

int nps_adc(nps_handle h) {
    if (!h) return NPS_ERR_ARG;
    if (h->ext) return adc_ext(h);
    NPS_HIP(h, hipSetDevice(h->device));
    const uint32_t n = h->n;
    int rc;
    size_t nn = (size_t)std::max<uint32_t>(n, 1);
    // alive0 qa dead dead_rank(n+1) dead_list Q K(n+1) qual_list flag(4)
    if ((rc = ensure(h, h->adc_u32, 4 * (8 * nn + 2 + 4 + 64))) != NPS_OK) return rc;
    if ((rc = ensure(h, h->adc_tmp, 128 * nn)) != NPS_OK) return rc;
    if ((rc = ensure(h, h->adc_scan, scan_temp_bytes(n) + 64)) != NPS_OK) return rc;
    uint32_t* base = (uint32_t*)h->adc_u32.p;
    AdcBuffers b;
    b.alive0 = base;
    b.qa = base + nn;
    b.dead = base + 2 * nn;
    b.dead_rank = base + 3 * nn;        // n + 1
    b.dead_list = base + 4 * nn + 1;
    b.Q = base + 5 * nn + 1;
    b.K = base + 6 * nn + 1;            // n + 1
    b.qual_list = base + 7 * nn + 2;
    b.flag = base + 8 * nn + 2;
    b.tmp = (float*)h->adc_tmp.p;
    b.scan_tmp = h->adc_scan.p;
    int64_t st[5];
    if (h->timing) NPS_HIP(h, hipEventRecord(h->ev[5], h->stream));
    int r = run_adc_ref(h->stream, h->splats, h->m, h->v, h->accum, n, h->adam_t, b, st);
    if (h->timing) {
        NPS_HIP(h, hipEventRecord(h->ev[6], h->stream));
        NPS_HIP(h, hipEventSynchronize(h->ev[6]));
        hipEventElapsedTime(&h->timing_ms[5], h->ev[5], h->ev[6]);
    }
    NPS_HIP(h, hipGetLastError());
    for (int k = 0; k < 5; ++k) h->stats_host[4 + k] = st[k];
    if (r != 0) return set_err(h, NPS_ERR_CAPACITY, "ADC cascade chain deeper than 256");
    return NPS_OK;
}

int nps_dispatch(nps_handle h) {
    if (!h) return NPS_ERR_ARG;
    int rc = nps_backward(h);
    if (rc) return rc;
    if ((rc = nps_step(h)) != NPS_OK) return rc;
    h->step_count++;
    if (h->cfg.adc_enabled && h->step_count % h->cfg.adc_period == 0) return nps_adc(h);
    return NPS_OK;
}

int nps_sync(nps_handle h) {
    if (!h) return NPS_ERR_ARG;
    NPS_HIP(h, hipSetDevice(h->device));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nps_read_image(nps_handle h, float* c_pred, float* d_pred) {
    if (!h) return NPS_ERR_ARG;
    if (!h->have_backward) return set_err(h, NPS_ERR_STATE, "no backward has run");
    size_t px = (size_t)h->W * h->H;
    if (c_pred) NPS_COPY(h, c_pred, h->ext ? h->xv.c_img.p : h->c_img.p, 12 * px, hipMemcpyDeviceToHost);
    if (d_pred && h->ext) memset(d_pred, 0, 4 * px);
    else if (d_pred) NPS_COPY(h, d_pred, h->d_img.p, 4 * px, hipMemcpyDeviceToHost);
    return NPS_OK;
}

int nps_read_loss(nps_handle h, double loss[2]) {
    if (!h || !loss) return NPS_ERR_ARG;
    if (!h->have_backward) return set_err(h, NPS_ERR_STATE, "no backward has run");
    std::vector<double> t(2 * (size_t)h->last_tiles);
    NPS_COPY(h, t.data(), h->ext ? h->xv.tile_loss.p : h->tile_loss.p, sizeof(double) * t.size(),
             hipMemcpyDeviceToHost);
    double a = 0.0, b = 0.0;
    for (size_t k = 0; k < h->last_tiles; ++k) {
        a += t[2 * k];
        b += t[2 * k + 1];
    }
    loss[0] = a;
    loss[1] = b;
    return NPS_OK;
}

int nps_read_tiles(nps_handle h, uint32_t* ranges, uint32_t* ids, uint32_t* full_counts, int32_t* num_tiles,
                   int64_t* num_ids) {
    if (!h) return NPS_ERR_ARG;
    if (!h->have_backward) return set_err(h, NPS_ERR_STATE, "no backward has run");
    if (num_tiles) *num_tiles = (int32_t)h->last_tiles;
    if (num_ids) *num_ids = (int64_t)(h->ext ? h->xv.m : h->sc.m);
    if (h->ext) {
        std::vector<uint32_t> r(2 * (size_t)h->last_tiles);
        NPS_COPY(h, r.data(), h->xv.ranges.p, 8 * (size_t)h->last_tiles, hipMemcpyDeviceToHost);
        if (ranges) memcpy(ranges, r.data(), 8 * (size_t)h->last_tiles);
        if (full_counts)
            for (size_t t = 0; t < h->last_tiles; ++t) full_counts[t] = r[2 * t + 1] - r[2 * t];
        if (ids && h->xv.m) NPS_COPY(h, ids, h->xv.ids, 4 * (size_t)h->xv.m, hipMemcpyDeviceToHost);
        return NPS_OK;
    } else {
        if (ranges) NPS_COPY(h, ranges, h->sc.ranges.p, 8 * (size_t)h->last_tiles, hipMemcpyDeviceToHost);
        if (full_counts)
            NPS_COPY(h, full_counts, h->sc.full.p, 4 * (size_t)h->last_tiles, hipMemcpyDeviceToHost);
    }
    if (ids && h->sc.m) {
        std::vector<uint64_t> k(h->sc.m);
        NPS_COPY(h, k.data(), h->sc.in_alt ? h->sc.keys_alt.p : h->sc.keys.p, 8 * (size_t)h->sc.m,
                 hipMemcpyDeviceToHost);
        for (size_t j = 0; j < k.size(); ++j) ids[j] = (uint32_t)k[j];
    }
    return NPS_OK;
}

int nps_read_tile_work(nps_handle h, uint32_t* work) {
    if (!h || !work) return NPS_ERR_ARG;
    if (!h->ext || !h->have_backward) return set_err(h, NPS_ERR_STATE, "no EXT backward/forward has run");
    NPS_COPY(h, work, h->xv.tile_work.p, 8 * (size_t)h->last_tiles, hipMemcpyDeviceToHost);
    return NPS_OK;
}

int nps_read_stats(nps_handle h, int64_t stats[12]) {
    if (!h || !stats) return NPS_ERR_ARG;
    const DevBuf& sb = h->ext ? h->xv.stats : h->stats;
    if (h->ext && h->have_backward && sb.p) {
        // the last EXT composite's per-tile counter slots, summed now
        NPS_HIP(h, hipSetDevice(h->device));
        launch_fill_u32(h->stream, (uint32_t*)sb.p, 0u, 16);
        if (h->xv.part_tiles)
            launch_sum_parts(h->stream, (const unsigned long long*)h->xv.part.p, h->xv.part_tiles,
                             PartMap{4, 4, {0, 4, 2, 3, 0, 0, 0, 0}}, (unsigned long long*)sb.p);
        NPS_HIP(h, hipGetLastError());
    }
    if (h->have_backward && sb.p) {
        unsigned long long s[5] = {0, 0, 0, 0, 0};
        NPS_COPY(h, s, sb.p, sizeof(s), hipMemcpyDeviceToHost);
        h->stats_host[2] = (int64_t)s[0];
        h->stats_host[1] = (int64_t)s[1];
        h->stats_host[3] = (int64_t)s[2];
        h->stats_host[9] = (int64_t)s[3];
        h->stats_host[10] = (int64_t)s[4];
    }
    for (int k = 0; k < 12; ++k) stats[k] = h->stats_host[k];
    return NPS_OK;
}

int nps_enable_timing(nps_handle h, int32_t on) {
    if (!h) return NPS_ERR_ARG;
    h->timing = on != 0;
    return NPS_OK;
}

int nps_read_timing(nps_handle h, float ms[8]) {
    if (!h || !ms) return NPS_ERR_ARG;
    for (int k = 0; k < 8; ++k) ms[k] = h->timing_ms[k];
    return NPS_OK;
}

// ---------------- forward render manager ----------------

int npf_create(nps_handle owner, npf_handle* out) {
    if (!owner || !out) return NPS_ERR_ARG;
    if (owner->ext) return set_err(owner, NPS_ERR_STATE, "npf_* renders REF splats; EXT renders with nps_forward");
    npf_ctx* f = new npf_ctx();
    f->owner = owner;
    owner->renders.push_back(f);
    *out = f;
    return NPS_OK;
}

}  // extern "C"

static void npf_release(npf_ctx* f) {
    if (!f->owner) return;
    (void)hipSetDevice(f->owner->device);
    (void)hipStreamSynchronize(f->owner->stream);
    free_buf(f->color); free_buf(f->depth); free_buf(f->rec);
    free_scratch(f->sc);
    f->owner = nullptr;
}

extern "C" {

void npf_destroy(npf_handle f) {
    if (!f) return;
    if (f->owner) {
        auto& r = f->owner->renders;
        r.erase(std::remove(r.begin(), r.end(), f), r.end());
        npf_release(f);
    }
    delete f;
}

int npf_write_vp(npf_handle f, const float vp[16]) {
    if (!f || !vp) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;  // its splats' handle is destroyed
    memcpy(f->vp, vp, sizeof(f->vp));
    f->has_vp = true;
    return NPS_OK;
}

int npf_set_target(npf_handle f, int32_t w, int32_t hh) {
    if (!f) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;  // its splats' handle is destroyed
    nps_ctx* h = f->owner;
    if (w <= 0 || hh <= 0) return set_err(h, NPS_ERR_ARG, "bad render target size");
    NPS_HIP(h, hipSetDevice(h->device));
    size_t px = (size_t)w * hh;
    int rc;
    if ((rc = ensure(h, f->color, 16 * px)) != NPS_OK) return rc;
    if ((rc = ensure(h, f->depth, 16 * px)) != NPS_OK) return rc;
    f->W = (uint32_t)w;
    f->H = (uint32_t)hh;
    return NPS_OK;
}

int npf_dispatch(npf_handle f, int32_t clear, int32_t quantize) {
    if (!f) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;  // its splats' handle is destroyed
    nps_ctx* h = f->owner;
    if (!f->has_vp || f->W == 0) return set_err(h, NPS_ERR_STATE, "render target or VP not set");
    NPS_HIP(h, hipSetDevice(h->device));
    const uint32_t n = h->n;
    ViewParams v = make_view(f->vp, nullptr, f->W, f->H);
    int rc;
    if ((rc = ensure(h, f->rec, 64 * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, f->sc.counts, 4 * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    if ((rc = ensure(h, f->sc.rects, 16 * (size_t)std::max<uint32_t>(n, 1))) != NPS_OK) return rc;
    f->sc.have_full = false;
    f->sc.splat_block = 0;
    launch_preprocess_render(h->stream, h->splats, n, v, (float4*)f->rec.p, (uint32_t*)f->sc.counts.p,
                             (uint4*)f->sc.rects.p);
    if ((rc = build_tile_lists(h, f->sc, v, n, -1, -1)) != NPS_OK) return rc;
    RenderArgs a;
    a.rec = (const float4*)f->rec.p;
    a.keys = (const uint64_t*)(f->sc.in_alt ? f->sc.keys_alt.p : f->sc.keys.p);
    a.ranges = (const uint32_t*)f->sc.ranges.p;
    a.color = (float4*)f->color.p;
    a.depth = (float4*)f->depth.p;
    a.clear = clear != 0;
    a.quantize = quantize != 0;
    launch_render(h->stream, v, a);
    NPS_HIP(h, hipGetLastError());
    return NPS_OK;
}

int npf_read_target(npf_handle f, float* color, float* depth) {
    if (!f) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;  // its splats' handle is destroyed
    nps_ctx* h = f->owner;
    size_t px = (size_t)f->W * f->H;
    if (color) NPS_COPY(h, color, f->color.p, 16 * px, hipMemcpyDeviceToHost);
    if (depth) NPS_COPY(h, depth, f->depth.p, 16 * px, hipMemcpyDeviceToHost);
    return NPS_OK;
}

int npf_write_target(npf_handle f, const float* color, const float* depth) {
    if (!f) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;  // its splats' handle is destroyed
    nps_ctx* h = f->owner;
    size_t px = (size_t)f->W * f->H;
    if (color) NPS_COPY(h, f->color.p, color, 16 * px, hipMemcpyHostToDevice);
    if (depth) NPS_COPY(h, f->depth.p, depth, 16 * px, hipMemcpyHostToDevice);
    return NPS_OK;
}

// ---------------- multi-view step ----------------

}  // extern "C"

struct npt_view {
    int scene;
    float vp[16], vp_inv[16];
    const float* tgt;
    uint32_t W, H;
    ExtTune tune;
};

struct npt_lane {
    hipStream_t st = nullptr;
    ExtView x;
};

struct npt_ctx {
    std::vector<nps_handle> scenes;
    std::vector<npt_view> views;
    std::vector<int> order;    // step order of the views (largest first)
    std::vector<int> lane_of;  // lane of each view
    std::vector<npt_lane> lanes;
    int device = 0;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> lane_done, chain_done;  // chain_done: one per scene
    hipEvent_t joined = nullptr;
    DevBuf view_loss, flags;  // 2 doubles per view; 16 words per lane
    DevBuf view_g2, view_counts;    // per view: 9n doubles, n words (graph mode)
    std::vector<size_t> view_base;  // splat offset of each view in view_g2 / view_counts
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool planned = false;
    bool detached = false;  // a scene was destroyed: every call but npt_destroy is refused
    int lanes_used = 0;        // lanes the views are dealt to (<= lanes.size(); set by npt_autotune)
    int64_t graph_nodes = 0;   // nodes of the last captured graph
    float tune_ms[4] = {-1.0f, -1.0f, -1.0f, -1.0f};  // last npt_autotune's candidates (timing graphs)
    double cap_scale = 1.0;
    std::string err;
};

namespace {

int npt_fail(npt_handle t, int code, const std::string& msg) {
    t->err = msg;
    return code;
}

#define NPT_HIP(t, expr)                                                                                 \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return npt_fail((t), NPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

ViewParams npt_view_params(const npt_view& w) { return make_view(w.vp, w.vp_inv, w.W, w.H, (uint32_t)w.tune.tile); }

// Node limit of a captured step graph: the tests capture and replay a step of
// about 98.5 % of it (about 200 small views), the batched configuration of 4
// scenes and 64 views (chain rule and Adam included) is below it; a larger
// plan is refused rather than captured untested.
constexpr size_t kMaxGraphNodes = 8192;

// Views dealt to the first lanes_used lanes, most pixels first, each to the
// least-loaded lane (longest-processing-time order); t->order is the step order.
void npt_assign_lanes(npt_handle t) {
    const int V = (int)t->views.size();
    const int L = std::max(1, std::min(t->lanes_used > 0 ? t->lanes_used : (int)t->lanes.size(), (int)t->lanes.size()));
    t->order.resize(V);
    for (int i = 0; i < V; ++i) t->order[i] = i;
    std::stable_sort(t->order.begin(), t->order.end(), [&](int a, int b) {
        return (uint64_t)t->views[a].W * t->views[a].H > (uint64_t)t->views[b].W * t->views[b].H;
    });
    t->lane_of.assign(V, 0);
    std::vector<uint64_t> load(L, 0);
    for (int vi : t->order) {
        int best = 0;
        for (int l = 1; l < L; ++l)
            if (load[l] < load[best]) best = l;
        t->lane_of[vi] = best;
        load[best] += (uint64_t)t->views[vi].W * t->views[vi].H;
    }
}

void npt_drop_graph(npt_handle t) {
    if (t->exec) (void)hipGraphExecDestroy(t->exec);
    if (t->graph) (void)hipGraphDestroy(t->graph);
    t->exec = nullptr;
    t->graph = nullptr;
    t->planned = false;
}

// Adam (and densification on its cadence) for every scene after the views.
int npt_finish_scenes(npt_handle t, bool adam_done) {
    for (nps_handle h : t->scenes) {
        if (!adam_done) {
            launch_adam_ext(h->stream, h->splats, h->sh, h->gradx, h->m, h->v, h->n, h->deg, (float*)h->adam_dev.p,
                            nullptr, 0);
            NPT_HIP(t, hipGetLastError());
        }
        h->adam_t += 1.0f;
        h->step_count++;
        if (h->cfg.adc_enabled && h->step_count % h->cfg.adc_period == 0) {
            NPT_HIP(t, hipStreamSynchronize(h->stream));
            int rc = nps_adc(h);
            if (rc != NPS_OK) return npt_fail(t, rc, h->err);
        }
        NPT_HIP(t, hipStreamSynchronize(h->stream));
    }
    return NPS_OK;
}

// Views one after another, each with its own launches and key-count readback.
// serial: each view completes before the next one is launched.
int npt_step_sequential(npt_handle t, bool serial) {
    for (nps_handle h : t->scenes) NPT_HIP(t, hipStreamSynchronize(h->stream));
    for (int vi : t->order) {
        const npt_view& w = t->views[vi];
        nps_handle h = t->scenes[w.scene];
        ViewParams v = npt_view_params(w);
        float cam[3];
        ext_camera_centre(w.vp_inv, cam);
        int rc = run_ext_view(scene_of(h), v, w.tune.batch, cam, (const float4*)w.tgt, h->xv, h->stream, true, false,
                              t->err, nullptr, nullptr);
        if (rc != NPS_OK) return rc;
        launch_sum_tile_loss(h->stream, (const double*)h->xv.tile_loss.p, v.num_tiles,
                             (double*)t->view_loss.p + 2 * vi);
        if (serial) NPT_HIP(t, hipStreamSynchronize(h->stream));
    }
    return npt_finish_scenes(t, false);
}

// Sizes every lane for its views (key counts measured by a forward-only eager
// run) and captures the whole step as one graph.

int npt_plan(npt_handle t, bool timing) {
    npt_drop_graph(t);
    const int L = (int)t->lanes.size();
    std::vector<uint64_t> mmax(L, 0);
    for (int vi : t->order) {
        const npt_view& w = t->views[vi];
        nps_handle h = t->scenes[w.scene];
        npt_lane& ln = t->lanes[t->lane_of[vi]];
        ViewParams v = npt_view_params(w);
        float cam[3];
        ext_camera_centre(w.vp_inv, cam);
        int rc = run_ext_view(scene_of(h), v, w.tune.batch, cam, (const float4*)w.tgt, ln.x, ln.st, false, false,
                              t->err, nullptr, nullptr);
        if (rc != NPS_OK) return rc;
        const int l = t->lane_of[vi];
        mmax[l] = std::max(mmax[l], ln.x.m);
    }
    for (int l = 0; l < L; ++l) {
        npt_lane& ln = t->lanes[l];
        ln.x.cap = (uint64_t)((double)ext_key_capacity(mmax[l]) * t->cap_scale);
        ln.x.flags_at = (uint32_t*)t->flags.p + 16 * l;
    }
    // every buffer of a lane large enough for each of its views (tile and batch differ by view)
    for (int vi : t->order) {
        const npt_view& w = t->views[vi];
        npt_lane& ln = t->lanes[t->lane_of[vi]];
        int rc = ext_view_reserve(ln.x, t->scenes[w.scene]->n, w.W, w.H, ln.x.cap, true, w.tune, ln.st, t->err);
        if (rc != NPS_OK) return rc;
    }
    for (auto& ln : t->lanes) NPT_HIP(t, hipStreamSynchronize(ln.st));
    t->view_base.assign(t->views.size(), 0);
    size_t total = 0;
    for (size_t vi = 0; vi < t->views.size(); ++vi) {
        t->view_base[vi] = total;
        total += std::max<uint32_t>(t->scenes[t->views[vi].scene]->n, 1);
    }
    if (ensure_dev(t->view_g2, sizeof(double) * 9 * total, t->main, t->err) != NPS_OK) return NPS_ERR_HIP;
    if (ensure_dev(t->view_counts, sizeof(uint32_t) * total, t->main, t->err) != NPS_OK) return NPS_ERR_HIP;
    for (nps_handle h : t->scenes) NPT_HIP(t, hipStreamSynchronize(h->stream));

    NPT_HIP(t, hipStreamBeginCapture(t->main, hipStreamCaptureModeThreadLocal));
    int rc = NPS_OK;
    [&]() {
        launch_fill_u32(t->main, (uint32_t*)t->flags.p, 0u, 16 * (uint64_t)t->lanes.size());
        if (hipEventRecord(t->fork, t->main) != hipSuccess) { rc = npt_fail(t, NPS_ERR_HIP, "fork"); return; }
        for (auto& ln : t->lanes)
            if (hipStreamWaitEvent(ln.st, t->fork, 0) != hipSuccess) { rc = npt_fail(t, NPS_ERR_HIP, "join"); return; }
        for (int vi : t->order) {
            const npt_view& w = t->views[vi];
            nps_handle h = t->scenes[w.scene];
            npt_lane& ln = t->lanes[t->lane_of[vi]];
            ViewParams v = npt_view_params(w);
            float cam[3];
            ext_camera_centre(w.vp_inv, cam);
            ViewOut out;
            out.counts = (uint32_t*)t->view_counts.p + t->view_base[vi];
            out.g2 = (double*)t->view_g2.p + 9 * t->view_base[vi];
            out.defer_chain = true;
            rc = run_ext_view(scene_of(h), v, w.tune.batch, cam, (const float4*)w.tgt, ln.x, ln.st, true, true, t->err,
                              &out, nullptr);
            if (rc != NPS_OK) return;
            launch_sum_tile_loss(ln.st, (const double*)ln.x.tile_loss.p, v.num_tiles,
                                 (double*)t->view_loss.p + 2 * vi);
        }
        for (size_t l = 0; l < t->lanes.size(); ++l) {
            if (hipEventRecord(t->lane_done[l], t->lanes[l].st) != hipSuccess ||
                hipStreamWaitEvent(t->main, t->lane_done[l], 0) != hipSuccess) {
                rc = npt_fail(t, NPS_ERR_HIP, "join");
                return;
            }
        }
        if (timing) return;  // timing graph of the autotuner: the views only
        // chain rule: each scene's views in step order on one stream, scenes side by side
        if (hipEventRecord(t->joined, t->main) != hipSuccess) { rc = npt_fail(t, NPS_ERR_HIP, "fork"); return; }
        const size_t S = t->scenes.size(), L = t->lanes.size();
        for (size_t l = 0; l < std::min(S, L); ++l)
            if (hipStreamWaitEvent(t->lanes[l].st, t->joined, 0) != hipSuccess) { rc = npt_fail(t, NPS_ERR_HIP, "fork"); return; }
        for (size_t sidx = 0; sidx < S; ++sidx) {
            nps_handle h = t->scenes[sidx];
            hipStream_t st = t->lanes[sidx % L].st;
            for (int vi : t->order) {
                const npt_view& w = t->views[vi];
                if (w.scene != (int)sidx) continue;
                ViewParams v = npt_view_params(w);
                float cam[3];
                ext_camera_centre(w.vp_inv, cam);
                // gated on the lanes' overflow flags: after an overflow the host
                // redoes the step view by view, so nothing may be added here
                launch_ext_chain(st, h->splats, h->sh, h->n, h->deg, v, cam,
                                 (const double*)t->view_g2.p + 9 * t->view_base[vi],
                                 (const uint32_t*)t->view_counts.p + t->view_base[vi], h->gradx, h->accum, false,
                                 (const uint32_t*)t->flags.p, (uint32_t)t->lanes.size());
            }
        }
        for (size_t l = 0; l < std::min(S, L); ++l) {
            if (hipEventRecord(t->chain_done[l], t->lanes[l].st) != hipSuccess ||
                hipStreamWaitEvent(t->main, t->chain_done[l], 0) != hipSuccess) {
                rc = npt_fail(t, NPS_ERR_HIP, "join");
                return;
            }
        }
        for (nps_handle h : t->scenes)
            launch_adam_ext(t->main, h->splats, h->sh, h->gradx, h->m, h->v, h->n, h->deg, (float*)h->adam_dev.p,
                            (const uint32_t*)t->flags.p, (uint32_t)t->lanes.size());
    }();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(t->main, &g);
    if (rc != NPS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return npt_fail(t, NPS_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    size_t nodes = 0;
    if (hipGraphGetNodes(g, nullptr, &nodes) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return npt_fail(t, NPS_ERR_HIP, "hipGraphGetNodes");
    }
    t->graph_nodes = (int64_t)nodes;
    if (nodes > kMaxGraphNodes) {
        (void)hipGraphDestroy(g);
        return npt_fail(t, NPS_ERR_CAPACITY,
                        "captured step has " + std::to_string(nodes) + " nodes, above the " +
                            std::to_string(kMaxGraphNodes) + " of the largest graph tested; use npt_step mode 0");
    }
    t->graph = g;
    NPT_HIP(t, hipGraphInstantiate(&t->exec, t->graph, nullptr, nullptr, 0));
    t->planned = !timing;
    return NPS_OK;
}

int npt_step_graph(npt_handle t) {
    if (!t->planned) {
        int rc = npt_plan(t, false);
        if (rc != NPS_OK) return rc;
    }
    for (nps_handle h : t->scenes) NPT_HIP(t, hipStreamSynchronize(h->stream));
    NPT_HIP(t, hipGraphLaunch(t->exec, t->main));
    std::vector<uint32_t> fl(16 * t->lanes.size());
    NPT_HIP(t, hipMemcpyAsync(fl.data(), t->flags.p, 4 * fl.size(), hipMemcpyDeviceToHost, t->main));
    NPT_HIP(t, hipStreamSynchronize(t->main));
    bool overflow = false;
    for (size_t l = 0; l < t->lanes.size(); ++l) overflow = overflow || fl[16 * l] != 0;
    if (overflow) {
        // a view had more keys than its lane holds: the chain-rule kernels and
        // Adam saw the flag and changed nothing (gradients, densification
        // statistics, parameters); clear the gradients anyway, grow the lanes,
        // redo the step view by view and capture again next time
        for (nps_handle h : t->scenes) {
            launch_fill_u32(h->stream, (uint32_t*)h->gradx, 0u, 64 * (uint64_t)h->n);
            NPT_HIP(t, hipStreamSynchronize(h->stream));
        }
        t->cap_scale *= 1.5;
        npt_drop_graph(t);
        return npt_step_sequential(t, false);
    }
    return npt_finish_scenes(t, true);
}

}  // namespace

extern "C" {

int npt_create(nps_handle* scenes, int32_t n_scenes, int32_t lanes, npt_handle* out) {
    if (!out || !scenes || n_scenes <= 0 || lanes <= 0 || lanes > 16) return NPS_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < n_scenes; ++i)
        if (!scenes[i] || !scenes[i]->ext || scenes[i]->device != scenes[0]->device) return NPS_ERR_ARG;
    npt_ctx* t = new npt_ctx();
    t->scenes.assign(scenes, scenes + n_scenes);
    t->device = scenes[0]->device;
    auto fail = [&]() {
        npt_destroy(t);
        return NPS_ERR_HIP;
    };
    if (hipSetDevice(t->device) != hipSuccess) return fail();
    if (hipStreamCreateWithFlags(&t->main, hipStreamNonBlocking) != hipSuccess) return fail();
    if (hipEventCreateWithFlags(&t->fork, hipEventDisableTiming) != hipSuccess) return fail();
    if (hipEventCreateWithFlags(&t->joined, hipEventDisableTiming) != hipSuccess) return fail();
    t->lanes.resize(lanes);
    t->lane_done.assign(lanes, nullptr);
    t->chain_done.assign(lanes, nullptr);
    for (int l = 0; l < lanes; ++l) {
        if (hipStreamCreateWithFlags(&t->lanes[l].st, hipStreamNonBlocking) != hipSuccess) return fail();
        if (hipEventCreateWithFlags(&t->lane_done[l], hipEventDisableTiming) != hipSuccess) return fail();
        if (hipEventCreateWithFlags(&t->chain_done[l], hipEventDisableTiming) != hipSuccess) return fail();
    }
    if (ensure_dev(t->flags, 64 * (size_t)lanes, t->main, t->err) != NPS_OK) return fail();
    launch_fill_u32(t->main, (uint32_t*)t->flags.p, 0u, 16 * (uint64_t)lanes);
    if (hipStreamSynchronize(t->main) != hipSuccess) return fail();
    for (nps_handle h : t->scenes)
        if (std::find(h->steps.begin(), h->steps.end(), t) == h->steps.end()) h->steps.push_back(t);
    *out = t;
    return NPS_OK;
}

void npt_destroy(npt_handle t) {
    if (!t) return;
    if (!t->detached)
        for (nps_handle h : t->scenes) {
            auto& v = h->steps;
            v.erase(std::remove(v.begin(), v.end(), t), v.end());
        }
    (void)hipSetDevice(t->device);
    if (t->main) (void)hipStreamSynchronize(t->main);
    npt_drop_graph(t);
    for (auto& ln : t->lanes) {
        if (ln.st) {
            (void)hipStreamSynchronize(ln.st);
            (void)hipStreamDestroy(ln.st);
        }
        free_ext_view(ln.x);
    }
    for (hipEvent_t e : t->lane_done)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : t->chain_done)
        if (e) (void)hipEventDestroy(e);
    if (t->fork) (void)hipEventDestroy(t->fork);
    if (t->joined) (void)hipEventDestroy(t->joined);
    free_buf(t->view_loss);
    free_buf(t->view_g2);
    free_buf(t->view_counts);
    free_buf(t->flags);
    if (t->main) (void)hipStreamDestroy(t->main);
    delete t;
}

}  // extern "C"

// A scene of the step is being destroyed: the step stops using the scenes (its
// graph may hold their buffers), unregisters from the others and refuses every
// later call but npt_destroy.
static void npt_detach(npt_ctx* t) {
    (void)hipSetDevice(t->device);
    if (t->main) (void)hipStreamSynchronize(t->main);
    for (auto& ln : t->lanes)
        if (ln.st) (void)hipStreamSynchronize(ln.st);
    npt_drop_graph(t);
    for (nps_handle h : t->scenes) {
        auto& v = h->steps;
        v.erase(std::remove(v.begin(), v.end(), t), v.end());
    }
    t->detached = true;
    t->err = "a scene of this step was destroyed";
}

extern "C" {

const char* npt_last_error(npt_handle t) { return t ? t->err.c_str() : "invalid handle"; }

int npt_add_view(npt_handle t, int32_t scene, const float vp[16], const float vp_inv[16], const float* rgba_dev,
                 int32_t w, int32_t hgt) {
    if (!t) return NPS_ERR_ARG;
    if (t->detached) return NPS_ERR_STATE;
    if (scene < 0 || scene >= (int)t->scenes.size() || !vp || !vp_inv || !rgba_dev || w <= 0 || hgt <= 0)
        return npt_fail(t, NPS_ERR_ARG, "bad view");
    npt_view v;
    v.scene = scene;
    memcpy(v.vp, vp, sizeof(v.vp));
    memcpy(v.vp_inv, vp_inv, sizeof(v.vp_inv));
    v.tgt = rgba_dev;
    v.W = (uint32_t)w;
    v.H = (uint32_t)hgt;
    v.tune = t->scenes[scene]->tune;
    t->views.push_back(v);
    // step order: most pixels first (longest-processing-time order), then index;
    // lanes are filled greedily with the least-loaded lane first
    npt_assign_lanes(t);
    if (ensure_dev(t->view_loss, 16 * t->views.size(), t->main, t->err) != NPS_OK) return NPS_ERR_HIP;
    npt_drop_graph(t);
    return NPS_OK;
}

int npt_step(npt_handle t, int32_t mode) {
    if (!t) return NPS_ERR_ARG;
    if (t->detached) return NPS_ERR_STATE;
    if (t->views.empty()) return npt_fail(t, NPS_ERR_STATE, "no views");
    NPT_HIP(t, hipSetDevice(t->device));
    // the step's chain kernels add into the scenes' gradients (never store)
    for (nps_handle h : t->scenes) h->gradx_zero = false;
    if (mode == 0 || mode == 2) return npt_step_sequential(t, mode == 2);
    if (mode == 1) return npt_step_graph(t);
    return npt_fail(t, NPS_ERR_ARG, "mode must be 0 (per view), 1 (graph) or 2 (per view, serial)");
}

int npt_autotune(npt_handle t) {
    if (!t) return NPS_ERR_ARG;
    if (t->detached) return NPS_ERR_STATE;
    if (t->views.empty()) return npt_fail(t, NPS_ERR_STATE, "no views");
    NPT_HIP(t, hipSetDevice(t->device));
    std::vector<hipEvent_t> ev(8, nullptr);
    for (auto& e : ev) NPT_HIP(t, hipEventCreate(&e));
    int rc = NPS_OK;
    // 1. per (scene, resolution) class, the fastest (tile, batch) of one view alone
    const std::vector<ExtTune> untuned = [&] {
        std::vector<ExtTune> u;
        for (const auto& w : t->views) u.push_back(t->scenes[w.scene]->tune);
        return u;
    }();
    std::vector<ExtTune> tuned = untuned;
    std::vector<bool> done(t->views.size(), false);
    for (int vi : t->order) {
        if (done[vi]) continue;
        const npt_view& w = t->views[vi];
        float t9[9];
        ExtTune best = w.tune;
        rc = ext_time_variants(scene_of(t->scenes[w.scene]), w.vp, w.vp_inv, w.W, w.H, (const float4*)w.tgt,
                               t->lanes[0].x, t->lanes[0].st, ev.data(), t->err, t9, &best);
        if (rc != NPS_OK) break;
        // the same choice for every view of this scene at this resolution
        for (size_t k = 0; k < t->views.size(); ++k) {
            const npt_view& o = t->views[k];
            if (o.scene == w.scene && o.W == w.W && o.H == w.H) {
                tuned[k] = best;
                done[k] = true;
            }
        }
    }
    // 2. in context: the whole step's views as one captured graph (no chain rule,
    // no Adam: the scenes are untouched) for {scene defaults, per-class choices} x
    // {all lanes, half of them}; the fastest replay wins
    const int L = (int)t->lanes.size();
    const int lane_opts[2] = {L, std::max(1, L / 2)};
    int best_c = -1;
    float best_ms = 1e30f;
    for (int c = 0; rc == NPS_OK && c < 4; ++c) {
        t->tune_ms[c] = -1.0f;
        if (c % 2 == 1 && lane_opts[1] == lane_opts[0]) continue;
        const std::vector<ExtTune>& tu = c < 2 ? untuned : tuned;
        for (size_t k = 0; k < t->views.size(); ++k) t->views[k].tune = tu[k];
        t->lanes_used = lane_opts[c % 2];
        npt_assign_lanes(t);
        if ((rc = npt_plan(t, true)) != NPS_OK) break;
        float tmin = 1e30f;
        for (int rep = 0; rep < 4 && rc == NPS_OK; ++rep) {  // a warm-up replay, then the best of 3
            if (hipEventRecord(ev[0], t->main) != hipSuccess || hipGraphLaunch(t->exec, t->main) != hipSuccess ||
                hipEventRecord(ev[1], t->main) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess) {
                rc = npt_fail(t, NPS_ERR_HIP, "autotune replay");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
            if (rep > 0) tmin = std::min(tmin, ms);
        }
        npt_drop_graph(t);
        t->tune_ms[c] = tmin;
        if (tmin < best_ms) {
            best_ms = tmin;
            best_c = c;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (best_c < 0) best_c = 0;
    const std::vector<ExtTune>& tu = best_c < 2 ? untuned : tuned;
    for (size_t k = 0; k < t->views.size(); ++k) t->views[k].tune = tu[k];
    t->lanes_used = lane_opts[best_c % 2];
    npt_assign_lanes(t);
    npt_drop_graph(t);
    return rc;
}

int npt_graph_info(npt_handle t, int64_t info[2], float tune_ms[4]) {
    if (!t) return NPS_ERR_ARG;
    if (info) {
        info[0] = t->graph_nodes;
        info[1] = t->lanes_used > 0 ? t->lanes_used : (int64_t)t->lanes.size();
    }
    if (tune_ms)
        for (int c = 0; c < 4; ++c) tune_ms[c] = t->tune_ms[c];
    return NPS_OK;
}

int npt_read_tune(npt_handle t, int32_t* tile_batch2v) {
    if (!t || !tile_batch2v) return NPS_ERR_ARG;
    for (size_t k = 0; k < t->views.size(); ++k) {
        tile_batch2v[2 * k] = t->views[k].tune.tile;
        tile_batch2v[2 * k + 1] = t->views[k].tune.batch;
    }
    return NPS_OK;
}

int npt_read_losses(npt_handle t, double* loss2v) {
    if (!t || !loss2v) return NPS_ERR_ARG;
    if (t->detached) return NPS_ERR_STATE;
    NPT_HIP(t, hipMemcpy(loss2v, t->view_loss.p, 16 * t->views.size(), hipMemcpyDeviceToHost));
    return NPS_OK;
}

// ---------------- target blur ----------------

}  // extern "C"

// Fixed scratch of npb_blur_dev_noalloc: the 8-bit linear temp of the two
// blur passes (uchar4 per pixel), at most kNoAllocPixels pixels.  Every GPU the
// library runs on holds its own copy until the process ends; calls on one GPU
// share it, so calls that overlap on one GPU race.  Nothing in it has to start
// at zero: the horizontal pass writes every pixel the vertical pass reads.
namespace {
constexpr int64_t kNoAllocPixels = 2097152;  // 2^21 pixels (8 MiB): 1920x1080 fits
constexpr int kMaxScratchDevices = 64;
}  // namespace
static __device__ __align__(256) uchar4 scratch[kNoAllocPixels];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up
// once per device; the host-side name is only a stand-in for it.
int resolve_scratch(int device, void** out, std::string& err) {
    if (device < 0 || device >= kMaxScratchDevices) {
        err = "device index out of range for the no-allocation scratch";
        return NPS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return NPS_ERR_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return NPS_OK;
}
}  // namespace

extern "C" {

static int npb_fail(npb_handle b, int code, const std::string& msg) {
    b->err = msg;
    return code;
}

static int npb_ensure(npb_handle b, DevBuf& d, size_t bytes) {
    if (bytes <= d.bytes && d.p) return NPS_OK;
    if (d.p) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipFree(d.p);
        d.p = nullptr;
        d.bytes = 0;
    }
    hipError_t e = hipMalloc(&d.p, bytes);
    if (e != hipSuccess) {
        d.p = nullptr;
        return npb_fail(b, NPS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    d.bytes = bytes;
    return NPS_OK;
}

int npb_create(int32_t device, npb_handle* out) {
    if (!out) return NPS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return NPS_ERR_HIP;
    npb_ctx* b = new npb_ctx();
    b->device = device;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return NPS_ERR_HIP;
    }
    // set-up of npb_blur_dev_noalloc: the scratch address on this device
    void* sp = nullptr;
    if (resolve_scratch(device, &sp, b->err) != NPS_OK) {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return NPS_ERR_HIP;
    }
    *out = b;
    return NPS_OK;
}

void npb_destroy(npb_handle b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_buf(b->src); free_buf(b->dst); free_buf(b->temp);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* npb_last_error(npb_handle b) { return b ? b->err.c_str() : "invalid handle"; }

static int npb_run(npb_handle b, const float* src, float* dst, int32_t w, int32_t hh, int32_t radius, float sigma,
                   int32_t is_srgb, bool dev) {
    if (!b) return NPS_ERR_ARG;
    if (!src || !dst || w <= 0 || hh <= 0 || radius < 0) return npb_fail(b, NPS_ERR_ARG, "bad blur arguments");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    const size_t bytes = 16 * (size_t)w * hh;
    int rc;
    if ((rc = npb_ensure(b, b->temp, bytes)) != NPS_OK) return rc;
    const float4* s = (const float4*)src;
    float4* d = (float4*)dst;
    if (!dev) {
        if ((rc = npb_ensure(b, b->src, bytes)) != NPS_OK) return rc;
        if ((rc = npb_ensure(b, b->dst, bytes)) != NPS_OK) return rc;
        if ((e = hipMemcpyAsync(b->src.p, src, bytes, hipMemcpyHostToDevice, b->stream)) != hipSuccess)
            return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
        s = (const float4*)b->src.p;
        d = (float4*)b->dst.p;
    }
    launch_blur(b->stream, s, d, (float4*)b->temp.p, w, hh, radius, sigma, is_srgb ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    if (!dev && (e = hipMemcpyAsync(dst, b->dst.p, bytes, hipMemcpyDeviceToHost, b->stream)) != hipSuccess)
        return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    return NPS_OK;
}

int npb_blur(npb_handle b, const float* src, float* dst, int32_t w, int32_t hh, int32_t radius, float sigma,
             int32_t is_srgb) {
    return npb_run(b, src, dst, w, hh, radius, sigma, is_srgb, false);
}

int npb_blur_dev(npb_handle b, const float* src, float* dst, int32_t w, int32_t hh, int32_t radius, float sigma,
                 int32_t is_srgb) {
    return npb_run(b, src, dst, w, hh, radius, sigma, is_srgb, true);
}

int npb_blur_dev_noalloc(npb_handle b, const float* src, float* dst, int32_t w, int32_t hh, int32_t radius,
                         float sigma, int32_t is_srgb) {
    if (!b) return NPS_ERR_ARG;
    if (!src || !dst || w <= 0 || hh <= 0 || radius < 0) return npb_fail(b, NPS_ERR_ARG, "bad blur arguments");
    if ((int64_t)w * hh > kNoAllocPixels)
        return npb_fail(b, NPS_ERR_ARG, "npb_blur_dev_noalloc: " + std::to_string(w) + " x " + std::to_string(hh) +
                                            " pixels exceed the limit of " + std::to_string(kNoAllocPixels) +
                                            " (use npb_blur_dev)");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    void* sp = nullptr;
    int rc = resolve_scratch(b->device, &sp, b->err);
    if (rc != NPS_OK) return rc;
    launch_blur_u8temp(b->stream, (const float4*)src, (float4*)dst, (uchar4*)sp, w, hh, radius, sigma,
                       is_srgb ? 1 : 0);
    if ((e = hipGetLastError()) != hipSuccess) return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return npb_fail(b, NPS_ERR_HIP, hipGetErrorString(e));
    return NPS_OK;
}

int npd_compose(int32_t device, const float* const* textures, const int32_t* dims, const float enabled[3],
                float* canvas, int32_t W, int32_t H) {
    if (!textures || !dims || !enabled || !canvas || W <= 0 || H <= 0) return NPS_ERR_ARG;
    for (int k = 0; k < 8; ++k)
        if (!textures[k] || dims[2 * k] <= 0 || dims[2 * k + 1] <= 0) return NPS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return NPS_ERR_HIP;
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return NPS_ERR_HIP;
    DisplayArgs a;
    void* bufs[9] = {nullptr};
    int rc = NPS_OK;
    for (int k = 0; k < 8 && rc == NPS_OK; ++k) {
        const size_t bytes = 16 * (size_t)dims[2 * k] * dims[2 * k + 1];
        if (hipMalloc(&bufs[k], bytes) != hipSuccess ||
            hipMemcpyAsync(bufs[k], textures[k], bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = NPS_ERR_HIP;
        a.tex[k] = DisplayTex{(const float4*)bufs[k], dims[2 * k], dims[2 * k + 1]};
    }
    a.edge_on = enabled[0] > 0.5f;
    a.base_color_on = enabled[1] > 0.5f;
    a.color_on = enabled[2] > 0.5f;
    const size_t cbytes = 16 * (size_t)W * H;
    if (rc == NPS_OK && hipMalloc(&bufs[8], cbytes) != hipSuccess) rc = NPS_ERR_HIP;
    if (rc == NPS_OK) {
        launch_compose(st, a, (float4*)bufs[8], W, H);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(canvas, bufs[8], cbytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = NPS_ERR_HIP;
    }
    (void)hipStreamSynchronize(st);
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    (void)hipStreamDestroy(st);
    return rc;
}

int nps_dev_alloc(nps_handle h, int64_t bytes, void** out) {
    if (!h || !out || bytes < 0) return NPS_ERR_ARG;
    NPS_HIP(h, hipSetDevice(h->device));
    *out = nullptr;
    NPS_HIP(h, hipMalloc(out, (size_t)std::max<int64_t>(bytes, 1)));
    return NPS_OK;
}

int nps_dev_free(nps_handle h, void* p) {
    if (!h) return NPS_ERR_ARG;
    NPS_HIP(h, hipSetDevice(h->device));
    NPS_HIP(h, hipStreamSynchronize(h->stream));
    if (p) NPS_HIP(h, hipFree(p));
    return NPS_OK;
}

int nps_dev_upload(nps_handle h, void* dst, const void* src, int64_t bytes) {
    if (!h || !dst || !src || bytes < 0) return NPS_ERR_ARG;
    NPS_COPY(h, dst, src, (size_t)bytes, hipMemcpyHostToDevice);
    return NPS_OK;
}

int nps_debug_math(int32_t device, int32_t op, const float* x, const float* y, float* out, int32_t n) {
    if (!x || !out || n < 0 || ((op == 2 || op == 8) && !y)) return NPS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return NPS_ERR_HIP;
    float *dx = nullptr, *dy = nullptr, *dout = nullptr;
    size_t bytes = 4 * (size_t)std::max(n, 1);
    int rc = NPS_OK;
    if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
        hipMalloc(&dout, bytes) != hipSuccess) {
        rc = NPS_ERR_HIP;
    } else {
        hipMemcpy(dx, x, 4 * (size_t)n, hipMemcpyHostToDevice);
        if (y) hipMemcpy(dy, y, 4 * (size_t)n, hipMemcpyHostToDevice);
        launch_debug_math(nullptr, op, dx, dy, dout, (uint32_t)n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = NPS_ERR_HIP;
        else hipMemcpy(out, dout, 4 * (size_t)n, hipMemcpyDeviceToHost);
    }
    if (dx) hipFree(dx);
    if (dy) hipFree(dy);
    if (dout) hipFree(dout);
    return rc;
}

int nps_debug_depth_sort(int32_t device, const uint64_t* keys, uint32_t n, int32_t narrow, uint64_t* out,
                         uint32_t flags_out[2]) {
    if ((!keys && n) || (!out && n) || !flags_out || n > 0x7FFFFFFFu) return NPS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return NPS_ERR_HIP;
    flags_out[0] = flags_out[1] = 0u;
    if (n == 0) return NPS_OK;
    // per 256 keys: the least depth bits of the drawn ones (the preprocess's
    // workgroup minima)
    const uint32_t npart = (n + 255) / 256;
    std::vector<uint32_t> part(npart, 0xFFFFFFFFu);
    for (uint32_t i = 0; i < n; ++i) part[i / 256] = std::min(part[i / 256], (uint32_t)(keys[i] >> 32));
    const size_t tmp_bytes = std::max(radix_temp_bytes(n, true), depth_sort_temp_bytes(n));
    uint64_t *dk = nullptr, *da = nullptr;
    uint32_t *dp = nullptr, *df = nullptr;
    void* tmp = nullptr;
    int rc = NPS_OK;
    if (hipMalloc(&dk, 8 * (size_t)n) != hipSuccess || hipMalloc(&da, 8 * (size_t)n) != hipSuccess ||
        hipMalloc(&dp, 4 * (size_t)npart) != hipSuccess || hipMalloc(&df, 64) != hipSuccess ||
        hipMalloc(&tmp, tmp_bytes) != hipSuccess) {
        rc = NPS_ERR_HIP;
    } else {
        hipMemcpy(dk, keys, 8 * (size_t)n, hipMemcpyHostToDevice);
        hipMemcpy(dp, part.data(), 4 * (size_t)npart, hipMemcpyHostToDevice);
        launch_fill_u32(nullptr, df, 0u, 16);
        bool alt = false;
        if (narrow) depth_sort_narrow(nullptr, dk, da, n, tmp, dp, npart, df, nullptr, nullptr);
        else alt = radix_sort_u64(nullptr, dk, da, nullptr, nullptr, n, 32, 64, tmp, nullptr, true);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            rc = NPS_ERR_HIP;
        } else {
            hipMemcpy(out, alt ? da : dk, 8 * (size_t)n, hipMemcpyDeviceToHost);
            hipMemcpy(flags_out, df, 8, hipMemcpyDeviceToHost);
        }
    }
    if (dk) hipFree(dk);
    if (da) hipFree(da);
    if (dp) hipFree(dp);
    if (df) hipFree(df);
    if (tmp) hipFree(tmp);
    return rc;
}

}  // extern "C"
