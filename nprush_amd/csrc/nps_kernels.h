// nps_kernels.h -- host-side launchers of the gfx950 kernels (one translation
// unit per subsystem; every launcher enqueues on the given stream).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nps_common.h"
#include "nps_ext.h"

namespace nps {

struct CompositeArgs {
    const FwdRec* fwd;
    const BwdRec* bwd;
    const uint64_t* keys;     // sorted (tile << 32 | splat id)
    const uint32_t* ranges;   // [2 * num_tiles]
    const float4* tgt_rgba;   // H*W
    const float* tgt_depth;   // H*W
    int32_t* grads;           // 16N fixed point
    float* c_img;             // H*W*3
    float* d_img;             // H*W
    double* tile_loss;        // 2 * num_tiles
    float* ckpt;              // num_tiles * nchunk_max * 256
    unsigned long long* stats;  // [0] pairs, [1] tiles over the cap, [2] pairs in the reverse pass,
                                // [3] reverse wave-entries, [4] forward wave-entries evaluating pow/exp
    unsigned long long* part;   // 4 per tile (PartMap slots): the counters stats[0], [2], [3], [4]
    uint32_t cap;
    uint32_t nchunk_max;
};

// k_bin.hip
void launch_preprocess(hipStream_t st, const float* splats, uint32_t n, const ViewParams& v, FwdRec* fwd,
                       BwdRec* bwd, uint32_t* counts, uint4* rects);
// Per-tile id-block cut of capped lists.  diff: 4 * nsb * (tiles_x+1) * (tiles_y+1) + 16 * (tiles_x+1) i32
// scratch (overwritten); kept_counts[n] = keys each splat emits under the cut.
void launch_tile_cut(hipStream_t st, const ViewParams& v, uint32_t n, int32_t* diff, uint32_t splat_block,
                     uint32_t cap, uint32_t* cut, uint32_t* full, unsigned long long* over_cap,
                     const uint32_t* counts, const uint4* rects, uint32_t* kept_counts);
void launch_emit_keys(hipStream_t st, uint32_t n, const uint32_t* counts, const uint64_t* offsets,
                      const uint4* rects, uint32_t tiles_x, uint32_t num_tiles, const uint32_t* cut,
                      uint32_t splat_block, uint64_t* keys);
void launch_tile_ranges(hipStream_t st, const uint64_t* keys, uint64_t m, uint32_t* ranges,
                        const uint64_t* m_dev = nullptr);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint64_t n);

// k_sort.hip
// Diagnostic counters without same-address global atomics (a device-scope
// atomic on one address from every wave of a launch serialises across the
// XCDs and cost more than the work it counted): workgroup slot s of a kernel
// stores its partial counts in part[K * s + k]; launch_sum_parts then adds the
// sum over the nslot slots of each k < nk to stats[dst[k]] (one workgroup).
struct PartMap {
    uint32_t K, nk;
    int32_t dst[8];
};
void launch_sum_parts(hipStream_t st, const unsigned long long* part, uint32_t nslot, const PartMap& m,
                      unsigned long long* stats);

size_t scan_temp_bytes(uint64_t n);
void exclusive_scan_u32_to_u64(hipStream_t st, const uint32_t* in, uint64_t* out, uint64_t n, void* temp);
void exclusive_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint64_t n, void* temp);
size_t radix_temp_bytes(uint64_t m, bool wide = false);
// wide: digits of up to 11 bits (2048 buckets, fewer passes) instead of 8.
// gather_src/dst (optional): the last pass also writes gather_dst[r] =
// gather_src[low 32 bits of the key placed at r].
bool radix_sort_u64(hipStream_t st, uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt,
                    uint64_t m, int bit_lo, int bit_hi, void* temp, const uint64_t* m_dev = nullptr,
                    bool wide = false,
                    const uint32_t* gather_src = nullptr, uint32_t* gather_dst = nullptr);
// Depth order of m keys depth_bits << 32 | id (0xFFFFFFFF: not drawn) in two
// 12-bit passes over min(depth_bits - base, 2^24 - 1), base = the least of
// wmin_part[0..npart) (the preprocess's per-workgroup least drawn depth bits).
// A drawn depth at or beyond base + 2^24 - 1 stores 1 to flags[0] and flags[1]
// (the result is then not the depth order).  Result in keys; the last pass
// gathers as radix_sort_u64 does.  temp: depth_sort_temp_bytes(m).
// The sort's counters that start at zero: depth_sort_zero_words words at
// depth_sort_zero_region(temp); with counters_clear the caller has cleared
// them on the stream, else the sort clears them with a launch of its own.
size_t depth_sort_temp_bytes(uint64_t m);
uint32_t* depth_sort_zero_region(void* temp);
constexpr uint32_t depth_sort_zero_words = 512;
void depth_sort_narrow(hipStream_t st, uint64_t* keys, uint64_t* keys_alt, uint32_t m, void* temp,
                       const uint32_t* wmin_part, uint32_t npart, uint32_t* flags, const uint32_t* gather_src,
                       uint32_t* gather_dst, bool counters_clear = false);


// k_composite.hip
void launch_composite_ref(hipStream_t st, const ViewParams& v, const CompositeArgs& a);

// k_step.hip
void launch_adam_ref(hipStream_t st, float* splats, int32_t* grads, float* m, float* v, float* accum, uint32_t n,
                     float t);

// k_adc.hip
struct AdcBuffers {
    uint32_t* alive0;     // n
    uint32_t* qa;         // n
    uint32_t* dead;       // n
    uint32_t* dead_rank;  // n + 1
    uint32_t* dead_list;  // n
    uint32_t* Q;          // n
    uint32_t* K;          // n + 1
    uint32_t* qual_list;  // n
    float* tmp;           // 32 * n
    uint32_t* flag;       // 4
    void* scan_tmp;
};
// Runs the whole ADC; returns 0, or -1 when a cascade chain exceeds the depth
// bound.  stats: qualifiers, children, splits, clones, rounds.
int run_adc_ref(hipStream_t st, float* splats, float* m, float* v, float* accum, uint32_t n, float adam_t,
                const AdcBuffers& b, int64_t stats[5]);

// k_render.hip
struct RenderArgs {
    const float4* rec;        // 4 float4 per splat
    const uint64_t* keys;
    const uint32_t* ranges;
    float4* color;            // H*W
    float4* depth;            // H*W
    int clear;
    int quantize;
};
void launch_preprocess_render(hipStream_t st, const float* splats, uint32_t n, const ViewParams& v, float4* rec,
                              uint32_t* counts, uint4* rects);
void launch_render(hipStream_t st, const ViewParams& v, const RenderArgs& a);

// k_ext.hip
struct ExtCompositeArgs {
    const ExtRec* rec;
    const uint32_t* ids;        // splat ids in (tile, depth, id) order
    const uint32_t* ranges;     // [2 * num_tiles]
    const float4* tgt_rgba;     // H*W
    double* g2;                 // 9 per splat: d scr.xy, d conic A,2B,C, d rgb, d opacity
    float* c_img;               // H*W*3
    double* tile_loss;          // 2 * num_tiles: L2, L1
    float* ckpt;                // per batch slot: batch/8 chunks x tile^2 pixels of T
    uint64_t* act;              // per batch slot and wave: batch entry bits
    unsigned long long* stats;  // [0] forward pairs, [2] reverse pairs, [3] reverse wave-entries,
                                // [4] forward wave-entries
    const uint32_t* order;      // workgroup -> tile (launch_tile_order), ~0 = none
    uint32_t* tile_work;        // 2 per tile: entries staged by the forward, reverse extent
    unsigned long long* part;   // 4 per tile (PartMap slots): the counters stats[0], [4], [2], [3]
    int backward;               // 0: image and loss only
    int col_major;              // lists stored column by column (ext_bin), else tile by tile
};
// EXT binning without a key sort (k_bin_* in k_ext.hip): from the depth order
// (order: depth_bits << 32 | id, sorted) and the tile rectangles to splat ids
// grouped by tile, each tile in
// (depth, id) order (tiles column by column), and the tile ranges.
// ext_bin_fits: at most 256 tile rows and columns, the 4-byte items hold the
// column and the id, and the counts fit 32 bits (else the caller sorts keys).
// ext_bin_rows (row counts and their offsets) needs no key count; the last
// offset, roff[ext_bin_row_words(n, tiles_y) - 1], is the key count (0 when n
// is 0).  ext_bin then does the rest, ending with the ranges and the
// composite's workgroup order (launch_tile_order's output).
// rcnt / roff: ext_bin_row_words(...) words each, rscan: scan_temp_bytes of
// that; ccnt / coff: ext_bin_col_words(...) words each, cscan likewise.
bool ext_bin_fits(uint32_t n, uint32_t tiles_x, uint32_t tiles_y, uint64_t cap);
size_t ext_bin_row_words(uint32_t n, uint32_t tiles_y);
size_t ext_bin_col_words(uint32_t tiles_x, uint64_t cap);
void ext_bin_rows(hipStream_t st, uint32_t n, const uint64_t* order, const uint4* rects, uint32_t tiles_y,
                  uint32_t* rcnt, uint32_t* roff, void* rscan);
void ext_bin(hipStream_t st, uint32_t n, const uint64_t* order, const uint4* rects,
             uint32_t tiles_x, uint32_t tiles_y, uint64_t cap, const uint32_t* m_dev, const uint32_t* roff,
             uint32_t* items, uint32_t* ids, uint32_t* ranges, uint32_t* tile_order, uint32_t* ccnt, uint32_t* coff,
             uint32_t* tile_cnt, void* cscan, uint32_t* overflow, bool tile_cnt_clear = false);
void launch_keys_to_ids(hipStream_t st, const uint64_t* keys, uint64_t m, const uint64_t* m_dev, uint32_t* ids);
// Entries of the composite's workgroup order (= its grid size).
uint32_t ext_order_slots(uint32_t tiles_x, uint32_t tiles_y);
void launch_tile_order(hipStream_t st, const uint32_t* ranges, uint32_t tiles_x, uint32_t tiles_y, uint32_t* order);
void launch_ext_preprocess(hipStream_t st, const float* splats, const float* sh, uint32_t n, int deg,
                           const ViewParams& v, const float cam[3], ExtRec* rec, uint32_t* counts, uint4* rects,
                           uint64_t* dkey, uint32_t* wmin_part = nullptr, uint32_t* clr0 = nullptr,
                           uint32_t nclr0 = 0, uint32_t* clr1 = nullptr, uint32_t nclr1 = 0,
                           uint32_t* clr2 = nullptr, uint32_t nclr2 = 0);
// keys beyond `cap` are not written; *overflow is then set to 1
void launch_ext_keys(hipStream_t st, uint32_t n, const uint64_t* order, const uint32_t* rank_counts,
                     const uint4* rects, uint32_t tiles_x, const uint64_t* offsets, uint64_t* keys, uint64_t cap,
                     uint32_t* overflow);
// tile edge v.tile in {8, 16, 32}, records staged `batch` in {32, 64, 128} at a time
bool ext_tune_valid(int tile, int batch);
void launch_composite_ext(hipStream_t st, const ViewParams& v, const ExtCompositeArgs& a, int batch);
// batch slots of a view: keys/batch + tiles + 1; checkpoint floats and mask words per slot set
size_t ext_ckpt_floats(size_t slots, int tile, int batch);
size_t ext_act_words(size_t slots, int tile, int batch);
// grads_zero: the caller knows the 64N gradients are all zero (then they are
// stored, not read and added).  flags/nflags: key-overflow flags (16 words
// apart) of a captured step; if any is set the launch changes nothing.
void launch_ext_chain(hipStream_t st, const float* splats, const float* sh, uint32_t n, int deg, const ViewParams& v,
                      const float cam[3], const double* g2, const uint32_t* counts, float* grads, float* accum,
                      bool grads_zero = false, const uint32_t* flags = nullptr, uint32_t nflags = 0,
                      double* g2_clear = nullptr);
// ctl = {t, 1 - 0.9^t, 1 - 0.999^t, skip} on the device: t advances by one
// unless one of the nflags flags (16 words apart) is set, then nothing changes.
void launch_adam_ext(hipStream_t st, float* splats, float* sh, float* grads, float* m, float* v, uint32_t n, int deg,
                     float* ctl, const uint32_t* flags, uint32_t nflags);

void launch_sum_tile_loss(hipStream_t st, const double* tile_loss, uint32_t ntiles, double* out2);

// EXT densification (k_ext.hip)
struct AdcExtParams {
    float threshold;      // mean |dL/d ndc| per view that qualifies a splat
    float extent;         // scene extent
    float prune_opacity;  // sigmoid(logit) below: dead slot
    float percent_dense;  // split when the largest scale exceeds percent_dense * extent
};
struct AdcExtBuffers {
    uint32_t *dead, *qual, *dead_list, *qual_list;  // n each
    uint32_t *dead_rank, *qual_rank;                // n + 1 each (the scan writes the total at [n])
    uint32_t* cnt;  // 4: qualifiers, dead slots, splits, clones
    void* scan_tmp;
};
void run_adc_ext(hipStream_t st, float* splats, float* sh, float* m, float* v, float* accum, uint32_t n,
                 const AdcExtParams& prm, uint32_t seed, const AdcExtBuffers& b);
// qualifier/dead lists as run_adc_ext builds them, and (feat != null) 32 features
// per qualifier in placement order
void launch_adc_ext_features(hipStream_t st, const float* splats, const float* sh, const float* accum,
                             const float* m, const float* v, uint32_t n, const AdcExtParams& prm, const float cam[3],
                             const AdcExtBuffers& b, float* feat);
// adds 14 deltas per placement k < nrows to the two children of the last densification
void launch_adc_ext_deltas(hipStream_t st, float* splats, const AdcExtBuffers& b, const float* delta, uint32_t nrows);

// k_blur.hip: both passes; temp is an H*W float4 scratch image
void launch_blur(hipStream_t st, const float4* src, float4* dst, float4* temp, int W, int H, int radius, float sigma,
                 int is_srgb);
// The same two passes with the 8-bit linear temp stored as uchar4 (k of k/255):
// W*H*4 bytes of temp, bit-identical output.
void launch_blur_u8temp(hipStream_t st, const float4* src, float4* dst, uchar4* temp, int W, int H, int radius,
                        float sigma, int is_srgb);

// k_display.hip: the debug display compositor (splat_render.wgsl).  Textures in
// binding order: target, splat view, target depth, target edges, edge-layer
// Bezier view, base-colour Bezier view, colour Bezier view, path trace.
struct DisplayTex {
    const float4* p;
    int w, h;
};
struct DisplayArgs {
    DisplayTex tex[8];
    int edge_on, base_color_on, color_on;
};
void launch_compose(hipStream_t st, const DisplayArgs& a, float4* canvas, int W, int H);

// k_debug.hip
void launch_debug_math(hipStream_t st, int op, const float* x, const float* y, float* out, uint32_t n);

}  // namespace nps
