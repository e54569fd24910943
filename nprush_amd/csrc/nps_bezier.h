// nps_bezier.h -- records and launch arguments of the Bezier stroke path
// (bezier_backward.wgsl, bezier_step.wgsl, bezier_adc.wgsl) and of its display
// raster (bezier_forward.wgsl).
//
// Curve layout in HBM (the reference's Bezier struct, bezier_backward.wgsl:4-10):
// 20 floats = p0.xyz width | p1.xyz softness | p2.xyz pad | p3.xyz pad | rgba.
// Gradients: 18 i32 per curve in the order p0.xyz p1.xyz p2.xyz p3.xyz rgb a
// width softness, fixed point 1e4 (positions, width, softness) and 1e5 (colour,
// opacity) (bezier_backward.wgsl:536-558).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nps_common.h"

namespace nps {

constexpr int kBzSeg = 16;            // polyline segments per curve (bezier_backward.wgsl:52)
constexpr float kBzKill = 0.005f;     // BEZIER_KILL_ALPHA_THRESH (constants.ts)
constexpr int kBzCap = 1024;          // BEZIER_MAX_TILE_BEZIERS
constexpr int kBzParams = 18;

// Per curve and view, read by every pixel of the curve's tiles.  The forward
// part is what the per-pixel distance loop needs; the reverse part adds the
// backprojection of screen gradients to world space.
struct BzFwd {
    float x[kBzSeg + 1], y[kBzSeg + 1];  // P0 = projected p0, Pk = bezier_at(k / 16), screen space
    float w[4];                          // clip.w of the control points
    float width, soft, op, r, g, b;      // width, softness floored at 0.001
    float cx0, cy0, cx1, cy1;            // box of P0..P16
    float rcull;                         // beyond this distance from the box every alpha is exactly 0
    float pad[3];
    float gb[16];                        // boxes (x0, y0, x1, y1) of the points of segments 4g+1..4g+4
};
struct BzBwd {
    float q0x, q0y;   // bezier_at(0): the reverse pass's first segment start
    float J0[12];     // d screen.x / d world (x, y, z) per control point
    float J1[12];     // d screen.y / d world
    float pad[2];
};
static_assert(sizeof(BzFwd) == 272, "BzFwd layout");
static_assert(sizeof(BzBwd) == 112, "BzBwd layout");

// BezierUniforms scalars (bezier_backward.wgsl:20-32) and AdamState.no_kill.
struct BzParams {
    float mode;         // > 0.5: colour mode (background textures, depth loss); else edge mode
    float max_width;    // > 0: fine layer (softness and direction regularisers, width cap)
    float prune_alpha;  // > 0 overrides the prune threshold 0.001
    float prune_width;
    float bg_penalty;   // opacity penalty on background pixels
    float no_kill;      // >= 0.5: no loss-based or off-screen kills
};

struct BzCompositeArgs {
    const BzFwd* fwd;
    const BzBwd* bwd;
    const uint64_t* keys;      // sorted (tile << 32 | curve id)
    const uint32_t* ranges;    // [2 * num_tiles]
    const float4* tgt;         // H*W rgba (edge mode: .r is the edge target)
    const float* tgt_depth;    // H*W
    const float4* bg;          // H*W rgba
    const float* bg_depth;     // H*W
    const float4* normal;      // H*W rgba
    int32_t* grads;            // 18 n
    unsigned long long* loss_acc;  // n, 2^-32 fixed point
    int32_t* pixel_loss;       // H*W, fixed point 1e4
    float* c_img;              // H*W*3
    float* d_img;              // H*W
    float* t_img;              // H*W final transmittance
    double* tile_loss;         // 2 * num_tiles: sum |C - tgt|^2, sum |C - tgt|
    float* ckpt;               // num_tiles * nchunk_max * 256
    float* pa;                 // per (key, pixel of its tile): alpha, depth and segment of the forward,
    float* pd;                 //   written for entries where some lane of the wave has alpha > 0
    uint8_t* pk;
    uint32_t* act;             // per tile, wave and 32 entries: some pixel has alpha >= 0.001
    uint32_t* nzm;             // per tile, wave and 32 entries: pair values stored (some alpha > 0)
    float* seeds;              // 9 planes of num_tiles * 256: dC rgb, dT, dD, is_bg, flow xy, has flow
    unsigned long long* part;  // 5 per tile (PartMap slots): the counters stats[0], [4], [2], [3], [5]
    unsigned long long* stats; // [0] pairs, [2] reverse pairs, [3] reverse wave-entries, [4] forward and
                               // [5] reverse wave-entries reloaded from the forward's pair buffers
    BzParams prm;
    uint32_t cap;
    uint32_t nchunk_max;
    int backward;              // 0: image only
};

struct BzAdcBuffers {
    uint32_t* dead_list;  // n
    uint32_t* cand;       // n
    int64_t* stats;       // 6 (device)
};

// Display raster (bezier_forward.wgsl): per curve and view, what each pixel of
// the curve's tiles reads.  Widths are floored at 1e-4 there, 0.001 in training.
struct BzDraw {
    float x[kBzSeg + 1], y[kBzSeg + 1];  // P0 = projected p0, Pk = bezier_at(k / 16)
    float width, soft, op, r, g, b;
    float qx0, qy0, qx1, qy1;            // the quad: control-point hull -+ (width + softness), clamped to +-4
    float cx0, cy0, cx1, cy1;            // box of P0..P16
    float rcull;                         // beyond this distance from the box every alpha is exactly 0
    float pad[3];
    float gb[16];                        // boxes (x0, y0, x1, y1) of the points of segments 4g+1..4g+4
};
static_assert(sizeof(BzDraw) == 272, "BzDraw layout");

struct BzDrawArgs {
    const BzDraw* rec;
    const uint64_t* keys;    // sorted (tile << 32 | curve id)
    const uint32_t* ranges;  // [2 * num_tiles]
    const float* brush;      // bh rows of bw texel values (r / 255)
    int32_t bw, bh;
    float4* rgba;            // H*W, read when !clear
    int clear, quantize;
};

void launch_bz_draw_preprocess(hipStream_t st, const float* curves, uint32_t n, const ViewParams& v, BzDraw* rec,
                               uint32_t* counts, uint4* rects);
void launch_bz_draw(hipStream_t st, const ViewParams& v, const BzDrawArgs& a);

void launch_bz_preprocess(hipStream_t st, const float* curves, uint32_t n, const ViewParams& v, BzFwd* fwd,
                          BzBwd* bwd, uint32_t* counts, uint4* rects);
void launch_bz_composite(hipStream_t st, const ViewParams& v, const BzCompositeArgs& a);
// t: the step number of this update (the reference's adam.t + 1)
void launch_bz_step(hipStream_t st, float* curves, int32_t* grads, float* m, float* v, float* grad_accum, uint32_t n,
                    const float vp[16], const BzParams& prm, float t, float pixel_count);
void launch_bz_adc(hipStream_t st, float* curves, float* m, float* v, float* grad_accum, unsigned long long* loss_acc,
                   int32_t* pixel_loss, uint32_t n, uint32_t W, uint32_t H, const float vp[16], const float vp_inv[16],
                   const BzParams& prm, float adam_t, const BzAdcBuffers& b);

}  // namespace nps
