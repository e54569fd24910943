// k_display.hip -- the debug display compositor of the viewer's final pass
// (src/lib/gpu/splat_render.wgsl:24-181): one thread per canvas pixel, the
// fragment shader's arithmetic in its written order (FP32, no contraction).
//
// Canvas pixel (c, r) is shaded at uv = ((c + 0.5) / W, (r + 0.5) / H), the
// interpolated uv of a full-screen quad at the pixel centre.  The bottom strip
// (uv.y > 1 - 0.18) shows six panels (splat view, target depth, target edges,
// edge layer, base-colour layer, colour layer), each fitted into its panel with
// pillar- or letterboxing; the main area shows the path-traced image (or the
// raster target where it has no sample yet, alpha 0) on the left and the splat
// view with the three Bezier layers composited over it on the right.
// Texel reads clamp to the texture (WGSL leaves an out-of-range textureLoad
// implementation-defined).
#include "nps_common.h"
#include "nps_kernels.h"

namespace nps {

namespace {
constexpr float kStrip = 0.18f;   // SPLAT_RENDER_STRIP_HEIGHT (gpu/constants.ts:38)
constexpr float kPanels = 6.0f;   // SPLAT_RENDER_NUM_PANELS (gpu/constants.ts:39)

__device__ __forceinline__ float4 tex_load(const DisplayTex& t, int x, int y) {
    x = min(max(x, 0), t.w - 1);
    y = min(max(y, 0), t.h - 1);
    return t.p[(size_t)y * t.w + x];
}
// vec2i(uv * dims): truncation toward zero
__device__ __forceinline__ float4 tex_at(const DisplayTex& t, float u, float v) {
    return tex_load(t, (int)(u * (float)t.w), (int)(v * (float)t.h));
}
// fitInPanel (splat_render.wgsl:44-62); false where the panel shows background
__device__ __forceinline__ bool fit_in_panel(float pu, float pv, float panel_aspect, float src_aspect, float& fu,
                                             float& fv) {
    const float rel = panel_aspect / src_aspect;
    if (rel > 1.0f) {
        const float content_width = 1.0f / rel;
        const float margin = (1.0f - content_width) * 0.5f;
        if (pu < margin || pu > 1.0f - margin) return false;
        fu = (pu - margin) / content_width;
        fv = pv;
    } else {
        const float content_height = rel;
        const float margin = (1.0f - content_height) * 0.5f;
        if (pv < margin || pv > 1.0f - margin) return false;
        fu = pu;
        fv = (pv - margin) / content_height;
    }
    return true;
}
}  // namespace

__global__ __launch_bounds__(256) void k_compose(DisplayArgs a, float4* __restrict__ canvas, int W, int H) {
    const int c = blockIdx.x * 16 + (threadIdx.x & 15);
    const int r = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (c >= W || r >= H) return;
    const float ux = ((float)c + 0.5f) / (float)W, uy = ((float)r + 0.5f) / (float)H;
    const DisplayTex& target = a.tex[0];
    const DisplayTex& splat = a.tex[1];
    const float splat_aspect = (float)target.w / (float)target.h;
    const float screen_aspect = (splat_aspect * 2.0f) * (1.0f - kStrip);
    float4 out;
    if (uy > 1.0f - kStrip) {
        const float strip_v = (uy - (1.0f - kStrip)) / kStrip;
        const float panel_idx = floorf(ux * kPanels);
        const float panel_u = ux * kPanels - floorf(ux * kPanels);
        const float4 border = make_float4(0.3f, 0.3f, 0.3f, 1.0f);
        const float4 bg = make_float4(0.05f, 0.05f, 0.05f, 1.0f);
        if (panel_u < 0.005f || panel_u > 0.995f || strip_v < 0.01f) {
            out = border;
        } else {
            const float panel_aspect = screen_aspect / (kPanels * kStrip);
            const DisplayTex& dep = a.tex[2];
            const DisplayTex& edg = a.tex[3];
            const float depth_aspect = (float)dep.w / (float)dep.h;
            const float edge_aspect = (float)edg.w / (float)edg.h;
            // panel -> (texture, its aspect, grey from .r)
            int k;
            float src_aspect;
            bool grey;
            if (panel_idx < 0.5f) { k = 1; src_aspect = splat_aspect; grey = false; }
            else if (panel_idx < 1.5f) { k = 2; src_aspect = depth_aspect; grey = true; }
            else if (panel_idx < 2.5f) { k = 3; src_aspect = edge_aspect; grey = true; }
            else if (panel_idx < 3.5f) { k = 4; src_aspect = splat_aspect; grey = true; }
            else if (panel_idx < 4.5f) { k = 5; src_aspect = splat_aspect; grey = false; }
            else { k = 6; src_aspect = splat_aspect; grey = false; }
            float fu = 0.0f, fv = 0.0f;
            if (!fit_in_panel(panel_u, strip_v, panel_aspect, src_aspect, fu, fv)) {
                out = bg;
            } else {
                const float4 t = tex_at(a.tex[k], fu, fv);
                out = grey ? make_float4(t.x, t.x, t.x, 1.0f) : t;
            }
        }
    } else {
        const float main_v = uy / (1.0f - kStrip);
        if (ux < 0.498f) {
            const float tx = ux * 2.0f;
            const float4 pt = tex_at(a.tex[7], tx, main_v);
            const float4 raster = tex_at(target, tx, main_v);
            out = pt.w > 0.0f ? pt : raster;
        } else if (ux < 0.502f) {
            out = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        } else {
            const float ru = (ux - 0.5f) * 2.0f;
            const float4 base = tex_at(splat, ru, main_v);
            float e = tex_at(a.tex[4], ru, main_v).x;
            e = fminf(fmaxf(e, 0.0f), 1.0f);
            const float4 bcb = tex_at(a.tex[5], ru, main_v);
            const float4 cb = tex_at(a.tex[6], ru, main_v);
            float x = base.x, y = base.y, z = base.z;
            if (a.base_color_on) {
                x = x * (1.0f - bcb.w) + bcb.x;
                y = y * (1.0f - bcb.w) + bcb.y;
                z = z * (1.0f - bcb.w) + bcb.z;
            }
            if (a.color_on) {
                x = x * (1.0f - cb.w) + cb.x;
                y = y * (1.0f - cb.w) + cb.y;
                z = z * (1.0f - cb.w) + cb.z;
            }
            if (a.edge_on) {  // mix(composite, 1, e) = composite (1 - e) + 1 e
                x = x * (1.0f - e) + 1.0f * e;
                y = y * (1.0f - e) + 1.0f * e;
                z = z * (1.0f - e) + 1.0f * e;
            }
            out = make_float4(x, y, z, 1.0f);
        }
    }
    canvas[(size_t)r * W + c] = out;
}

void launch_compose(hipStream_t st, const DisplayArgs& a, float4* canvas, int W, int H) {
    if (W <= 0 || H <= 0) return;
    dim3 g((W + 15) / 16, (H + 15) / 16), b(256);
    hipLaunchKernelGGL(k_compose, g, b, 0, st, a, canvas, W, H);
}

}  // namespace nps
