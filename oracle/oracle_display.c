/*
 * oracle_display.c -- CPU oracle of the viewer's debug display compositor
 * (test infrastructure only): src/lib/gpu/splat_render.wgsl:24-181 restated
 * per canvas pixel, FP32 in the shader's written order (no contraction).
 * Pixel (c, r) is shaded at uv = ((c + 0.5) / W, (r + 0.5) / H); texel reads
 * clamp to the texture (an out-of-range textureLoad is implementation-defined
 * in WGSL).  Textures in binding order: target, splat view, target depth,
 * target edges, edge-layer Bezier view, base-colour Bezier view, colour Bezier
 * view, path trace (rgba floats, dims[2k] x dims[2k+1]).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define STRIP 0.18f   /* SPLAT_RENDER_STRIP_HEIGHT, gpu/constants.ts:38 */
#define PANELS 6.0f   /* SPLAT_RENDER_NUM_PANELS, gpu/constants.ts:39 */

typedef struct { const float* p; int w, h; } dtex;

static void load(const dtex* t, int x, int y, float out[4]) {
    if (x < 0) x = 0;
    if (x > t->w - 1) x = t->w - 1;
    if (y < 0) y = 0;
    if (y > t->h - 1) y = t->h - 1;
    const float* q = t->p + 4 * ((size_t)y * t->w + x);
    out[0] = q[0]; out[1] = q[1]; out[2] = q[2]; out[3] = q[3];
}

static void at(const dtex* t, float u, float v, float out[4]) {
    load(t, (int)(u * (float)t->w), (int)(v * (float)t->h), out);
}

/* fitInPanel (splat_render.wgsl:44-62): 0 where the panel shows background */
static int fit(float pu, float pv, float panel_aspect, float src_aspect, float* fu, float* fv) {
    float rel = panel_aspect / src_aspect;
    if (rel > 1.0f) {
        float cw = 1.0f / rel;
        float margin = (1.0f - cw) * 0.5f;
        if (pu < margin || pu > 1.0f - margin) return 0;
        *fu = (pu - margin) / cw;
        *fv = pv;
    } else {
        float ch = rel;
        float margin = (1.0f - ch) * 0.5f;
        if (pv < margin || pv > 1.0f - margin) return 0;
        *fu = pu;
        *fv = (pv - margin) / ch;
    }
    return 1;
}

int ora_compose_f32(const float* const* textures, const int32_t* dims, const float* enabled, float* canvas, int W,
                    int H) {
    dtex t[8];
    for (int k = 0; k < 8; ++k) { t[k].p = textures[k]; t[k].w = dims[2 * k]; t[k].h = dims[2 * k + 1]; }
    const float splat_aspect = (float)t[0].w / (float)t[0].h;
    const float screen_aspect = (splat_aspect * 2.0f) * (1.0f - STRIP);
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            float ux = ((float)c + 0.5f) / (float)W, uy = ((float)r + 0.5f) / (float)H;
            float o[4];
            if (uy > 1.0f - STRIP) {
                float sv = (uy - (1.0f - STRIP)) / STRIP;
                float pidx = floorf(ux * PANELS);
                float pu = ux * PANELS - floorf(ux * PANELS);
                if (pu < 0.005f || pu > 0.995f) { o[0] = o[1] = o[2] = 0.3f; o[3] = 1.0f; }
                else if (sv < 0.01f) { o[0] = o[1] = o[2] = 0.3f; o[3] = 1.0f; }
                else {
                    float panel_aspect = screen_aspect / (PANELS * STRIP);
                    float depth_aspect = (float)t[2].w / (float)t[2].h;
                    float edge_aspect = (float)t[3].w / (float)t[3].h;
                    int k, grey;
                    float sa;
                    if (pidx < 0.5f) { k = 1; sa = splat_aspect; grey = 0; }
                    else if (pidx < 1.5f) { k = 2; sa = depth_aspect; grey = 1; }
                    else if (pidx < 2.5f) { k = 3; sa = edge_aspect; grey = 1; }
                    else if (pidx < 3.5f) { k = 4; sa = splat_aspect; grey = 1; }
                    else if (pidx < 4.5f) { k = 5; sa = splat_aspect; grey = 0; }
                    else { k = 6; sa = splat_aspect; grey = 0; }
                    float fu = 0, fv = 0;
                    if (!fit(pu, sv, panel_aspect, sa, &fu, &fv)) { o[0] = o[1] = o[2] = 0.05f; o[3] = 1.0f; }
                    else {
                        at(&t[k], fu, fv, o);
                        if (grey) { o[1] = o[0]; o[2] = o[0]; o[3] = 1.0f; }
                    }
                }
            } else {
                float mv = uy / (1.0f - STRIP);
                if (ux < 0.498f) {
                    float tx = ux * 2.0f, pt[4], ra[4];
                    at(&t[7], tx, mv, pt);
                    at(&t[0], tx, mv, ra);
                    for (int q = 0; q < 4; ++q) o[q] = pt[3] > 0.0f ? pt[q] : ra[q];
                } else if (ux < 0.502f) {
                    o[0] = o[1] = o[2] = o[3] = 1.0f;
                } else {
                    float ru = (ux - 0.5f) * 2.0f, base[4], e4[4], bc[4], cb[4];
                    at(&t[1], ru, mv, base);
                    at(&t[4], ru, mv, e4);
                    at(&t[5], ru, mv, bc);
                    at(&t[6], ru, mv, cb);
                    float e = fminf(fmaxf(e4[0], 0.0f), 1.0f);
                    for (int q = 0; q < 3; ++q) {
                        float x = base[q];
                        if (enabled[1] > 0.5f) x = x * (1.0f - bc[3]) + bc[q];
                        if (enabled[2] > 0.5f) x = x * (1.0f - cb[3]) + cb[q];
                        if (enabled[0] > 0.5f) x = x * (1.0f - e) + 1.0f * e;
                        o[q] = x;
                    }
                    o[3] = 1.0f;
                }
            }
            float* dst = canvas + 4 * ((size_t)r * W + c);
            dst[0] = o[0]; dst[1] = o[1]; dst[2] = o[2]; dst[3] = o[3];
        }
    return 0;
}
