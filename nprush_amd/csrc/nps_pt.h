// nps_pt.h -- path-traced target generator (path_trace.wgsl, bvh.ts,
// path_trace_resolve.wgsl): scene layout and the launch arguments.
//
// Vertices: 12 f32 each [px py pz, nx ny nz, r g b a, u v] (alpha < 0.5 marks a
// mirror); triangles: 3 vertex indices each.  BVH nodes: 8 words,
// [min_x min_y min_z data0][max_x max_y max_z data1] read as two float4;
// internal: data0 = left child, data1 = right child; leaf: data0 = first
// triangle of the reordered list, data1 = count | kPtLeafFlag.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nps_math.h"

namespace nps {

constexpr uint32_t kPtLeafFlag = 0x80000000u;
constexpr uint32_t kPtMaxStack = 64;           // path_trace.wgsl:16
constexpr float kPtPi = 3.14159265358979f;     // path_trace.wgsl:188 (f32)
constexpr int kPtMaxLeafTris = 4;              // bvh.ts:15
constexpr int kPtSahBins = 8;                  // bvh.ts:16

struct PtArgs {
    float ivp[16];              // inverse view-projection, column-major
    uint32_t frame, W, H;
    uint32_t nnodes;            // 0: empty scene (every path escapes)
    const float4* verts;        // 3 float4 per vertex
    const float4* nodes;        // 2 float4 per node
    const uint32_t* tris;       // 3 per triangle, leaf order
    const float4* env;          // ew x eh texels (rgb used)
    int ew, eh;
    float4* accum;              // W x H: radiance sums and sample count
};

struct EnvArgs {
    float ivp[16];
    const float4* env;
    int ew, eh;
    uint32_t W, H;
    float4* out;  // W x H rgba8 values (k/255)
};

// Texel coordinate of the sampler: NaN (a degenerate direction) reads as 0 and
// the range is bounded to +-2^20, so the texel index is always defined.
__device__ __forceinline__ float pt_tex_coord(float x) {
    x = x == x ? x : 0.0f;
    return fminf(fmaxf(x, -1048576.0f), 1048576.0f);
}

// Linear texture filtering (one mip level) as FP32 bilinear filtering of the
// texels at x = u ew - 1/2, y = v eh - 1/2, clamp-to-edge in v and, with
// kRepeatU, repeat in u (the path tracer's environment sampler,
// GpuPathTracePipelineManager.ts:205-208), else clamp-to-edge (the mesh's
// matcap sampler, GpuMeshRenderPipelineManager.ts:254).
template <bool kRepeatU = true>
__device__ __forceinline__ float3 pt_env_bilinear_raw(const float4* __restrict__ env, int ew, int eh, float u,
                                                      float v) {
    const float x = pt_tex_coord(u * (float)ew - 0.5f), y = pt_tex_coord(v * (float)eh - 0.5f);
    const float x0f = floorf(x), y0f = floorf(y);
    const float fx = x - x0f, fy = y - y0f;
    const int ix = (int)x0f, iy = (int)y0f;
    const int x0 = kRepeatU ? ((ix % ew) + ew) % ew : min(max(ix, 0), ew - 1);
    const int x1 = kRepeatU ? (((ix + 1) % ew) + ew) % ew : min(max(ix + 1, 0), ew - 1);
    const int y0 = min(max(iy, 0), eh - 1), y1 = min(max(iy + 1, 0), eh - 1);
    const float4 t00 = env[(size_t)y0 * ew + x0], t10 = env[(size_t)y0 * ew + x1];
    const float4 t01 = env[(size_t)y1 * ew + x0], t11 = env[(size_t)y1 * ew + x1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    float3 r;
    r.x = (t00.x * gx + t10.x * fx) * gy + (t01.x * gx + t11.x * fx) * fy;
    r.y = (t00.y * gx + t10.y * fx) * gy + (t01.y * gx + t11.y * fx) * fy;
    r.z = (t00.z * gx + t10.z * fx) * gy + (t01.z * gx + t11.z * fx) * fy;
    return r;
}

__device__ __forceinline__ float pt_reinhard(float c) { return c / (c + 1.0f); }

// rgba8unorm store: clamp to [0, 1], round to nearest (ties to even), as k/255.
__device__ __forceinline__ float pt_unorm8(float x) {
    if (!(x > 0.0f)) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return rintf(x * 255.0f) / 255.0f;
}

// Mesh raster target (mesh.wgsl vert/frag, GpuMeshRenderPipelineManager's
// pipeline: back faces culled, depth test "less", over the envmap background),
// restated per pixel centre as the closest front-facing triangle along the
// centre's ray, with the varyings blended by its barycentric weights (the
// perspective-correct interpolation weights at that point).
struct RasterArgs {
    float vp[16], view[16], ivp[16];
    float shading;              // < 0.5: normal colours, else matcap x vertex colour
    uint32_t W, H, nnodes;
    const float4* verts;
    const float4* nodes;
    const uint32_t* tris;
    const float4* env;          // the matcap / environment texture
    int ew, eh;
    float4* color;              // W x H rgba8 values (k/255)
    float4* depth;              // W x H: (d, d, d, 1), d = 1 - 0.1 / max(w, 0.1) blended, rgba8
};

void launch_pt_trace(hipStream_t st, const PtArgs& a);
void launch_mesh_raster(hipStream_t st, const RasterArgs& a);
void launch_pt_resolve(hipStream_t st, const float4* accum, uint32_t npix, float4* out);
void launch_env_background(hipStream_t st, const float* ivp, const float4* env, int ew, int eh, uint32_t W,
                           uint32_t H, float4* out);

}  // namespace nps
