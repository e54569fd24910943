// npp_api.cpp -- C ABI of the path-traced target generator (npp_*), the
// counterpart of GpuPathTracePipelineManager (src/lib/gpu/GpuPathTracePipelineManager.ts),
// and its host-side SAH BVH builder (src/lib/gpu/bvh.ts).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <string>
#include <vector>

#include "../../include/nprush_hip.h"
#include "nps_host.h"
#include "nps_kernels.h"
#include "nps_pt.h"

using namespace nps;

namespace {

// ------------------------------------------------------------------ BVH build
// The builder of bvh.ts in double precision (its arithmetic is JavaScript's):
// per triangle a box padded by 1e-4 and its centre; per node the union box
// (stored as float32), then a leaf at <= 4 triangles, else the best of 8 SAH
// bins per axis (cost (SA_l n_l + SA_r n_r) / SA_parent, first minimum over
// axis x, y, z and split from the right), a split at the centroid bin, and a
// split at the middle of the list when the centroids do not spread, the best
// cost is not below the triangle count, or one side would be empty.  Nodes are
// numbered in depth-first order (a node before its children, the left subtree
// before the right); leaves append their triangles to the reordered list.
struct Box {
    double mn[3], mx[3];
};
struct BTri {
    uint32_t i0, i1, i2;
    Box b;
    double c[3];
};

inline Box empty_box() {
    const double inf = std::numeric_limits<double>::infinity();
    return Box{{inf, inf, inf}, {-inf, -inf, -inf}};
}
// Math.min / Math.max of two numbers: NaN-propagating, -0 below +0.
inline double jmin(double a, double b) {
    if (a != a || b != b) return NAN;
    if (a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}
inline double jmax(double a, double b) {
    if (a != a || b != b) return NAN;
    if (a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}
inline Box box_union(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = jmin(a.mn[k], b.mn[k]);
        r.mx[k] = jmax(a.mx[k], b.mx[k]);
    }
    return r;
}
inline double box_area(const Box& a) {
    const double dx = a.mx[0] - a.mn[0], dy = a.mx[1] - a.mn[1], dz = a.mx[2] - a.mn[2];
    return 2 * (dx * dy + dy * dz + dz * dx);
}

struct BvhBuilder {
    std::vector<uint32_t> nodes;  // 8 words per node
    std::vector<uint32_t> tris;   // 3 per triangle, leaf order

    static uint32_t f32_bits(double v) {
        const float f = (float)v;
        uint32_t u;
        memcpy(&u, &f, 4);
        return u;
    }
    uint32_t alloc(const Box& b) {
        const uint32_t idx = (uint32_t)(nodes.size() / 8);
        nodes.resize(nodes.size() + 8, 0u);
        uint32_t* n = &nodes[8 * (size_t)idx];
        n[0] = f32_bits(b.mn[0]);
        n[1] = f32_bits(b.mn[1]);
        n[2] = f32_bits(b.mn[2]);
        n[4] = f32_bits(b.mx[0]);
        n[5] = f32_bits(b.mx[1]);
        n[6] = f32_bits(b.mx[2]);
        return idx;
    }
    static int bin_of(double c, double cmin, double cmax) {
        const double b = floor(((c - cmin) / (cmax - cmin)) * kPtSahBins);
        return (int)std::min(b, (double)(kPtSahBins - 1));
    }
    uint32_t build(const std::vector<BTri>& t) {
        Box bounds = empty_box(), cb = empty_box();
        for (const BTri& x : t) {
            bounds = box_union(bounds, x.b);
            cb = box_union(cb, Box{{x.c[0], x.c[1], x.c[2]}, {x.c[0], x.c[1], x.c[2]}});
        }
        const uint32_t idx = alloc(bounds);
        if ((int)t.size() <= kPtMaxLeafTris) {
            const uint32_t start = (uint32_t)(tris.size() / 3);
            for (const BTri& x : t) {
                tris.push_back(x.i0);
                tris.push_back(x.i1);
                tris.push_back(x.i2);
            }
            nodes[8 * (size_t)idx + 3] = start;
            nodes[8 * (size_t)idx + 7] = (uint32_t)t.size() | kPtLeafFlag;
            return idx;
        }
        double best_cost = std::numeric_limits<double>::infinity();
        int best_axis = 0, best_split = 0;
        const double parent_sa = box_area(bounds);
        for (int axis = 0; axis < 3; ++axis) {
            const double cmin = cb.mn[axis], cmax = cb.mx[axis];
            if (cmax - cmin < 1e-8) continue;
            Box bb[kPtSahBins];
            int bc[kPtSahBins] = {};
            for (int i = 0; i < kPtSahBins; ++i) bb[i] = empty_box();
            for (const BTri& x : t) {
                const int b = bin_of(x.c[axis], cmin, cmax);
                bb[b] = box_union(bb[b], x.b);
                bc[b]++;
            }
            Box lbox[kPtSahBins - 1];
            int lcnt[kPtSahBins - 1];
            Box lb = empty_box();
            int lc = 0;
            for (int i = 0; i < kPtSahBins - 1; ++i) {
                lb = box_union(lb, bb[i]);
                lc += bc[i];
                lbox[i] = lb;
                lcnt[i] = lc;
            }
            Box rb = empty_box();
            int rc = 0;
            for (int i = kPtSahBins - 2; i >= 0; --i) {
                rb = box_union(rb, bb[i + 1]);
                rc += bc[i + 1];
                const double cost = (box_area(lbox[i]) * lcnt[i] + box_area(rb) * rc) / parent_sa;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = i;
                }
            }
        }
        const double cmin = cb.mn[best_axis], cmax = cb.mx[best_axis];
        std::vector<BTri> left, right;
        const size_t mid = t.size() / 2;
        if (cmax - cmin < 1e-8 || best_cost >= (double)t.size()) {
            left.assign(t.begin(), t.begin() + (ptrdiff_t)mid);
            right.assign(t.begin() + (ptrdiff_t)mid, t.end());
        } else {
            for (const BTri& x : t) (bin_of(x.c[best_axis], cmin, cmax) <= best_split ? left : right).push_back(x);
            if (left.empty() || right.empty()) {
                left.assign(t.begin(), t.begin() + (ptrdiff_t)mid);
                right.assign(t.begin() + (ptrdiff_t)mid, t.end());
            }
        }
        const uint32_t li = build(left);
        const uint32_t ri = build(right);
        nodes[8 * (size_t)idx + 3] = li;
        nodes[8 * (size_t)idx + 7] = ri;
        return idx;
    }
};

void build_bvh(const float* verts, const uint32_t* idx, uint32_t ntris, std::vector<uint32_t>& nodes,
               std::vector<uint32_t>& tris) {
    std::vector<BTri> t(ntris);
    const double eps = 1e-4;
    for (uint32_t k = 0; k < ntris; ++k) {
        BTri& x = t[k];
        x.i0 = idx[3 * (size_t)k];
        x.i1 = idx[3 * (size_t)k + 1];
        x.i2 = idx[3 * (size_t)k + 2];
        for (int a = 0; a < 3; ++a) {
            const double p = verts[12 * (size_t)x.i0 + a], q = verts[12 * (size_t)x.i1 + a],
                         r = verts[12 * (size_t)x.i2 + a];
            x.b.mn[a] = jmin(jmin(p, q), r) - eps;
            x.b.mx[a] = jmax(jmax(p, q), r) + eps;
            x.c[a] = (x.b.mn[a] + x.b.mx[a]) * 0.5;
        }
    }
    BvhBuilder b;
    b.build(t);
    nodes.swap(b.nodes);
    tris.swap(b.tris);
}

}  // namespace

struct npp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf verts, nodes, tris, env, accum, out, bg, rc, rd;
    std::vector<uint32_t> host_nodes, host_tris;
    bool have_meshes = false, have_env = false;
    int ew = 0, eh = 0;
    uint32_t W = 0, H = 0, frame = 0;
    float ivp[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

namespace {

int perr(npp_ctx* p, int code, const std::string& msg) {
    p->err = msg;
    return code;
}

#define NPP_HIP(p, expr)                                                                                     \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return perr((p), NPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

int npp_create(int32_t device, npp_handle* out) {
    if (!out) return NPS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return NPS_ERR_HIP;
    npp_ctx* p = new npp_ctx();
    p->device = device;
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return NPS_ERR_HIP;
    }
    *out = p;
    return NPS_OK;
}

void npp_destroy(npp_handle p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    (void)hipStreamSynchronize(p->stream);
    free_buf(p->verts); free_buf(p->nodes); free_buf(p->tris); free_buf(p->env);
    free_buf(p->accum); free_buf(p->out); free_buf(p->bg); free_buf(p->rc); free_buf(p->rd);
    (void)hipStreamDestroy(p->stream);
    delete p;
}

const char* npp_last_error(npp_handle p) { return p ? p->err.c_str() : "invalid handle"; }

int npp_set_env(npp_handle p, const float* rgba, int32_t ew, int32_t eh) {
    if (!p) return NPS_ERR_ARG;
    if (!rgba || ew <= 0 || eh <= 0) return perr(p, NPS_ERR_ARG, "environment: null texels or empty size");
    NPP_HIP(p, hipSetDevice(p->device));
    const size_t bytes = 16 * (size_t)ew * eh;
    int rc;
    if ((rc = ensure_dev(p->env, bytes, p->stream, p->err)) != NPS_OK) return rc;
    NPP_HIP(p, hipMemcpyAsync(p->env.p, rgba, bytes, hipMemcpyHostToDevice, p->stream));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    p->ew = ew;
    p->eh = eh;
    p->have_env = true;
    return NPS_OK;
}

int npp_set_meshes(npp_handle p, const float* vertices, int32_t nverts, const uint32_t* indices, int32_t ntris) {
    if (!p) return NPS_ERR_ARG;
    if (nverts < 0 || ntris < 0 || (nverts > 0 && !vertices) || (ntris > 0 && !indices))
        return perr(p, NPS_ERR_ARG, "meshes: negative count or null array");
    for (size_t k = 0; k < 3 * (size_t)ntris; ++k)
        if (indices[k] >= (uint32_t)nverts) return perr(p, NPS_ERR_ARG, "meshes: vertex index out of range");
    NPP_HIP(p, hipSetDevice(p->device));
    build_bvh(vertices, indices, (uint32_t)ntris, p->host_nodes, p->host_tris);
    int rc;
    const size_t vb = 48 * (size_t)std::max(nverts, 1), nb = 4 * p->host_nodes.size(),
                 tb = 4 * std::max<size_t>(p->host_tris.size(), 1);
    if ((rc = ensure_dev(p->verts, vb, p->stream, p->err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(p->nodes, nb, p->stream, p->err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(p->tris, tb, p->stream, p->err)) != NPS_OK) return rc;
    if (nverts > 0) NPP_HIP(p, hipMemcpyAsync(p->verts.p, vertices, 48 * (size_t)nverts, hipMemcpyHostToDevice, p->stream));
    NPP_HIP(p, hipMemcpyAsync(p->nodes.p, p->host_nodes.data(), nb, hipMemcpyHostToDevice, p->stream));
    if (!p->host_tris.empty())
        NPP_HIP(p, hipMemcpyAsync(p->tris.p, p->host_tris.data(), 4 * p->host_tris.size(), hipMemcpyHostToDevice,
                                  p->stream));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    p->have_meshes = true;
    return NPS_OK;
}

int npp_set_output_size(npp_handle p, int32_t w, int32_t hgt) {
    if (!p) return NPS_ERR_ARG;
    if (w <= 0 || hgt <= 0) return perr(p, NPS_ERR_ARG, "output size must be positive");
    if ((uint32_t)w == p->W && (uint32_t)hgt == p->H) return NPS_OK;
    NPP_HIP(p, hipSetDevice(p->device));
    const size_t px = (size_t)w * hgt;
    int rc;
    if ((rc = ensure_dev(p->accum, 16 * px, p->stream, p->err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(p->out, 16 * px, p->stream, p->err)) != NPS_OK) return rc;
    launch_fill_u32(p->stream, (uint32_t*)p->accum.p, 0u, 4 * (uint64_t)px);
    launch_fill_u32(p->stream, (uint32_t*)p->out.p, 0u, 4 * (uint64_t)px);
    NPP_HIP(p, hipGetLastError());
    p->W = (uint32_t)w;
    p->H = (uint32_t)hgt;
    p->frame = 0;
    return NPS_OK;
}

int npp_reset(npp_handle p) {
    if (!p) return NPS_ERR_ARG;
    if (!p->W) return perr(p, NPS_ERR_STATE, "no output size (npp_set_output_size)");
    NPP_HIP(p, hipSetDevice(p->device));
    launch_fill_u32(p->stream, (uint32_t*)p->accum.p, 0u, 4 * (uint64_t)p->W * p->H);
    NPP_HIP(p, hipGetLastError());
    p->frame = 0;
    return NPS_OK;
}

int npp_write_inv_vp(npp_handle p, const float inv_vp[16]) {
    if (!p || !inv_vp) return NPS_ERR_ARG;
    memcpy(p->ivp, inv_vp, sizeof(p->ivp));
    return NPS_OK;
}

int npp_dispatch(npp_handle p) {
    if (!p) return NPS_ERR_ARG;
    if (!p->have_meshes) return perr(p, NPS_ERR_STATE, "no meshes (npp_set_meshes)");
    if (!p->have_env) return perr(p, NPS_ERR_STATE, "no environment (npp_set_env)");
    if (!p->W) return perr(p, NPS_ERR_STATE, "no output size (npp_set_output_size)");
    NPP_HIP(p, hipSetDevice(p->device));
    PtArgs a;
    memcpy(a.ivp, p->ivp, sizeof(a.ivp));
    a.frame = p->frame;
    a.W = p->W;
    a.H = p->H;
    a.nnodes = (uint32_t)(p->host_nodes.size() / 8);
    a.verts = (const float4*)p->verts.p;
    a.nodes = (const float4*)p->nodes.p;
    a.tris = (const uint32_t*)p->tris.p;
    a.env = (const float4*)p->env.p;
    a.ew = p->ew;
    a.eh = p->eh;
    a.accum = (float4*)p->accum.p;
    launch_pt_trace(p->stream, a);
    launch_pt_resolve(p->stream, (const float4*)p->accum.p, p->W * p->H, (float4*)p->out.p);
    NPP_HIP(p, hipGetLastError());
    p->frame++;
    return NPS_OK;
}

int npp_read_output(npp_handle p, float* rgba) {
    if (!p || !rgba) return NPS_ERR_ARG;
    if (!p->W) return perr(p, NPS_ERR_STATE, "no output size (npp_set_output_size)");
    NPP_HIP(p, hipSetDevice(p->device));
    NPP_HIP(p, hipMemcpyAsync(rgba, p->out.p, 16 * (size_t)p->W * p->H, hipMemcpyDeviceToHost, p->stream));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    return NPS_OK;
}

int npp_read_accum(npp_handle p, float* accum) {
    if (!p || !accum) return NPS_ERR_ARG;
    if (!p->W) return perr(p, NPS_ERR_STATE, "no output size (npp_set_output_size)");
    NPP_HIP(p, hipSetDevice(p->device));
    NPP_HIP(p, hipMemcpyAsync(accum, p->accum.p, 16 * (size_t)p->W * p->H, hipMemcpyDeviceToHost, p->stream));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    return NPS_OK;
}

int npp_output_dev(npp_handle p, const float** rgba_dev) {
    if (!p || !rgba_dev) return NPS_ERR_ARG;
    if (!p->W) return perr(p, NPS_ERR_STATE, "no output size (npp_set_output_size)");
    NPP_HIP(p, hipSetDevice(p->device));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    *rgba_dev = (const float*)p->out.p;
    return NPS_OK;
}

int npp_read_bvh(npp_handle p, uint32_t* nodes, uint32_t* tris, int32_t counts[2]) {
    if (!p || !counts) return NPS_ERR_ARG;
    if (!p->have_meshes) return perr(p, NPS_ERR_STATE, "no meshes (npp_set_meshes)");
    counts[0] = (int32_t)(p->host_nodes.size() / 8);
    counts[1] = (int32_t)(p->host_tris.size() / 3);
    if (nodes) memcpy(nodes, p->host_nodes.data(), 4 * p->host_nodes.size());
    if (tris) memcpy(tris, p->host_tris.data(), 4 * p->host_tris.size());
    return NPS_OK;
}

int npp_env_background(npp_handle p, const float inv_vp[16], float* rgba, int32_t w, int32_t hgt) {
    if (!p || !inv_vp || !rgba || w <= 0 || hgt <= 0) return NPS_ERR_ARG;
    if (!p->have_env) return perr(p, NPS_ERR_STATE, "no environment (npp_set_env)");
    NPP_HIP(p, hipSetDevice(p->device));
    const size_t bytes = 16 * (size_t)w * hgt;
    int rc;
    if ((rc = ensure_dev(p->bg, bytes, p->stream, p->err)) != NPS_OK) return rc;
    launch_env_background(p->stream, inv_vp, (const float4*)p->env.p, p->ew, p->eh, (uint32_t)w, (uint32_t)hgt,
                          (float4*)p->bg.p);
    NPP_HIP(p, hipGetLastError());
    NPP_HIP(p, hipMemcpyAsync(rgba, p->bg.p, bytes, hipMemcpyDeviceToHost, p->stream));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    return NPS_OK;
}

int npp_raster(npp_handle p, const float vp[16], const float view[16], const float inv_vp[16], float shading_mode,
               float* color, float* depth, int32_t w, int32_t hgt) {
    if (!p || !vp || !view || !inv_vp || !color || !depth || w <= 0 || hgt <= 0) return NPS_ERR_ARG;
    if (!p->have_meshes) return perr(p, NPS_ERR_STATE, "no meshes (npp_set_meshes)");
    if (!p->have_env) return perr(p, NPS_ERR_STATE, "no environment / matcap texture (npp_set_env)");
    NPP_HIP(p, hipSetDevice(p->device));
    const size_t bytes = 16 * (size_t)w * hgt;
    int rc;
    if ((rc = ensure_dev(p->rc, bytes, p->stream, p->err)) != NPS_OK) return rc;
    if ((rc = ensure_dev(p->rd, bytes, p->stream, p->err)) != NPS_OK) return rc;
    RasterArgs a;
    memcpy(a.vp, vp, sizeof(a.vp));
    memcpy(a.view, view, sizeof(a.view));
    memcpy(a.ivp, inv_vp, sizeof(a.ivp));
    a.shading = shading_mode;
    a.W = (uint32_t)w;
    a.H = (uint32_t)hgt;
    a.nnodes = (uint32_t)(p->host_nodes.size() / 8);
    a.verts = (const float4*)p->verts.p;
    a.nodes = (const float4*)p->nodes.p;
    a.tris = (const uint32_t*)p->tris.p;
    a.env = (const float4*)p->env.p;
    a.ew = p->ew;
    a.eh = p->eh;
    a.color = (float4*)p->rc.p;
    a.depth = (float4*)p->rd.p;
    launch_mesh_raster(p->stream, a);
    NPP_HIP(p, hipGetLastError());
    NPP_HIP(p, hipMemcpyAsync(color, p->rc.p, bytes, hipMemcpyDeviceToHost, p->stream));
    NPP_HIP(p, hipMemcpyAsync(depth, p->rd.p, bytes, hipMemcpyDeviceToHost, p->stream));
    NPP_HIP(p, hipStreamSynchronize(p->stream));
    return NPS_OK;
}

}  // extern "C"
