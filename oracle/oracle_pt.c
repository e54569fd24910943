/* oracle_pt.c -- CPU restatement of the path-traced target generator, test
 * infrastructure only (the checker of the GPU path, never the thing shipped):
 *   ora_pt_bvh     : the SAH BVH builder of src/lib/gpu/bvh.ts:56-223 (double
 *                    precision, as its JavaScript arithmetic);
 *   ora_pt_trace   : one frame of path_trace.wgsl:214-268 (closest hit by the
 *                    BVH walk of :146-185, watertight triangle test :60-129,
 *                    slab test :132-142, PCG :36-45, environment :190-196,
 *                    cosine hemisphere :200-210);
 *   ora_pt_resolve : path_trace_resolve.wgsl:21-30;
 *   ora_env_background : envmap.wgsl:38-60 per pixel centre.
 * FP32 with -ffp-contract=off, every expression in the shader's order;
 * dot(a,b) = (ax bx + ay by) + az bz, M v = ((c0 x + c1 y) + c2 z) + c3 w,
 * normalize(v) = v / sqrt(dot(v, v)), WGSL min/max (a NaN operand loses);
 * atan2/asin/sin/cos are ref_math.h's.  The environment sampler is FP32
 * bilinear filtering at x = u W - 1/2, y = v H - 1/2 (repeat u, clamp v; a NaN
 * coordinate reads as 0 and coordinates are bounded to +-2^20): the hardware
 * filter weights are implementation-defined (parity unpinned there). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ref_math.h"

#define PT_LEAF 0x80000000u
#define PT_STACK 64
#define PT_LEAF_TRIS 4
#define PT_BINS 8
static const float PT_PI = 3.14159265358979f;

/* ------------------------------------------------------------ BVH (bvh.ts) */
typedef struct { double lo[3], hi[3]; } pbox;
typedef struct { uint32_t i[3]; pbox b; double c[3]; } ptri;

static pbox pbox_empty(void) {
    pbox r;
    for (int k = 0; k < 3; ++k) { r.lo[k] = INFINITY; r.hi[k] = -INFINITY; }
    return r;
}
/* JavaScript's Math.min / Math.max: NaN wins, -0 counts below +0 */
static double jsmin(double a, double b) {
    if (a != a || b != b) return NAN;
    if (a == 0.0 && b == 0.0) return signbit(b) ? b : a;
    return b < a ? b : a;
}
static double jsmax(double a, double b) {
    if (a != a || b != b) return NAN;
    if (a == 0.0 && b == 0.0) return signbit(b) ? a : b;
    return b > a ? b : a;
}
static pbox pbox_join(pbox a, pbox b) {
    pbox r;
    for (int k = 0; k < 3; ++k) { r.lo[k] = jsmin(a.lo[k], b.lo[k]); r.hi[k] = jsmax(a.hi[k], b.hi[k]); }
    return r;
}
static double pbox_sa(pbox a) {
    double dx = a.hi[0] - a.lo[0], dy = a.hi[1] - a.lo[1], dz = a.hi[2] - a.lo[2];
    return 2 * (dx * dy + dy * dz + dz * dx);
}
static uint32_t f32bits(double v) { float f = (float)v; uint32_t u; memcpy(&u, &f, 4); return u; }

typedef struct { uint32_t* nodes; int nn, cap; uint32_t* tris; int nt; const ptri* T; } pbuild;

static int pbin(double c, double lo, double hi) {
    double b = floor(((c - lo) / (hi - lo)) * PT_BINS);
    if (b > PT_BINS - 1) b = PT_BINS - 1;
    return (int)b;
}

/* idx: n triangle numbers in list order; returns the node's number */
static int pbuild_node(pbuild* B, const int* idx, int n) {
    pbox bounds = pbox_empty(), cb = pbox_empty();
    for (int k = 0; k < n; ++k) {
        const ptri* t = &B->T[idx[k]];
        pbox cc;
        for (int a = 0; a < 3; ++a) { cc.lo[a] = t->c[a]; cc.hi[a] = t->c[a]; }
        bounds = pbox_join(bounds, t->b);
        cb = pbox_join(cb, cc);
    }
    if (B->nn == B->cap) {
        B->cap = B->cap ? 2 * B->cap : 64;
        B->nodes = (uint32_t*)realloc(B->nodes, (size_t)B->cap * 32);
    }
    int me = B->nn++;
    uint32_t* nd = B->nodes + 8 * (size_t)me;
    for (int a = 0; a < 3; ++a) { nd[a] = f32bits(bounds.lo[a]); nd[4 + a] = f32bits(bounds.hi[a]); }
    if (n <= PT_LEAF_TRIS) {
        nd[3] = (uint32_t)B->nt;
        nd[7] = (uint32_t)n | PT_LEAF;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < 3; ++j) B->tris[3 * (size_t)B->nt + 3 * (size_t)k + j] = B->T[idx[k]].i[j];
        B->nt += n;
        return me;
    }
    double best = INFINITY, psa = pbox_sa(bounds);
    int baxis = 0, bsplit = 0;
    for (int a = 0; a < 3; ++a) {
        double lo = cb.lo[a], hi = cb.hi[a];
        if (hi - lo < 1e-8) continue;
        pbox bb[PT_BINS];
        int bc[PT_BINS];
        for (int i = 0; i < PT_BINS; ++i) { bb[i] = pbox_empty(); bc[i] = 0; }
        for (int k = 0; k < n; ++k) {
            int b = pbin(B->T[idx[k]].c[a], lo, hi);
            bb[b] = pbox_join(bb[b], B->T[idx[k]].b);
            bc[b] += 1;
        }
        pbox lbs[PT_BINS - 1], acc = pbox_empty();
        int lcs[PT_BINS - 1], lc = 0;
        for (int i = 0; i < PT_BINS - 1; ++i) { acc = pbox_join(acc, bb[i]); lc += bc[i]; lbs[i] = acc; lcs[i] = lc; }
        pbox rb = pbox_empty();
        int rc = 0;
        for (int i = PT_BINS - 2; i >= 0; --i) {
            rb = pbox_join(rb, bb[i + 1]);
            rc += bc[i + 1];
            double cost = (pbox_sa(lbs[i]) * lcs[i] + pbox_sa(rb) * rc) / psa;
            if (cost < best) { best = cost; baxis = a; bsplit = i; }
        }
    }
    double lo = cb.lo[baxis], hi = cb.hi[baxis];
    int* L = (int*)malloc(sizeof(int) * (size_t)n);
    int* R = (int*)malloc(sizeof(int) * (size_t)n);
    int nl = 0, nr = 0;
    if (hi - lo < 1e-8 || best >= (double)n) {
        for (int k = 0; k < n; ++k) { if (k < n / 2) L[nl++] = idx[k]; else R[nr++] = idx[k]; }
    } else {
        for (int k = 0; k < n; ++k) {
            if (pbin(B->T[idx[k]].c[baxis], lo, hi) <= bsplit) L[nl++] = idx[k];
            else R[nr++] = idx[k];
        }
        if (nl == 0 || nr == 0) {
            nl = nr = 0;
            for (int k = 0; k < n; ++k) { if (k < n / 2) L[nl++] = idx[k]; else R[nr++] = idx[k]; }
        }
    }
    int li = pbuild_node(B, L, nl);
    int ri = pbuild_node(B, R, nr);
    free(L);
    free(R);
    B->nodes[8 * (size_t)me + 3] = (uint32_t)li;
    B->nodes[8 * (size_t)me + 7] = (uint32_t)ri;
    return me;
}

/* nodes_out: capacity 8 * (2 ntris + 1) words (a binary tree with at most
 * ntris leaves... every leaf holds >= 1 triangle unless the scene is empty);
 * tris_out: 3 ntris.  Returns the node count. */
int ora_pt_bvh(const float* verts, int nverts, const uint32_t* indices, int ntris, uint32_t* nodes_out,
               uint32_t* tris_out) {
    (void)nverts;
    ptri* T = (ptri*)malloc(sizeof(ptri) * (size_t)(ntris > 0 ? ntris : 1));
    for (int k = 0; k < ntris; ++k) {
        for (int j = 0; j < 3; ++j) T[k].i[j] = indices[3 * (size_t)k + j];
        for (int a = 0; a < 3; ++a) {
            double p = verts[12 * (size_t)T[k].i[0] + a], q = verts[12 * (size_t)T[k].i[1] + a],
                   r = verts[12 * (size_t)T[k].i[2] + a];
            T[k].b.lo[a] = jsmin(jsmin(p, q), r) - 1e-4;
            T[k].b.hi[a] = jsmax(jsmax(p, q), r) + 1e-4;
            T[k].c[a] = (T[k].b.lo[a] + T[k].b.hi[a]) * 0.5;
        }
    }
    int* idx = (int*)malloc(sizeof(int) * (size_t)(ntris > 0 ? ntris : 1));
    for (int k = 0; k < ntris; ++k) idx[k] = k;
    pbuild B = {0};
    B.T = T;
    B.tris = tris_out;
    pbuild_node(&B, idx, ntris);
    memcpy(nodes_out, B.nodes, (size_t)B.nn * 32);
    int nn = B.nn;
    free(B.nodes);
    free(idx);
    free(T);
    return nn;
}

/* ------------------------------------------------------------ path tracer */
typedef struct { float x, y, z; } f3;
static f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f3 vadd(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static f3 vsub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static f3 vscale(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static float vdot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static f3 vcross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static f3 vnorm(f3 v) { float l = sqrtf(vdot(v, v)); return mk(v.x / l, v.y / l, v.z / l); }
static float vget(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
static float wmax(float a, float b) { if (b != b) return a; if (a != a) return b; return a > b ? a : b; }
static float wmin(float a, float b) { if (b != b) return a; if (a != a) return b; return a < b ? a : b; }

static f3 unproj(const float* m, float x, float y, float z) {
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = ((m[i] * x + m[4 + i] * y) + m[8 + i] * z) + m[12 + i] * 1.0f;
    return mk(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}

uint32_t ora_pt_pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static void prand2(uint32_t* seed, float* a, float* b) {
    *seed = ora_pt_pcg(*seed);
    uint32_t x = *seed;
    *seed = ora_pt_pcg(*seed);
    uint32_t y = *seed;
    *a = (float)x / 4294967296.0f;
    *b = (float)y / 4294967296.0f;
}

typedef struct { int hit; float t; f3 n; float c[4]; float bu, bv, bw; uint32_t tri[3]; } phit;

/* cull: skip a triangle whose ccw normal (p1 - p0) x (p2 - p0) does not face
 * the ray (the raster's back-face cull, GpuMeshRenderPipelineManager.ts:99-103) */
static void ptri_test(f3 ro, f3 rd, const float* V, uint32_t i0, uint32_t i1, uint32_t i2, float tmax, phit* best,
                      int cull) {
    const float* q[3] = { V + 12 * (size_t)i0, V + 12 * (size_t)i1, V + 12 * (size_t)i2 };
    if (cull) {
        f3 e1 = vsub(mk(q[1][0], q[1][1], q[1][2]), mk(q[0][0], q[0][1], q[0][2]));
        f3 e2 = vsub(mk(q[2][0], q[2][1], q[2][2]), mk(q[0][0], q[0][1], q[0][2]));
        if (!(vdot(vcross(e1, e2), rd) < 0.0f)) return;
    }
    float ax = fabsf(rd.x), ay = fabsf(rd.y), az = fabsf(rd.z);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > wmax(ax, ay)) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (vget(rd, kz) < 0.0f) { int t = kx; kx = ky; ky = t; }
    float Sz = 1.0f / vget(rd, kz);
    float Sx = vget(rd, kx) * Sz, Sy = vget(rd, ky) * Sz;
    f3 A = vsub(mk(q[0][0], q[0][1], q[0][2]), ro);
    f3 B = vsub(mk(q[1][0], q[1][1], q[1][2]), ro);
    f3 C = vsub(mk(q[2][0], q[2][1], q[2][2]), ro);
    float Ax = vget(A, kx) - Sx * vget(A, kz), Ay = vget(A, ky) - Sy * vget(A, kz);
    float Bx = vget(B, kx) - Sx * vget(B, kz), By = vget(B, ky) - Sy * vget(B, kz);
    float Cx = vget(C, kx) - Sx * vget(C, kz), Cy = vget(C, ky) - Sy * vget(C, kz);
    float U = Cx * By - Cy * Bx, Vv = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return;
    float det = (U + Vv) + W;
    if (det == 0.0f) return;
    float Az = Sz * vget(A, kz), Bz = Sz * vget(B, kz), Cz = Sz * vget(C, kz);
    float T = (U * Az + Vv * Bz) + W * Cz;
    float inv = 1.0f / det;
    float t = T * inv;
    if (t < 1e-4f || t >= tmax) return;
    float u = U * inv, v = Vv * inv, w = W * inv;
    f3 n0 = mk(q[0][3], q[0][4], q[0][5]), n1 = mk(q[1][3], q[1][4], q[1][5]), n2 = mk(q[2][3], q[2][4], q[2][5]);
    best->hit = 1;
    best->t = t;
    best->n = vnorm(vadd(vadd(vscale(n0, u), vscale(n1, v)), vscale(n2, w)));
    for (int k = 0; k < 4; ++k) best->c[k] = (u * q[0][6 + k] + v * q[1][6 + k]) + w * q[2][6 + k];
    best->bu = u; best->bv = v; best->bw = w;
    best->tri[0] = i0; best->tri[1] = i1; best->tri[2] = i2;
}

static int pslab(f3 ro, f3 inv, const uint32_t* nd, float tmax) {
    float mn[3], mx[3];
    memcpy(mn, nd, 12);
    memcpy(mx, nd + 4, 12);
    float t0x = (mn[0] - ro.x) * inv.x, t0y = (mn[1] - ro.y) * inv.y, t0z = (mn[2] - ro.z) * inv.z;
    float t1x = (mx[0] - ro.x) * inv.x, t1y = (mx[1] - ro.y) * inv.y, t1z = (mx[2] - ro.z) * inv.z;
    float tmin = wmax(wmax(wmin(t0x, t1x), wmin(t0y, t1y)), wmin(t0z, t1z));
    float tm = wmin(wmin(wmax(t0x, t1x), wmax(t0y, t1y)), wmax(t0z, t1z));
    return tm >= 0.0f && tm >= tmin && tmin < tmax;
}

static phit pscene(f3 ro, f3 rd, const uint32_t* nodes, int nn, const uint32_t* tris, const float* V, int cull) {
    phit best;
    memset(&best, 0, sizeof(best));
    best.t = 1e30f;
    if (nn <= 0) return best;
    f3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    uint32_t st[PT_STACK];
    uint32_t top = 0;
    st[top++] = 0;
    while (top > 0) {
        uint32_t node = st[--top];
        const uint32_t* nd = nodes + 8 * (size_t)node;
        if (!pslab(ro, inv, nd, best.t)) continue;
        if (nd[7] & PT_LEAF) {
            uint32_t cnt = nd[7] & ~PT_LEAF;
            for (uint32_t k = 0; k < cnt; ++k) {
                size_t b = 3 * ((size_t)nd[3] + k);
                ptri_test(ro, rd, V, tris[b], tris[b + 1], tris[b + 2], best.t, &best, cull);
            }
        } else {
            if (top + 1 < PT_STACK) st[top++] = nd[7];
            if (top < PT_STACK) st[top++] = nd[3];
        }
    }
    return best;
}

/* texel coordinate: NaN reads as 0, bounded to +-2^20 (always a defined index) */
static float ptexc(float x) {
    x = x == x ? x : 0.0f;
    return x < -1048576.0f ? -1048576.0f : (x > 1048576.0f ? 1048576.0f : x);
}

static f3 penv_bilinear_m(const float* env, int ew, int eh, float u, float v, int repeat_u) {
    float x = ptexc(u * (float)ew - 0.5f), y = ptexc(v * (float)eh - 0.5f);
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    int ix = (int)x0f, iy = (int)y0f;
    int x0, x1;
    if (repeat_u) {
        x0 = ((ix % ew) + ew) % ew;
        x1 = (((ix + 1) % ew) + ew) % ew;
    } else {
        x0 = ix < 0 ? 0 : (ix > ew - 1 ? ew - 1 : ix);
        x1 = ix + 1 < 0 ? 0 : (ix + 1 > ew - 1 ? ew - 1 : ix + 1);
    }
    int y0 = iy < 0 ? 0 : (iy > eh - 1 ? eh - 1 : iy);
    int y1 = iy + 1 < 0 ? 0 : (iy + 1 > eh - 1 ? eh - 1 : iy + 1);
    const float* t00 = env + 4 * ((size_t)y0 * ew + x0);
    const float* t10 = env + 4 * ((size_t)y0 * ew + x1);
    const float* t01 = env + 4 * ((size_t)y1 * ew + x0);
    const float* t11 = env + 4 * ((size_t)y1 * ew + x1);
    float gx = 1.0f - fx, gy = 1.0f - fy, r[3];
    for (int k = 0; k < 3; ++k) r[k] = (t00[k] * gx + t10[k] * fx) * gy + (t01[k] * gx + t11[k] * fx) * fy;
    return mk(r[0], r[1], r[2]);
}

static f3 penv_bilinear(const float* env, int ew, int eh, float u, float v) {
    return penv_bilinear_m(env, ew, eh, u, v, 1);
}

static f3 penv(const float* env, int ew, int eh, f3 d) {
    float u = rm_atan2f(d.y, d.x) / (2.0f * PT_PI) + 0.5f;
    float z = wmin(wmax(d.z, -1.0f), 1.0f);
    float v = 0.5f - rm_asinf(z) / PT_PI;
    return penv_bilinear(env, ew, eh, u, v);
}

static f3 phemi(f3 n, uint32_t* seed) {
    float rx, ry;
    prand2(seed, &rx, &ry);
    float phi = (2.0f * PT_PI) * rx;
    float sr = sqrtf(ry);
    float x = rm_cosf(phi) * sr, y = rm_sinf(phi) * sr, z = sqrtf(wmax(0.0f, 1.0f - ry));
    f3 up = mk(0.0f, 1.0f, 0.0f);
    if (fabsf(n.y) > 0.99f) up = mk(1.0f, 0.0f, 0.0f);
    f3 t = vnorm(vcross(up, n));
    f3 b = vcross(n, t);
    return vnorm(vadd(vadd(vscale(t, x), vscale(b, y)), vscale(n, z)));
}

/* One frame: accum (W*H*4) += the pixel's path radiance and 1. */
void ora_pt_trace(const float* ivp, uint32_t frame, int W, int H, const float* verts, const uint32_t* nodes, int nn,
                  const uint32_t* tris, const float* env, int ew, int eh, float* accum) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            uint32_t pix = (uint32_t)py * (uint32_t)W + (uint32_t)px;
            uint32_t seed = ora_pt_pcg(pix ^ (frame * 2654435761u));
            float ja, jb;
            prand2(&seed, &ja, &jb);
            float jx = ja - 0.5f, jy = jb - 0.5f;
            float ux = (((float)px + 0.5f) + jx) / (float)W;
            float uy = (((float)py + 0.5f) + jy) / (float)H;
            float nx = ux * 2.0f - 1.0f, ny = 1.0f - uy * 2.0f;
            f3 ro = unproj(ivp, nx, ny, 0.0f), fw = unproj(ivp, nx, ny, 1.0f);
            f3 rd = vnorm(vsub(fw, ro));
            f3 rad = mk(0, 0, 0), thr = mk(1, 1, 1), o = ro;
            for (int bounce = 0; bounce < 8; ++bounce) {
                phit h = pscene(o, rd, nodes, nn, tris, verts, 0);
                if (!h.hit) {
                    f3 e = penv(env, ew, eh, rd);
                    rad = vadd(rad, mk(thr.x * e.x, thr.y * e.y, thr.z * e.z));
                    break;
                }
                f3 n = vdot(h.n, rd) > 0.0f ? mk(-h.n.x, -h.n.y, -h.n.z) : h.n;
                o = vadd(vadd(o, vscale(rd, h.t)), vscale(n, 5e-4f));
                if (h.c[3] < 0.5f) {
                    rd = vsub(rd, vscale(n, 2.0f * vdot(n, rd)));
                } else {
                    thr = mk(thr.x * h.c[0], thr.y * h.c[1], thr.z * h.c[2]);
                    rd = phemi(n, &seed);
                }
            }
            float* a = accum + 4 * (size_t)pix;
            a[0] = a[0] + rad.x;
            a[1] = a[1] + rad.y;
            a[2] = a[2] + rad.z;
            a[3] = a[3] + 1.0f;
        }
}

static float punorm8(float x) {
    if (!(x > 0.0f)) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return rintf(x * 255.0f) / 255.0f;
}
static float preinhard(float c) { return c / (c + 1.0f); }

void ora_pt_resolve(const float* accum, int npix, float* out) {
    for (int i = 0; i < npix; ++i) {
        const float* s = accum + 4 * (size_t)i;
        float n = wmax(s[3], 1.0f);
        for (int k = 0; k < 3; ++k) out[4 * (size_t)i + k] = punorm8(preinhard((s[k] / n) * 4.0f));
        out[4 * (size_t)i + 3] = 1.0f;
    }
}

void ora_env_background(const float* ivp, const float* env, int ew, int eh, int W, int H, float* out) {
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            float nx = (((float)px + 0.5f) / (float)W) * 2.0f - 1.0f;
            float ny = 1.0f - (((float)py + 0.5f) / (float)H) * 2.0f;
            f3 nw = unproj(ivp, nx, ny, 0.0f), fw = unproj(ivp, nx, ny, 1.0f);
            f3 c = penv(env, ew, eh, vnorm(vsub(fw, nw)));
            float* o = out + 4 * ((size_t)py * W + px);
            o[0] = punorm8(preinhard(c.x * 4.0f));
            o[1] = punorm8(preinhard(c.y * 4.0f));
            o[2] = punorm8(preinhard(c.z * 4.0f));
            o[3] = 1.0f;
        }
}

/* Closest hit of one ray by brute force over all triangles (for tests of the
 * BVH walk): the same triangle test, triangles in index order. */
void ora_pt_closest_brute(const float* ro3, const float* rd3, const float* verts, const uint32_t* indices, int ntris,
                          float* t_out, int* tri_out) {
    f3 ro = mk(ro3[0], ro3[1], ro3[2]), rd = mk(rd3[0], rd3[1], rd3[2]);
    phit best;
    memset(&best, 0, sizeof(best));
    best.t = 1e30f;
    int which = -1;
    for (int k = 0; k < ntris; ++k) {
        int before = best.hit;
        float tb = best.t;
        ptri_test(ro, rd, verts, indices[3 * (size_t)k], indices[3 * (size_t)k + 1], indices[3 * (size_t)k + 2], best.t,
                  &best, 0);
        if (best.hit && (!before || best.t != tb)) which = k;
    }
    *t_out = best.hit ? best.t : -1.0f;
    *tri_out = which;
}

/* Closest hit of one ray through the BVH: t (or -1) and the normal/colour. */
void ora_pt_closest_bvh(const float* ro3, const float* rd3, const float* verts, const uint32_t* nodes, int nn,
                        const uint32_t* tris, float* out8) {
    phit h = pscene(mk(ro3[0], ro3[1], ro3[2]), mk(rd3[0], rd3[1], rd3[2]), nodes, nn, tris, verts, 0);
    out8[0] = h.hit ? h.t : -1.0f;
    out8[1] = h.n.x; out8[2] = h.n.y; out8[3] = h.n.z;
    for (int k = 0; k < 4; ++k) out8[4 + k] = h.c[k];
}

/* Cosine-hemisphere directions about n from a seed (for distribution tests). */
void ora_pt_hemisphere(const float* n3, uint32_t seed, int count, float* dirs) {
    f3 n = mk(n3[0], n3[1], n3[2]);
    for (int k = 0; k < count; ++k) {
        f3 d = phemi(n, &seed);
        dirs[3 * (size_t)k] = d.x; dirs[3 * (size_t)k + 1] = d.y; dirs[3 * (size_t)k + 2] = d.z;
    }
}

void ora_atan2_asin(const float* y, const float* x, int n, float* atan2_out, float* asin_out) {
    for (int i = 0; i < n; ++i) {
        if (atan2_out) atan2_out[i] = rm_atan2f(y[i], x[i]);
        if (asin_out) asin_out[i] = rm_asinf(y[i]);
    }
}

/* ------------------------------------------------------------ mesh raster */
static float pclipw(const float* m, const float* p) { return ((m[3] * p[0] + m[7] * p[1]) + m[11] * p[2]) + m[15] * 1.0f; }
static f3 pviewdir(const float* m, const float* n) {
    float r[3];
    for (int i = 0; i < 3; ++i) r[i] = ((m[i] * n[0] + m[4 + i] * n[1]) + m[8 + i] * n[2]) + m[12 + i] * 0.0f;
    return mk(r[0], r[1], r[2]);
}
static float pviewdepth(float w) { return wmin(wmax(1.0f - 0.1f / wmax(w, 0.1f), 0.0f), 1.0f); }

/* mesh.wgsl vert/frag over the envmap background, per pixel centre: the
 * closest front-facing triangle along the centre's ray; varyings blended with
 * its barycentric weights. */
void ora_mesh_raster(const float* vp, const float* view, const float* ivp, float shading, int W, int H,
                     const float* verts, const uint32_t* nodes, int nn, const uint32_t* tris, const float* env, int ew,
                     int eh, float* color, float* depth) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            float nx = (((float)px + 0.5f) / (float)W) * 2.0f - 1.0f;
            float ny = 1.0f - (((float)py + 0.5f) / (float)H) * 2.0f;
            f3 nw = unproj(ivp, nx, ny, 0.0f), fw = unproj(ivp, nx, ny, 1.0f);
            f3 rd = vnorm(vsub(fw, nw));
            float* oc = color + 4 * ((size_t)py * W + px);
            float* od = depth + 4 * ((size_t)py * W + px);
            phit h = pscene(nw, rd, nodes, nn, tris, verts, 1);
            f3 fc;
            if (!h.hit) {
                fc = penv(env, ew, eh, rd);
                for (int k = 0; k < 3; ++k) od[k] = 1.0f;
            } else {
                const float* q[3] = { verts + 12 * (size_t)h.tri[0], verts + 12 * (size_t)h.tri[1], verts + 12 * (size_t)h.tri[2] };
                float u = h.bu, v = h.bv, w = h.bw;
                float d = (u * pviewdepth(pclipw(vp, q[0])) + v * pviewdepth(pclipw(vp, q[1]))) +
                          w * pviewdepth(pclipw(vp, q[2]));
                float dq = punorm8(d);
                for (int k = 0; k < 3; ++k) od[k] = dq;
                if (shading < 0.5f) {
                    f3 nn3 = vadd(vadd(vscale(mk(q[0][3], q[0][4], q[0][5]), u), vscale(mk(q[1][3], q[1][4], q[1][5]), v)),
                                  vscale(mk(q[2][3], q[2][4], q[2][5]), w));
                    f3 qq = vnorm(nn3);
                    fc = mk((qq.x + 1.0f) * 0.5f, (qq.y + 1.0f) * 0.5f, (qq.z + 1.0f) * 0.5f);
                } else {
                    f3 vn = vnorm(vadd(vadd(vscale(pviewdir(view, q[0] + 3), u), vscale(pviewdir(view, q[1] + 3), v)),
                                       vscale(pviewdir(view, q[2] + 3), w)));
                    float mu = vn.x * 0.5f + 0.5f, mv = vn.y * 0.5f + 0.5f;
                    f3 m = penv_bilinear_m(env, ew, eh, mu, 1.0f - mv, 0);
                    fc = mk(m.x * ((u * q[0][6] + v * q[1][6]) + w * q[2][6]), m.y * ((u * q[0][7] + v * q[1][7]) + w * q[2][7]),
                            m.z * ((u * q[0][8] + v * q[1][8]) + w * q[2][8]));
                }
            }
            oc[0] = punorm8(preinhard(fc.x * 4.0f));
            oc[1] = punorm8(preinhard(fc.y * 4.0f));
            oc[2] = punorm8(preinhard(fc.z * 4.0f));
            oc[3] = 1.0f;
            od[3] = 1.0f;
        }
}
