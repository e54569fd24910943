// nbz_api.cpp -- C ABI of the Bezier stroke optimizer (nbz_*, declared in
// include/nprush_hip.h), in the order of GpuBezierOptimizerManager.dispatch
// (src/lib/gpu/GpuBezierOptimizerManager.ts:370-406):
//   backward = curve preprocess -> per-tile id-block cut (first 1024 ascending
//              ids per tile) -> key scan and emission -> LSD radix sort on the
//              tile bits -> tile ranges -> fused composite / reverse pass
//   step     = Adam with the view's pixel count
//   adc      = every adc_period-th dispatch
// Everything is enqueued on the handle's stream (default priority).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/nprush_hip.h"
#include "nps_bezier.h"
#include "nps_host.h"
#include "nps_kernels.h"

using namespace nps;

struct nbz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    float* curves = nullptr;
    int32_t* grads = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* grad_accum = nullptr;
    unsigned long long* loss_acc = nullptr;
    float adam_t = 0.0f;
    int64_t step_count = 0;
    int32_t adc_period = 50;
    float vp[16] = {0}, vp_inv[16] = {0};
    BzParams prm = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t W = 0, H = 0;
    bool has_target = false;
    DevBuf own_tgt, own_tdepth, own_bg, own_bgd, own_normal;
    const float *tgt = nullptr, *tdepth = nullptr, *bg = nullptr, *bgd = nullptr, *normal = nullptr;
    DevBuf pixel_loss;
    uint32_t pl_w = 0, pl_h = 0;
    DevBuf fwd, bwd, ckpt, c_img, d_img, t_img, tile_loss, stats, pa, pd, pk, act, nzm, seeds, part;
    Scratch sc;
    std::vector<struct nbf_ctx*> draws;  // display-raster handles drawing from this one's curves
    uint32_t last_tiles = 0;
    uint32_t img_w = 0, img_h = 0;
    bool have_image = false;
    DevBuf adc_dead, adc_cand, adc_stats;
    bool have_adc_stats = false;
    bool timing = false;
    hipEvent_t ev[8] = {nullptr};
    float timing_ms[8] = {0};
    std::string err;
};

namespace {

int bz_err(nbz_ctx* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define NBZ_HIP(h, expr)                                                                                   \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return bz_err((h), NPS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int bz_ensure(nbz_ctx* h, DevBuf& b, size_t bytes) { return ensure_dev(b, bytes, h->stream, h->err); }

int check_n(nbz_ctx* h, int32_t n) {
    if (!h) return NPS_ERR_ARG;
    if (n != (int32_t)h->n) return bz_err(h, NPS_ERR_ARG, "count does not match the handle's curve count");
    return NPS_OK;
}

// pixel_loss follows the target size; a new size starts from zero
int bz_pixel_loss(nbz_ctx* h) {
    if (h->pl_w == h->W && h->pl_h == h->H && h->pixel_loss.p) return NPS_OK;
    const size_t px = (size_t)h->W * h->H;
    int rc = bz_ensure(h, h->pixel_loss, 4 * std::max<size_t>(px, 1));
    if (rc != NPS_OK) return rc;
    launch_fill_u32(h->stream, (uint32_t*)h->pixel_loss.p, 0u, std::max<size_t>(px, 1));
    h->pl_w = h->W;
    h->pl_h = h->H;
    return NPS_OK;
}

int bz_run(nbz_ctx* h, bool backward) {
    if (!h->has_target) return bz_err(h, NPS_ERR_STATE, "no backward target (nbz_set_backward_target)");
    NBZ_HIP(h, hipSetDevice(h->device));
    const uint32_t n = h->n;
    ViewParams v = make_view(h->vp, h->vp_inv, h->W, h->H);
    Scratch& sc = h->sc;
    int rc;
    const size_t n1 = std::max<uint32_t>(n, 1);
    if ((rc = bz_ensure(h, h->fwd, sizeof(BzFwd) * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->bwd, sizeof(BzBwd) * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, sc.counts, 4 * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, sc.rects, 16 * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->stats, 128)) != NPS_OK) return rc;
    if ((rc = bz_pixel_loss(h)) != NPS_OK) return rc;
    const uint32_t cap = (uint32_t)kBzCap;
    sc.splat_block = splat_block_of(n);
    const uint32_t nsb = ((uint32_t)n1 + sc.splat_block - 1) / sc.splat_block;
    const size_t diff_n = 4 * (size_t)nsb * (v.tiles_x + 1) * (v.tiles_y + 1) + 16 * (v.tiles_x + 1);
    const size_t t1 = std::max<uint32_t>(v.num_tiles, 1);
    if ((rc = bz_ensure(h, sc.diff, 4 * diff_n)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, sc.cut, 4 * t1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, sc.full, 4 * t1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, sc.kept, 4 * n1)) != NPS_OK) return rc;
    if (h->timing) NBZ_HIP(h, hipEventRecord(h->ev[0], h->stream));
    launch_fill_u32(h->stream, (uint32_t*)h->stats.p, 0u, 32);
    launch_bz_preprocess(h->stream, h->curves, n, v, (BzFwd*)h->fwd.p, (BzBwd*)h->bwd.p, (uint32_t*)sc.counts.p,
                         (uint4*)sc.rects.p);
    launch_tile_cut(h->stream, v, n, (int32_t*)sc.diff.p, sc.splat_block, cap, (uint32_t*)sc.cut.p,
                    (uint32_t*)sc.full.p, (unsigned long long*)h->stats.p + 1, (const uint32_t*)sc.counts.p,
                    (const uint4*)sc.rects.p, (uint32_t*)sc.kept.p);
    sc.have_full = true;
    if (h->timing) NBZ_HIP(h, hipEventRecord(h->ev[1], h->stream));
    if ((rc = build_tile_lists_on(h->stream, h->err, sc, v, n, h->timing ? h->ev[2] : nullptr, nullptr)) != NPS_OK)
        return rc;
    const uint32_t nchunk = (cap + kChunk - 1) / kChunk;
    const size_t px = (size_t)h->W * h->H;
    if ((rc = bz_ensure(h, h->ckpt, sizeof(float) * t1 * nchunk * kTilePx)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->c_img, 12 * std::max<size_t>(px, 1))) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->d_img, 4 * std::max<size_t>(px, 1))) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->t_img, 4 * std::max<size_t>(px, 1))) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->tile_loss, 16 * t1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->part, 40 * t1)) != NPS_OK) return rc;
    const size_t pslots = (size_t)std::max<uint64_t>(sc.m, 1) * kTilePx;
    if (backward) {
        if ((rc = bz_ensure(h, h->pa, 4 * pslots)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->pd, 4 * pslots)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->pk, pslots)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->act, 4 * t1 * 4 * (kBzCap / 32))) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->nzm, 4 * t1 * 4 * (kBzCap / 32))) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->seeds, 4 * 9 * t1 * kTilePx)) != NPS_OK) return rc;
    }
    if (h->timing) NBZ_HIP(h, hipEventRecord(h->ev[3], h->stream));
    BzCompositeArgs a;
    a.fwd = (const BzFwd*)h->fwd.p;
    a.bwd = (const BzBwd*)h->bwd.p;
    a.keys = (const uint64_t*)(sc.in_alt ? sc.keys_alt.p : sc.keys.p);
    a.ranges = (const uint32_t*)sc.ranges.p;
    a.tgt = (const float4*)h->tgt;
    a.tgt_depth = h->tdepth;
    a.bg = (const float4*)h->bg;
    a.bg_depth = h->bgd;
    a.normal = (const float4*)h->normal;
    a.grads = h->grads;
    a.loss_acc = h->loss_acc;
    a.pixel_loss = (int32_t*)h->pixel_loss.p;
    a.c_img = (float*)h->c_img.p;
    a.d_img = (float*)h->d_img.p;
    a.t_img = (float*)h->t_img.p;
    a.tile_loss = (double*)h->tile_loss.p;
    a.ckpt = (float*)h->ckpt.p;
    a.pa = (float*)h->pa.p;
    a.pd = (float*)h->pd.p;
    a.pk = (uint8_t*)h->pk.p;
    a.act = (uint32_t*)h->act.p;
    a.nzm = (uint32_t*)h->nzm.p;
    a.seeds = (float*)h->seeds.p;
    a.stats = (unsigned long long*)h->stats.p;
    a.part = (unsigned long long*)h->part.p;
    a.prm = h->prm;
    a.cap = cap;
    a.nchunk_max = nchunk;
    a.backward = backward ? 1 : 0;
    launch_bz_composite(h->stream, v, a);
    launch_sum_parts(h->stream, a.part, v.num_tiles, PartMap{5, backward ? 5u : 2u, {0, 4, 2, 3, 5, 0, 0, 0}},
                     a.stats);
    if (h->timing) NBZ_HIP(h, hipEventRecord(h->ev[4], h->stream));
    NBZ_HIP(h, hipGetLastError());
    h->last_tiles = v.num_tiles;
    h->img_w = h->W;
    h->img_h = h->H;
    h->have_image = true;
    if (h->timing) {
        NBZ_HIP(h, hipEventSynchronize(h->ev[4]));
        float t;
        hipEventElapsedTime(&t, h->ev[0], h->ev[1]); h->timing_ms[0] = t;
        hipEventElapsedTime(&t, h->ev[1], h->ev[2]); h->timing_ms[1] = t;
        hipEventElapsedTime(&t, h->ev[2], h->ev[3]); h->timing_ms[2] = t;
        hipEventElapsedTime(&t, h->ev[3], h->ev[4]); h->timing_ms[3] = t;
        hipEventElapsedTime(&t, h->ev[0], h->ev[4]); h->timing_ms[6] = t;
    }
    return NPS_OK;
}

int bz_copy(nbz_ctx* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    NBZ_HIP(h, hipSetDevice(h->device));
    if (bytes == 0) return NPS_OK;
    NBZ_HIP(h, hipMemcpyAsync(dst, src, bytes, kind, h->stream));
    NBZ_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int bz_set_target(nbz_ctx* h, const float* t, const float* td, const float* b, const float* bd, const float* nr,
                  int32_t w, int32_t hh, bool dev) {
    if (!h) return NPS_ERR_ARG;
    if (w <= 0 || hh <= 0) return bz_err(h, NPS_ERR_ARG, "target size must be positive");
    if (!t || !td || !b || !bd || !nr) return bz_err(h, NPS_ERR_ARG, "all five target images are required");
    NBZ_HIP(h, hipSetDevice(h->device));
    const size_t px = (size_t)w * hh;
    if (dev) {
        h->tgt = t;
        h->tdepth = td;
        h->bg = b;
        h->bgd = bd;
        h->normal = nr;
    } else {
        int rc;
        if ((rc = bz_ensure(h, h->own_tgt, 16 * px)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->own_tdepth, 4 * px)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->own_bg, 16 * px)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->own_bgd, 4 * px)) != NPS_OK) return rc;
        if ((rc = bz_ensure(h, h->own_normal, 16 * px)) != NPS_OK) return rc;
        NBZ_HIP(h, hipMemcpyAsync(h->own_tgt.p, t, 16 * px, hipMemcpyHostToDevice, h->stream));
        NBZ_HIP(h, hipMemcpyAsync(h->own_tdepth.p, td, 4 * px, hipMemcpyHostToDevice, h->stream));
        NBZ_HIP(h, hipMemcpyAsync(h->own_bg.p, b, 16 * px, hipMemcpyHostToDevice, h->stream));
        NBZ_HIP(h, hipMemcpyAsync(h->own_bgd.p, bd, 4 * px, hipMemcpyHostToDevice, h->stream));
        NBZ_HIP(h, hipMemcpyAsync(h->own_normal.p, nr, 16 * px, hipMemcpyHostToDevice, h->stream));
        NBZ_HIP(h, hipStreamSynchronize(h->stream));
        h->tgt = (const float*)h->own_tgt.p;
        h->tdepth = (const float*)h->own_tdepth.p;
        h->bg = (const float*)h->own_bg.p;
        h->bgd = (const float*)h->own_bgd.p;
        h->normal = (const float*)h->own_normal.p;
    }
    h->W = (uint32_t)w;
    h->H = (uint32_t)hh;
    h->has_target = true;
    return NPS_OK;
}

}  // namespace

struct nbf_ctx;
static void nbf_release(nbf_ctx* f);

extern "C" {

int nbz_create(int32_t device, int32_t num_beziers, nbz_handle* out) {
    if (!out || num_beziers < 0) return NPS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NPS_ERR_HIP;
    if (device < 0 || device >= ndev) return NPS_ERR_ARG;
    nbz_ctx* h = new nbz_ctx();
    h->device = device;
    h->n = (uint32_t)num_beziers;
    auto fail = [&](int code) {
        nbz_destroy(h);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(NPS_ERR_HIP);
    const size_t n1 = std::max<uint32_t>(h->n, 1);
    if (hipMalloc((void**)&h->curves, 80 * n1) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMalloc((void**)&h->grads, 72 * n1) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMalloc((void**)&h->m, 72 * n1) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMalloc((void**)&h->v, 72 * n1) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMalloc((void**)&h->grad_accum, 4 * n1) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMalloc((void**)&h->loss_acc, 8 * n1) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMemsetAsync(h->curves, 0, 80 * n1, h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMemsetAsync(h->grads, 0, 72 * n1, h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMemsetAsync(h->m, 0, 72 * n1, h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMemsetAsync(h->v, 0, 72 * n1, h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMemsetAsync(h->grad_accum, 0, 4 * n1, h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipMemsetAsync(h->loss_acc, 0, 8 * n1, h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    for (int k = 0; k < 8; ++k)
        if (hipEventCreate(&h->ev[k]) != hipSuccess) return fail(NPS_ERR_HIP);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(NPS_ERR_HIP);
    *out = h;
    return NPS_OK;
}

void nbz_destroy(nbz_handle h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    // display-raster handles of these curves: their device memory goes now, and
    // every later call on them returns NPS_ERR_STATE (nbf_destroy still frees them)
    for (nbf_ctx* f : h->draws) nbf_release(f);
    h->draws.clear();
    void* ps[] = {h->curves, h->grads, h->m, h->v, h->grad_accum, h->loss_acc};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    DevBuf* bs[] = {&h->own_tgt, &h->own_tdepth, &h->own_bg, &h->own_bgd, &h->own_normal, &h->pixel_loss,
                    &h->fwd, &h->bwd, &h->ckpt, &h->c_img, &h->d_img, &h->t_img, &h->tile_loss, &h->stats,
                    &h->pa, &h->pd, &h->pk, &h->act, &h->nzm, &h->seeds,
                    &h->adc_dead, &h->adc_cand, &h->adc_stats, &h->part};
    for (DevBuf* b : bs) free_buf(*b);
    free_scratch(h->sc);
    for (int k = 0; k < 8; ++k)
        if (h->ev[k]) (void)hipEventDestroy(h->ev[k]);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* nbz_last_error(nbz_handle h) { return h ? h->err.c_str() : "null handle"; }

int nbz_write_beziers(nbz_handle h, const float* c, int32_t n) {
    int rc = check_n(h, n);
    return rc != NPS_OK ? rc : bz_copy(h, h->curves, c, 80 * (size_t)n, hipMemcpyHostToDevice);
}
int nbz_read_beziers(nbz_handle h, float* c, int32_t n) {
    int rc = check_n(h, n);
    return rc != NPS_OK ? rc : bz_copy(h, c, h->curves, 80 * (size_t)n, hipMemcpyDeviceToHost);
}
int nbz_write_grads(nbz_handle h, const int32_t* g, int32_t n) {
    int rc = check_n(h, n);
    return rc != NPS_OK ? rc : bz_copy(h, h->grads, g, 72 * (size_t)n, hipMemcpyHostToDevice);
}
int nbz_read_grads(nbz_handle h, int32_t* g, int32_t n) {
    int rc = check_n(h, n);
    return rc != NPS_OK ? rc : bz_copy(h, g, h->grads, 72 * (size_t)n, hipMemcpyDeviceToHost);
}
int nbz_write_adam(nbz_handle h, const float* m, const float* v, float t, int32_t n) {
    int rc = check_n(h, n);
    if (rc != NPS_OK) return rc;
    if ((rc = bz_copy(h, h->m, m, 72 * (size_t)n, hipMemcpyHostToDevice)) != NPS_OK) return rc;
    if ((rc = bz_copy(h, h->v, v, 72 * (size_t)n, hipMemcpyHostToDevice)) != NPS_OK) return rc;
    h->adam_t = t;
    return NPS_OK;
}
int nbz_read_adam(nbz_handle h, float* m, float* v, float* t, int32_t n) {
    int rc = check_n(h, n);
    if (rc != NPS_OK) return rc;
    if (m && (rc = bz_copy(h, m, h->m, 72 * (size_t)n, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    if (v && (rc = bz_copy(h, v, h->v, 72 * (size_t)n, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    if (t) *t = h->adam_t;
    return NPS_OK;
}
int nbz_write_adc(nbz_handle h, const float* ga, const int64_t* la, int32_t n) {
    int rc = check_n(h, n);
    if (rc != NPS_OK) return rc;
    if ((rc = bz_copy(h, h->grad_accum, ga, 4 * (size_t)n, hipMemcpyHostToDevice)) != NPS_OK) return rc;
    return bz_copy(h, h->loss_acc, la, 8 * (size_t)n, hipMemcpyHostToDevice);
}
int nbz_read_adc(nbz_handle h, float* ga, int64_t* la, int32_t n) {
    int rc = check_n(h, n);
    if (rc != NPS_OK) return rc;
    if (ga && (rc = bz_copy(h, ga, h->grad_accum, 4 * (size_t)n, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    if (la && (rc = bz_copy(h, la, h->loss_acc, 8 * (size_t)n, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    return NPS_OK;
}
int nbz_write_pixel_loss(nbz_handle h, const int32_t* pl) {
    if (!h) return NPS_ERR_ARG;
    if (!h->has_target) return bz_err(h, NPS_ERR_STATE, "no target size (nbz_set_backward_target)");
    int rc = bz_pixel_loss(h);
    if (rc != NPS_OK) return rc;
    return bz_copy(h, h->pixel_loss.p, pl, 4 * (size_t)h->W * h->H, hipMemcpyHostToDevice);
}
int nbz_read_pixel_loss(nbz_handle h, int32_t* pl) {
    if (!h) return NPS_ERR_ARG;
    if (!h->has_target) return bz_err(h, NPS_ERR_STATE, "no target size (nbz_set_backward_target)");
    int rc = bz_pixel_loss(h);
    if (rc != NPS_OK) return rc;
    return bz_copy(h, pl, h->pixel_loss.p, 4 * (size_t)h->W * h->H, hipMemcpyDeviceToHost);
}
int nbz_write_vp_matrix(nbz_handle h, const float vp[16]) {
    if (!h || !vp) return NPS_ERR_ARG;
    memcpy(h->vp, vp, sizeof(h->vp));
    return NPS_OK;
}
int nbz_write_vp_inv_matrix(nbz_handle h, const float vp_inv[16]) {
    if (!h || !vp_inv) return NPS_ERR_ARG;
    memcpy(h->vp_inv, vp_inv, sizeof(h->vp_inv));
    return NPS_OK;
}
int nbz_write_mode(nbz_handle h, float mode) {
    if (!h) return NPS_ERR_ARG;
    h->prm.mode = mode;
    return NPS_OK;
}
int nbz_write_bg_penalty(nbz_handle h, float weight) {
    if (!h) return NPS_ERR_ARG;
    h->prm.bg_penalty = weight;
    return NPS_OK;
}
int nbz_write_max_width(nbz_handle h, float max_width) {
    if (!h) return NPS_ERR_ARG;
    h->prm.max_width = max_width;
    return NPS_OK;
}
int nbz_write_kill_thresholds(nbz_handle h, float alpha, float width) {
    if (!h) return NPS_ERR_ARG;
    h->prm.prune_alpha = alpha;
    h->prm.prune_width = width;
    return NPS_OK;
}
int nbz_set_adc_period(nbz_handle h, int32_t period) {
    if (!h || period <= 0) return NPS_ERR_ARG;
    h->adc_period = period;
    return NPS_OK;
}
int nbz_write_no_kill(nbz_handle h, int32_t no_kill) {
    if (!h) return NPS_ERR_ARG;
    h->prm.no_kill = no_kill ? 1.0f : 0.0f;
    return NPS_OK;
}
int nbz_set_backward_target(nbz_handle h, const float* t, const float* td, const float* b, const float* bd,
                            const float* nr, int32_t w, int32_t hh) {
    return bz_set_target(h, t, td, b, bd, nr, w, hh, false);
}
int nbz_set_backward_target_dev(nbz_handle h, const float* t, const float* td, const float* b, const float* bd,
                                const float* nr, int32_t w, int32_t hh) {
    return bz_set_target(h, t, td, b, bd, nr, w, hh, true);
}

int nbz_backward(nbz_handle h) {
    if (!h) return NPS_ERR_ARG;
    return bz_run(h, true);
}
int nbz_forward(nbz_handle h) {
    if (!h) return NPS_ERR_ARG;
    return bz_run(h, false);
}
int nbz_step(nbz_handle h) {
    if (!h) return NPS_ERR_ARG;
    NBZ_HIP(h, hipSetDevice(h->device));
    if (h->timing) NBZ_HIP(h, hipEventRecord(h->ev[5], h->stream));
    // dispatch writes pixel_count = width * height of the target (:375-380)
    const float pixel_count = (float)((double)h->W * (double)h->H);
    launch_bz_step(h->stream, h->curves, h->grads, h->m, h->v, h->grad_accum, h->n, h->vp, h->prm,
                   h->adam_t + 1.0f, pixel_count);
    h->adam_t += 1.0f;
    if (h->timing) {
        NBZ_HIP(h, hipEventRecord(h->ev[6], h->stream));
        NBZ_HIP(h, hipEventSynchronize(h->ev[6]));
        hipEventElapsedTime(&h->timing_ms[4], h->ev[5], h->ev[6]);
    }
    NBZ_HIP(h, hipGetLastError());
    return NPS_OK;
}
int nbz_adc(nbz_handle h) {
    if (!h) return NPS_ERR_ARG;
    if (!h->has_target) return bz_err(h, NPS_ERR_STATE, "no target size (nbz_set_backward_target)");
    NBZ_HIP(h, hipSetDevice(h->device));
    int rc;
    const size_t n1 = std::max<uint32_t>(h->n, 1);
    if ((rc = bz_pixel_loss(h)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->adc_dead, 4 * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->adc_cand, 4 * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, h->adc_stats, 64)) != NPS_OK) return rc;
    if (h->timing) NBZ_HIP(h, hipEventRecord(h->ev[5], h->stream));
    BzAdcBuffers b;
    b.dead_list = (uint32_t*)h->adc_dead.p;
    b.cand = (uint32_t*)h->adc_cand.p;
    b.stats = (int64_t*)h->adc_stats.p;
    launch_bz_adc(h->stream, h->curves, h->m, h->v, h->grad_accum, h->loss_acc, (int32_t*)h->pixel_loss.p, h->n,
                  h->W, h->H, h->vp, h->vp_inv, h->prm, h->adam_t, b);
    h->have_adc_stats = true;
    if (h->timing) {
        NBZ_HIP(h, hipEventRecord(h->ev[6], h->stream));
        NBZ_HIP(h, hipEventSynchronize(h->ev[6]));
        hipEventElapsedTime(&h->timing_ms[5], h->ev[5], h->ev[6]);
    }
    NBZ_HIP(h, hipGetLastError());
    return NPS_OK;
}
int nbz_dispatch(nbz_handle h) {
    if (!h) return NPS_ERR_ARG;
    int rc;
    if ((rc = nbz_backward(h)) != NPS_OK) return rc;
    if ((rc = nbz_step(h)) != NPS_OK) return rc;
    h->step_count += 1;
    if (h->step_count % h->adc_period == 0 && (rc = nbz_adc(h)) != NPS_OK) return rc;
    return NPS_OK;
}
int nbz_sync(nbz_handle h) {
    if (!h) return NPS_ERR_ARG;
    NBZ_HIP(h, hipSetDevice(h->device));
    NBZ_HIP(h, hipStreamSynchronize(h->stream));
    return NPS_OK;
}

int nbz_read_image(nbz_handle h, float* color, float* depth, float* trans) {
    if (!h) return NPS_ERR_ARG;
    if (!h->have_image) return bz_err(h, NPS_ERR_STATE, "no image yet (nbz_backward / nbz_forward)");
    const size_t px = (size_t)h->img_w * h->img_h;
    int rc;
    if (color && (rc = bz_copy(h, color, h->c_img.p, 12 * px, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    if (depth && (rc = bz_copy(h, depth, h->d_img.p, 4 * px, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    if (trans && (rc = bz_copy(h, trans, h->t_img.p, 4 * px, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
    return NPS_OK;
}
int nbz_read_loss(nbz_handle h, double loss[2]) {
    if (!h || !loss) return NPS_ERR_ARG;
    if (!h->have_image) return bz_err(h, NPS_ERR_STATE, "no image yet (nbz_backward / nbz_forward)");
    std::vector<double> tl(2 * (size_t)h->last_tiles);
    int rc = bz_copy(h, tl.data(), h->tile_loss.p, 16 * (size_t)h->last_tiles, hipMemcpyDeviceToHost);
    if (rc != NPS_OK) return rc;
    double a = 0.0, b = 0.0;
    for (uint32_t t = 0; t < h->last_tiles; ++t) {
        a += tl[2 * t];
        b += tl[2 * t + 1];
    }
    loss[0] = a;
    loss[1] = b;
    return NPS_OK;
}
int nbz_read_tiles(nbz_handle h, uint32_t* ranges, uint32_t* ids, int32_t* num_tiles, int64_t* num_keys) {
    if (!h) return NPS_ERR_ARG;
    if (!h->have_image) return bz_err(h, NPS_ERR_STATE, "no tile lists yet (nbz_backward / nbz_forward)");
    if (num_tiles) *num_tiles = (int32_t)h->last_tiles;
    if (num_keys) *num_keys = (int64_t)h->sc.m;
    int rc;
    if (ranges && (rc = bz_copy(h, ranges, h->sc.ranges.p, 8 * (size_t)h->last_tiles, hipMemcpyDeviceToHost)) != NPS_OK)
        return rc;
    if (ids && h->sc.m) {
        std::vector<uint64_t> keys(h->sc.m);
        const void* src = h->sc.in_alt ? h->sc.keys_alt.p : h->sc.keys.p;
        if ((rc = bz_copy(h, keys.data(), src, 8 * (size_t)h->sc.m, hipMemcpyDeviceToHost)) != NPS_OK) return rc;
        for (uint64_t k = 0; k < h->sc.m; ++k) ids[k] = (uint32_t)keys[k];
    }
    return NPS_OK;
}
int nbz_read_stats(nbz_handle h, int64_t stats[16]) {
    if (!h || !stats) return NPS_ERR_ARG;
    for (int k = 0; k < 16; ++k) stats[k] = 0;
    int rc;
    if (h->have_image) {
        uint64_t d[8];
        if ((rc = bz_copy(h, d, h->stats.p, sizeof(d), hipMemcpyDeviceToHost)) != NPS_OK) return rc;
        stats[0] = (int64_t)h->sc.m;
        stats[1] = (int64_t)d[1];
        stats[2] = (int64_t)d[0];
        stats[3] = (int64_t)d[2];
        stats[4] = (int64_t)d[3];
        stats[5] = (int64_t)d[4];
        stats[12] = (int64_t)d[5];
    }
    if (h->have_adc_stats) {
        int64_t a[6];
        if ((rc = bz_copy(h, a, h->adc_stats.p, sizeof(a), hipMemcpyDeviceToHost)) != NPS_OK) return rc;
        for (int k = 0; k < 6; ++k) stats[6 + k] = a[k];
    }
    return NPS_OK;
}
int nbz_enable_timing(nbz_handle h, int32_t on) {
    if (!h) return NPS_ERR_ARG;
    h->timing = on != 0;
    return NPS_OK;
}
int nbz_read_timing(nbz_handle h, float ms[8]) {
    if (!h || !ms) return NPS_ERR_ARG;
    for (int k = 0; k < 8; ++k) ms[k] = h->timing_ms[k];
    return NPS_OK;
}

int nbz_dev_alloc(nbz_handle h, int64_t bytes, void** out) {
    if (!h || !out || bytes < 0) return NPS_ERR_ARG;
    NBZ_HIP(h, hipSetDevice(h->device));
    *out = nullptr;
    NBZ_HIP(h, hipMalloc(out, (size_t)std::max<int64_t>(bytes, 1)));
    return NPS_OK;
}
int nbz_dev_free(nbz_handle h, void* p) {
    if (!h) return NPS_ERR_ARG;
    NBZ_HIP(h, hipSetDevice(h->device));
    NBZ_HIP(h, hipStreamSynchronize(h->stream));
    if (p) NBZ_HIP(h, hipFree(p));
    return NPS_OK;
}
int nbz_dev_upload(nbz_handle h, void* dst, const void* src, int64_t bytes) {
    if (!h || !dst || !src || bytes < 0) return NPS_ERR_ARG;
    return bz_copy(h, dst, src, (size_t)bytes, hipMemcpyHostToDevice);
}

}  // extern "C"

// ------------------------------------------------ display raster (nbf_*)

struct nbf_ctx {
    nbz_ctx* owner = nullptr;
    float vp[16] = {0};
    bool has_vp = false;
    uint32_t W = 0, H = 0;
    int32_t bw = 0, bh = 0;
    DevBuf rgba, rec, brush;
    Scratch sc;
};

// Device memory of a display-raster handle; the handle is then detached from
// its (possibly already destroyed) owner.
static void nbf_release(nbf_ctx* f) {
    if (!f->owner) return;
    (void)hipSetDevice(f->owner->device);
    (void)hipStreamSynchronize(f->owner->stream);
    free_buf(f->rgba);
    free_buf(f->rec);
    free_buf(f->brush);
    free_scratch(f->sc);
    f->owner = nullptr;
}

extern "C" {

int nbf_create(nbz_handle owner, const uint8_t* brush_rgba8, int32_t bw, int32_t bh, nbf_handle* out) {
    if (!owner || !out) return NPS_ERR_ARG;
    *out = nullptr;
    if (!brush_rgba8 || bw <= 0 || bh <= 0) return bz_err(owner, NPS_ERR_ARG, "brush texture missing or empty");
    NBZ_HIP(owner, hipSetDevice(owner->device));
    nbf_ctx* f = new nbf_ctx();
    f->owner = owner;
    // texel values of the red channel as an rgba8unorm texture returns them
    std::vector<float> tex((size_t)bw * bh);
    for (size_t i = 0; i < tex.size(); ++i) tex[i] = (float)brush_rgba8[4 * i] / 255.0f;
    int rc = bz_ensure(owner, f->brush, 4 * tex.size());
    if (rc == NPS_OK) rc = bz_copy(owner, f->brush.p, tex.data(), 4 * tex.size(), hipMemcpyHostToDevice);
    if (rc != NPS_OK) {
        nbf_destroy(f);
        return rc;
    }
    f->bw = bw;
    f->bh = bh;
    owner->draws.push_back(f);
    *out = f;
    return NPS_OK;
}

void nbf_destroy(nbf_handle f) {
    if (!f) return;
    if (f->owner) {
        auto& d = f->owner->draws;
        d.erase(std::remove(d.begin(), d.end(), f), d.end());
        nbf_release(f);
    }
    delete f;
}

int nbf_write_vp(nbf_handle f, const float vp[16]) {
    if (!f || !vp) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;  // its curves' handle is destroyed
    memcpy(f->vp, vp, sizeof(f->vp));
    f->has_vp = true;
    return NPS_OK;
}

int nbf_set_target(nbf_handle f, int32_t w, int32_t hh) {
    if (!f) return NPS_ERR_ARG;
    nbz_ctx* h = f->owner;
    if (!h) return NPS_ERR_STATE;
    if (w <= 0 || hh <= 0) return bz_err(h, NPS_ERR_ARG, "bad render target size");
    NBZ_HIP(h, hipSetDevice(h->device));
    int rc = bz_ensure(h, f->rgba, 16 * (size_t)w * hh);
    if (rc != NPS_OK) return rc;
    // a new size starts cleared to 0 (a dispatch with clear = 0 blends over it)
    if ((uint32_t)w != f->W || (uint32_t)hh != f->H) launch_fill_u32(h->stream, (uint32_t*)f->rgba.p, 0u, 4 * (uint64_t)w * hh);
    f->W = (uint32_t)w;
    f->H = (uint32_t)hh;
    return NPS_OK;
}

int nbf_dispatch(nbf_handle f, int32_t clear, int32_t quantize) {
    if (!f) return NPS_ERR_ARG;
    nbz_ctx* h = f->owner;
    if (!h) return NPS_ERR_STATE;
    if (!f->has_vp || f->W == 0) return bz_err(h, NPS_ERR_STATE, "render target or VP not set");
    NBZ_HIP(h, hipSetDevice(h->device));
    const uint32_t n = h->n;
    const size_t n1 = std::max<uint32_t>(n, 1);
    ViewParams v = make_view(f->vp, nullptr, f->W, f->H);
    int rc;
    if ((rc = bz_ensure(h, f->rec, sizeof(BzDraw) * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, f->sc.counts, 4 * n1)) != NPS_OK) return rc;
    if ((rc = bz_ensure(h, f->sc.rects, 16 * n1)) != NPS_OK) return rc;
    f->sc.have_full = false;
    f->sc.splat_block = 0;
    launch_bz_draw_preprocess(h->stream, h->curves, n, v, (BzDraw*)f->rec.p, (uint32_t*)f->sc.counts.p,
                              (uint4*)f->sc.rects.p);
    if ((rc = build_tile_lists_on(h->stream, h->err, f->sc, v, n, nullptr, nullptr)) != NPS_OK) return rc;
    BzDrawArgs a;
    a.rec = (const BzDraw*)f->rec.p;
    a.keys = (const uint64_t*)(f->sc.in_alt ? f->sc.keys_alt.p : f->sc.keys.p);
    a.ranges = (const uint32_t*)f->sc.ranges.p;
    a.brush = (const float*)f->brush.p;
    a.bw = f->bw;
    a.bh = f->bh;
    a.rgba = (float4*)f->rgba.p;
    a.clear = clear != 0;
    a.quantize = quantize != 0;
    launch_bz_draw(h->stream, v, a);
    NBZ_HIP(h, hipGetLastError());
    return NPS_OK;
}

int nbf_read_target(nbf_handle f, float* rgba) {
    if (!f || !rgba) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;
    return bz_copy(f->owner, rgba, f->rgba.p, 16 * (size_t)f->W * f->H, hipMemcpyDeviceToHost);
}

int nbf_write_target(nbf_handle f, const float* rgba) {
    if (!f || !rgba) return NPS_ERR_ARG;
    if (!f->owner) return NPS_ERR_STATE;
    return bz_copy(f->owner, f->rgba.p, rgba, 16 * (size_t)f->W * f->H, hipMemcpyHostToDevice);
}

}  // extern "C"
