/*
 * nprush_hip.h -- C ABI of libnprush_hip.so, the MI355X (gfx950) splat
 * rasterizer and optimizer loop.
 *
 * Plain C types only: pointers, sizes, int32 status codes.  No exceptions cross
 * the boundary.  Host pointers unless the name ends in _dev (device pointers on
 * the handle's GPU).  Work is enqueued on the handle's own HIP stream; every
 * nps_read_* call and nps_sync() wait for it.  One handle is not thread-safe;
 * separate handles are independent.  Status: 0 = ok, < 0 = error (see
 * nps_last_error()).
 *
 * The entry points replace the WebGPU objects of the reference
 * (wooningeire/nprush):
 *   GpuSplatOptimizerManager        src/lib/gpu/GpuSplatOptimizerManager.ts
 *   GpuSplatForwardPipelineManager  src/lib/gpu/GpuSplatForwardPipelineManager.ts
 * Each declaration names the reference method or buffer it stands for.
 */
#ifndef NPRUSH_HIP_H
#define NPRUSH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NPS_OK 0
#define NPS_ERR_ARG (-1)
#define NPS_ERR_HIP (-2)
#define NPS_ERR_STATE (-3)
#define NPS_ERR_CAPACITY (-4)

/* Per-splat layout (16 fp32, 64 B), splat_backward.wgsl:1-6:
 * [x y z sx] [r g b opacity] [qw qx qy qz] [sy shape_a shape_b sz]. */
#define NPS_PARAMS_PER_SPLAT 16

/*
 * Tunables, with the reference's values as defaults (nps_config_default).
 * Replaces GPU_CONSTANTS + injectWgslConstants (src/lib/gpu/constants.ts:5-55)
 * and the host cadence (GpuSplatOptimizerManager.ts:380-385).
 */
typedef struct nps_config {
    int32_t max_tile_splats; /* SPLAT_MAX_TILE_SPLATS = 1024 (constants.ts:35) */
    int32_t adc_period;      /* SPLAT_ADC_PERIOD = 25 (constants.ts:30; OPT:381) */
    int32_t adc_enabled;     /* 1: nps_dispatch runs ADC every adc_period steps */
    int32_t device;          /* HIP device ordinal */
    int32_t mode;            /* NPS_MODE_REF (0) or NPS_MODE_EXT (1) */
    int32_t sh_degree;       /* EXT: spherical-harmonic degree 0..3 */
    int32_t ext_tile;        /* EXT: composite tile edge 8, 16 (default) or 32 pixels */
    int32_t ext_batch;       /* EXT: records per LDS stage 32, 64 (default) or 128 */
    int32_t reserved[8];
} nps_config;

/*
 * Modes.  REF: the reference's splat model and loop (splat_backward.wgsl,
 * splat_step.wgsl, splat_adc.wgsl): 16-float splats with RGB colour and shape
 * parameters, ascending-id tile lists capped at max_tile_splats, L2 colour +
 * depth loss, i32 fixed-point gradients.
 * EXT: the same loop over 3D Gaussian splats with view-dependent colour:
 *   splat (16 fp32) [x y z opacity_logit] [log_sx log_sy log_sz 0] [qw qx qy qz] [0 0 0 0]
 *   sh (48 fp32)    16 RGB coefficients of the real SH basis (degree sh_degree)
 * exp scales, sigmoid opacity, normalised quaternion, 0.3 px screen filter,
 * colour max(SH(dir) + 0.5, 0), alpha = min(0.99, opacity exp(-r^2/2)) with
 * pairs below 1/255 skipped, tile lists ordered by view depth (then id) and not
 * capped, a pixel sees a splat only inside its 3-sigma box (so results do not
 * depend on the tile size), a pixel stops before the splat that would bring its
 * transmittance below 1e-4, background 0.05, L1 loss, FP32 gradients of all
 * 64 parameters.
 * EXT has no reference function (an extension of the same loop).
 */
#define NPS_MODE_REF 0
#define NPS_MODE_EXT 1

typedef struct nps_ctx* nps_handle;

/* Fill *cfg with the reference defaults. */
int nps_config_default(nps_config* cfg);

/* new GpuSplatOptimizerManager({device, format, numSplats}) (OPT:43-125):
 * allocates splats (16N f32), grads (16N i32), Adam m/v (16N f32 each) + t,
 * grad_accum (N f32).  Splats start zeroed (all dead); write them with
 * nps_write_splats. */
int nps_create(const nps_config* cfg, int32_t num_splats, nps_handle* out);
void nps_destroy(nps_handle h);
const char* nps_last_error(nps_handle h);
int32_t nps_num_splats(nps_handle h);

/* queue.writeBuffer(splatBuffer) / readback (OPT:88-93). n must equal num_splats. */
int nps_write_splats(nps_handle h, const float* splats16n, int32_t n);
int nps_read_splats(nps_handle h, float* splats16n, int32_t n);
/* gradBuffer (OPT:95-99): i32 fixed point, 16 per splat (REF mode). */
int nps_write_grads(nps_handle h, const int32_t* grads16n, int32_t n);
int nps_read_grads(nps_handle h, int32_t* grads16n, int32_t n);
/* adamBuffer (OPT:101-106): m, v and the step counter t.  16 floats per
 * splat each in REF mode, 64 (splat record then SH) in EXT mode. */
int nps_write_adam(nps_handle h, const float* m, const float* v, float t, int32_t n);
int nps_read_adam(nps_handle h, float* m, float* v, float* t, int32_t n);
/* adcBuffer (OPT:108-112): REF: grad_accum, N floats.  EXT: 2 floats per
 * splat, the sum over views of |dL/d ndc| of the splat centre and the number
 * of views in which it was on screen. */
int nps_write_accum(nps_handle h, const float* accum, int32_t n);
int nps_read_accum(nps_handle h, float* accum, int32_t n);

/* EXT mode only (no reference counterpart): SH coefficients (48 floats per
 * splat) and the FP32 gradients of the 64 parameters per splat (16 of the
 * splat record, then 48 SH).  Gradients add up over nps_backward calls and
 * nps_step clears them. */
int nps_write_sh(nps_handle h, const float* sh48n, int32_t n);
int nps_read_sh(nps_handle h, float* sh48n, int32_t n);
int nps_write_grads_f32(nps_handle h, const float* grads64n, int32_t n);
int nps_read_grads_f32(nps_handle h, float* grads64n, int32_t n);

/* writeSplatVPMatrix(mat, invMat, blurEnabled) (OPT:281-301): column-major
 * view-projection and its inverse. */
int nps_set_camera(nps_handle h, const float vp[16], const float vp_inv[16], int32_t blur_enabled);

/* setBackwardTarget(colorView, depthView, w, h) (OPT:303-317): rgba is H*W*4
 * floats (the .rgb of an rgba8unorm texel is used), depth is H*W floats (the
 * .r of the depth texel).  Copied into the handle's device buffers. */
int nps_set_target(nps_handle h, const float* rgba, const float* depth, int32_t w, int32_t hgt);
/* Same with device pointers (copied device-to-device). */
int nps_set_target_dev(nps_handle h, const float* rgba_dev, const float* depth_dev, int32_t w, int32_t hgt);

/* dispatchWorkgroups(ceil(W/16), ceil(H/16)) of splat_backward.wgsl (OPT:372-374):
 * binning, forward composite, L2 loss seeds, reverse pass, fixed-point grads. */
int nps_backward(nps_handle h);
/* EXT mode only: binning and the forward composite of the current camera into
 * an image the size of the last target (its contents unused): nps_read_image
 * and nps_read_loss then give the image and its loss against that target.  No
 * gradients.  REF mode renders with the npf_* entry points instead. */
int nps_forward(nps_handle h);
/* EXT mode (no reference counterpart): times the composite for tile edges
 * {8, 16, 32} x records per LDS stage {32, 64, 128} on the current camera and
 * target (binning, sort, work order and composite fwd+bwd; the chain rule is not
 * run, gradients stay untouched), keeps the fastest for this handle and returns
 * it; ms9[3 * tile_index + batch_index] receives the times (may be NULL). */
int nps_autotune(nps_handle h, int32_t* tile, int32_t* batch, float* ms9);
/* splat_step.wgsl (OPT:376-378): fused Adam; increments t. */
int nps_step(nps_handle h);
/* splat_adc.wgsl (OPT:380-385): clone/split into dead slots. */
int nps_adc(nps_handle h);
/* EXT mode: densification parameters (defaults 2e-4, 1.0, 0.005, 0.01).  No
 * reference function; nps_adc in EXT mode then runs 3D Gaussian splatting's
 * clone/split/prune, deterministic: a slot is dead if sigmoid(logit) <
 * prune_opacity; a live slot qualifies if summed |dL/d ndc| / views >=
 * grad_threshold; the k-th qualifier (ascending index) takes the k-th dead slot;
 * split (largest scale > percent_dense * scene_extent) into two children (one
 * in the parent's slot) at centre + R (z * s), z standard normal from a counter
 * hash seeded by the ADC call count, log-scales - ln 1.6; otherwise clone.
 * Children get zero Adam state, unfilled dead slots are parked at logit -1e4,
 * accumulators are reset.  stats[4..8] = qualifiers, children, splits, clones,
 * dead slots. */
int nps_set_adc_ext(nps_handle h, float grad_threshold, float scene_extent, float prune_opacity, float percent_dense);
/* Learned densification support (EXT mode, no reference counterpart), used
 * around nps_adc by a caller that proposes the children with the MLP (nml_*):
 * nps_adc_features - 32 features per parent the next nps_adc will densify, in
 *   its placement order (first max_rows rows copied; *nq = qualifiers): mean,
 *   world covariance (xx xy xz yy yz zz), opacity, DC colour, mean and summed
 *   |dL/d ndc| and view count, log-scales, largest scale, largest scale / split
 *   scale, split flag, opacity logit, distance to the current camera and its
 *   inverse, and m / (sqrt(v) + 1e-12) of the mean, log-scales and opacity
 *   logit (the direction of the parent's next Adam steps);
 * nps_adc_children - (parent slot, dead slot) of each placement of the last
 *   nps_adc (*count = min(qualifiers, dead slots));
 * nps_adc_apply_deltas - adds 14 deltas per placement (child in the parent
 *   slot: position xyz, log-scale xyz, opacity logit; then the child in the dead
 *   slot) to the children of the last nps_adc;
 * nps_snapshot - restore = 0 saves splats, SH, Adam m/v, Adam's step counter
 *   and bias corrections, the densification accumulators and the dispatch
 *   cadence's step count; restore = 1 puts them back (training steps taken in
 *   between are undone).  The gradient buffer is not saved: between steps it
 *   holds zeros. */
int nps_adc_features(nps_handle h, float* feats32, int32_t max_rows, int32_t* nq);
int nps_adc_children(nps_handle h, int32_t* parents, int32_t* children, int32_t max_rows, int32_t* count);
int nps_adc_apply_deltas(nps_handle h, const float* deltas14, int32_t rows);
int nps_snapshot(nps_handle h, int32_t restore);
/* the same on the device: the features stay in a handle buffer valid until the
 * next nps_adc_features call; the deltas are a device array (e.g. nml_forward_dev's output) */
int nps_adc_features_dev(nps_handle h, const float** feats32_dev, int32_t* nq);
int nps_adc_apply_deltas_dev(nps_handle h, const float* deltas14_dev, int32_t rows);
/* GpuSplatOptimizerManager.dispatch (OPT:365-388): backward, step, and ADC on
 * every adc_period-th call.  Returns NPS_ERR_STATE when no target is set
 * (the reference returns silently, OPT:366). */
int nps_dispatch(nps_handle h);
int nps_sync(nps_handle h);

/* Outputs of the last nps_backward.  The reference never materialises them
 * (splat_backward.wgsl:230-248); exposed for parity checks and metrics.
 * c_pred: H*W*3, d_pred: H*W (either may be NULL; EXT mode has no depth
 * output and fills d_pred with zeros). */
int nps_read_image(nps_handle h, float* c_pred, float* d_pred);
/* loss[0] = sum over pixels of |C-tgt|^2 + fg*(D-tgt_depth)^2 (the loss whose
 * gradient the reference seeds, BW:250-257); loss[1] = sum |C-tgt| (L1). */
int nps_read_loss(nps_handle h, double loss[2]);
/* Tile lists of the last nps_backward: ranges[2*T] = [start, end) into the
 * sorted list ids[M] (splat ids, ascending within a tile in REF mode; by view
 * depth, then id, in EXT mode, where full_counts are the range lengths).  Each tile's range
 * holds at least its first min(full, max_tile_splats) ids -- the ones the
 * composite uses (splat_backward.wgsl:199-210); ids of later id blocks that
 * cannot be among them are not emitted.  full_counts[T] = number of splats whose
 * bounding box overlaps the tile (uncapped).  Pass NULL to query sizes in
 * *num_tiles / *num_ids. */
int nps_read_tiles(nps_handle h, uint32_t* ranges, uint32_t* ids, uint32_t* full_counts, int32_t* num_tiles,
                   int64_t* num_ids);
/* Counters of the last backward/ADC: stats[0]=keys, [1]=tiles over the cap,
 * [2]=(pixel, splat) pairs composited, [3]=pairs taken by the reverse pass
 * (REF: alpha >= 0.001; EXT: composited), [4]=ADC qualifiers, [5]=ADC
 * children, [6]=ADC splits, [7]=ADC clones, [8]=ADC cascade rounds run (REF;
 * in groups of 4, the last group's final round changing nothing) or dead slots
 * (EXT), [9]=wave-entries walked by the reverse pass, [10]=forward wave-entries
 * that evaluate alphas (REF: pow/exp; EXT: entries that can reach the wave's 8x8
 * block), [11]=0. */
int nps_read_stats(nps_handle h, int64_t stats[12]);
/* EXT mode: per-tile work of the last nps_backward/nps_forward, 2 per tile:
 * entries of the tile list the forward staged before every pixel had stopped,
 * and the reverse pass's extent (last entry any pixel composited + 1).  No
 * reference counterpart (the cost model of the composite's work order). */
int nps_read_tile_work(nps_handle h, uint32_t* work2t);
/* Per-phase GPU time (ms, HIP events on the handle's stream) of the last
 * backward/step/adc when timing is enabled: [0]=preprocess+keys, [1]=sort,
 * [2]=ranges, [3]=composite fwd+bwd, [4]=step, [5]=adc, [6]=whole backward,
 * [7]=EXT per-splat chain rule. */
int nps_enable_timing(nps_handle h, int32_t on);
int nps_read_timing(nps_handle h, float ms[8]);

/*
 * new GpuSplatForwardPipelineManager({device, numSplats, splatBuffer})
 * (GpuSplatForwardPipelineManager.ts:17-92): borrows the optimizer handle's
 * splat buffer.  Instanced-quad render of splat_forward.wgsl, reverse index
 * order, "over" blending into colour and encoded-depth targets.  After
 * nps_destroy(splats_owner) every npf_* call returns NPS_ERR_STATE;
 * npf_destroy still frees the handle.
 */
typedef struct npf_ctx* npf_handle;
int npf_create(nps_handle splats_owner, npf_handle* out);
void npf_destroy(npf_handle f);
/* writeVPMatrix(mat) (FWDM:94-102) */
int npf_write_vp(npf_handle f, const float vp[16]);
/* setTarget(colorView, depthView, w, h) (FWDM:104-112): the targets live in
 * the handle (H*W*4 floats each). */
int npf_set_target(npf_handle f, int32_t w, int32_t hgt);
/* dispatch(encoder, clear) (FWDM:114-137).  quantize != 0 rounds every
 * channel to k/255 after each blend, as an rgba8unorm target does. */
int npf_dispatch(npf_handle f, int32_t clear, int32_t quantize);
int npf_read_target(npf_handle f, float* color, float* depth);
int npf_write_target(npf_handle f, const float* color, const float* depth);

/*
 * Multi-view training step over several EXT scenes (no reference counterpart:
 * the reference trains one view per frame, OPT:365-388).  A step composites
 * every registered view of every scene -- a scene's gradients add up over its
 * views in a fixed order (most pixels first, then registration order) -- then
 * runs Adam once per scene and densification on each scene's nps_config
 * cadence.  Scenes must be EXT handles on one device.
 *   mode 0: view after view, each with its own launches and key-count
 *           readback (the per-view path of nps_backward);
 *   mode 1: the whole step -- views spread over `lanes` concurrent graph
 *           branches, then the Adam updates -- as one hipGraph, captured on
 *           the first call after the view set changes and replayed after
 *           that.  Key buffers are sized from a measurement at capture time
 *           (+25 %); a view that outgrows them makes that step fall back to
 *           mode 0 and the next one re-capture.
 * Both modes give bit-identical results.
 */
/* After nps_destroy of any of its scenes a step refuses every call
 * (NPS_ERR_STATE) but npt_destroy. */
typedef struct npt_ctx* npt_handle;
int npt_create(nps_handle* scenes, int32_t n_scenes, int32_t lanes, npt_handle* out);
void npt_destroy(npt_handle t);
const char* npt_last_error(npt_handle t);
/* rgba_dev: the view's target, H*W*4 floats on the device (kept by pointer). */
int npt_add_view(npt_handle t, int32_t scene, const float vp[16], const float vp_inv[16], const float* rgba_dev,
                 int32_t w, int32_t hgt);
/* One step of every view (fwd+bwd) and Adam per scene.  mode 1: the whole step
 * as one captured graph over the lanes; mode 0: view by view, each launched on
 * its scene's stream after its key count is read back (views of different
 * scenes overlap on the GPU); mode 2: view by view, each completed before the
 * next is launched (strictly sequential).  Same results in every mode. */
int npt_step(npt_handle t, int32_t mode);
/* Autotunes in two stages: (1) as nps_autotune, one view of every (scene,
 * resolution) class alone, its choice for the whole class; (2) in context: the
 * step's views captured as one graph (no chain rule, no Adam, so the scenes are
 * untouched) for {the scenes' own settings, the per-class choices} x {all
 * lanes, half of them}, each replayed (best of 3); the fastest is kept.  Views
 * start with their scene's setting.  tile_batch2v[2 v], [2 v + 1]: the choice
 * of view v. */
int npt_autotune(npt_handle t);
int npt_read_tune(npt_handle t, int32_t* tile_batch2v);
/* Losses of every view in the last step: loss2v[2 v] = L2, loss2v[2 v + 1] = L1. */
int npt_read_losses(npt_handle t, double* loss2v);
/* info[0] = nodes of the last captured graph (a step graph above 8,192 nodes,
 * the largest tested, is refused with NPS_ERR_CAPACITY: use mode 0), info[1] =
 * lanes the views are dealt to; tune_ms[4] = npt_autotune's in-context replay
 * times of {own settings, per-class choices} x {all lanes, half} (-1: not run). */
int npt_graph_info(npt_handle t, int64_t info[2], float tune_ms[4]);

/*
 * new GpuBlurPipelineManager(device) (src/lib/gpu/GpuBlurPipelineManager.ts:9-49).
 * blur(encoder, srcView, dstView, tempView, w, h, radius, sigma, isSrgb)
 * (GpuBlurPipelineManager.ts:51-99, blur.wgsl:13-48): separable Gaussian over
 * i in [-radius, radius], weights exp(-i^2 / (2 sigma^2)), clamp to edge.  The
 * horizontal pass decodes sRGB when is_srgb and writes an 8-bit linear temp,
 * the vertical pass encodes sRGB when is_srgb and writes the 8-bit output
 * (alpha 1).  Images are H*W*4 floats holding rgba8unorm texel values (k/255).
 * The caller of the reference uses radius 16, sigma 8 (ViewerState.svelte.ts:28,
 * GpuRunner.svelte.ts:644-663): sRGB for colour, linear for depth.
 */
typedef struct npb_ctx* npb_handle;
int npb_create(int32_t device, npb_handle* out);
void npb_destroy(npb_handle b);
const char* npb_last_error(npb_handle b);
/* host pointers */
int npb_blur(npb_handle b, const float* src_rgba, float* dst_rgba, int32_t w, int32_t hgt, int32_t radius,
             float sigma, int32_t is_srgb);
/* device pointers on the handle's GPU (the output can feed nps_set_target_dev) */
int npb_blur_dev(npb_handle b, const float* src_dev, float* dst_dev, int32_t w, int32_t hgt, int32_t radius,
                 float sigma, int32_t is_srgb);
/* npb_blur_dev for callers that cannot allocate device memory (code under
 * hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and the same
 * output bits as npb_blur_dev; its 8-bit temp lives in a fixed device array of
 * 2,097,152 pixels (8 MiB per GPU, 1920x1080 fits) set up by npb_create, so it
 * calls no hipMalloc/hipFree.  w*hgt above that limit returns NPS_ERR_ARG.
 * Calls on one GPU share the array: calls that overlap on one GPU race. */
int npb_blur_dev_noalloc(npb_handle b, const float* src_dev, float* dst_dev, int32_t w, int32_t hgt,
                         int32_t radius, float sigma, int32_t is_srgb);

/* ---------------------------------------------------------------- Bezier
 * Stroke optimizer: replaces GpuBezierOptimizerManager
 * (src/lib/gpu/GpuBezierOptimizerManager.ts:17-406) and its three compute
 * pipelines (bezier_backward.wgsl, bezier_step.wgsl, bezier_adc.wgsl).
 * Curves: 20 floats each (p0.xyz width | p1.xyz softness | p2.xyz pad |
 * p3.xyz pad | rgba, bezier_backward.wgsl:4-10).  Gradients: 18 i32 per curve
 * (positions, rgb, a, width, softness; fixed point 1e4 / 1e5).  Textures are
 * H*W float arrays: rgba (4 floats per pixel) for target, background and
 * normal, one float per pixel for the two depths.
 * Defined where the reference is not: loss_accum is an int64 per curve in
 * 2^-32 fixed point (the reference's non-atomic float += races), adam.t is a
 * host counter, and pixel_loss and the densification's pixel grid have the
 * size of the backward target (the reference fixes 512^2 and 256^2). */
typedef struct nbz_ctx* nbz_handle;
/* new GpuBezierOptimizerManager({device, numBeziers}) (:46-55); curves start at
 * zero until nbz_write_beziers (the reference draws random squiggles, :59-95). */
int nbz_create(int32_t device, int32_t num_beziers, nbz_handle* out);
void nbz_destroy(nbz_handle h);
const char* nbz_last_error(nbz_handle h);
int nbz_write_beziers(nbz_handle h, const float* curves20n, int32_t n);   /* bezierBuffer */
int nbz_read_beziers(nbz_handle h, float* curves20n, int32_t n);
int nbz_write_grads(nbz_handle h, const int32_t* grads18n, int32_t n);   /* gradBuffer */
int nbz_read_grads(nbz_handle h, int32_t* grads18n, int32_t n);
/* adamBuffer: m, v (18 per curve), t */
int nbz_write_adam(nbz_handle h, const float* m, const float* v, float t, int32_t n);
int nbz_read_adam(nbz_handle h, float* m, float* v, float* t, int32_t n);
/* adcBuffer: grad_accum (float) and loss_accum (int64, 2^-32) per curve */
int nbz_write_adc(nbz_handle h, const float* grad_accum, const int64_t* loss_acc, int32_t n);
int nbz_read_adc(nbz_handle h, float* grad_accum, int64_t* loss_acc, int32_t n);
/* pixel loss map of the current target size (W*H i32, fixed point 1e4) */
int nbz_write_pixel_loss(nbz_handle h, const int32_t* pl);
int nbz_read_pixel_loss(nbz_handle h, int32_t* pl);
int nbz_write_vp_matrix(nbz_handle h, const float vp[16]);         /* writeVPMatrix (:269-277) */
int nbz_write_vp_inv_matrix(nbz_handle h, const float vp_inv[16]); /* writeVPInvMatrix (:279-288) */
int nbz_write_mode(nbz_handle h, float mode);                      /* writeMode (:290-296) */
int nbz_write_bg_penalty(nbz_handle h, float weight);              /* writeBgPenalty (:298-307) */
int nbz_write_max_width(nbz_handle h, float max_width);            /* writeMaxWidth (:309-316) */
int nbz_write_kill_thresholds(nbz_handle h, float alpha, float width); /* writeKillThresholds (:318-326) */
int nbz_set_adc_period(nbz_handle h, int32_t period);              /* setAdcPeriod (:328-330) */
int nbz_write_no_kill(nbz_handle h, int32_t no_kill);              /* writeNoKill (:332-339) */
/* setBackwardTarget (:341-368): host images are copied; the _dev variant keeps
 * the device pointers, which must stay valid while the target is in use. */
int nbz_set_backward_target(nbz_handle h, const float* target_rgba, const float* target_depth, const float* bg_rgba,
                            const float* bg_depth, const float* normal_rgba, int32_t w, int32_t hgt);
int nbz_set_backward_target_dev(nbz_handle h, const float* target_rgba, const float* target_depth,
                                const float* bg_rgba, const float* bg_depth, const float* normal_rgba, int32_t w,
                                int32_t hgt);
/* dispatch (:370-406): backward, step, and every adc_period-th call the ADC */
int nbz_dispatch(nbz_handle h);
/* the three passes on their own */
int nbz_backward(nbz_handle h);
int nbz_step(nbz_handle h);
int nbz_adc(nbz_handle h);
/* image only (no gradients): the backward's forward half */
int nbz_forward(nbz_handle h);
int nbz_sync(nbz_handle h);
/* Outputs of the last nbz_backward / nbz_forward: colour (H*W*3), encoded
 * depth (H*W) and final transmittance (H*W); loss = sum |C - target|^2 and
 * sum |C - target| over the colour channels. */
int nbz_read_image(nbz_handle h, float* color, float* depth, float* trans);
int nbz_read_loss(nbz_handle h, double loss[2]);
/* tile lists of the last backward (ascending curve ids, first 1024 per tile) */
int nbz_read_tiles(nbz_handle h, uint32_t* ranges, uint32_t* ids, int32_t* num_tiles, int64_t* num_keys);
/* [0] keys, [1] tiles over the cap, [2] pairs, [3] reverse pairs, [4] reverse
 * wave-entries, [5] forward wave-entries that evaluate the distance loop (the
 * rest are provably alpha 0), [6..11] last ADC: dead slots found, killed,
 * split/clone pass 1, spawned, split/clone pass 3, dead slots left, [12]
 * reverse-recompute wave-entries that evaluate the distance loop */
int nbz_read_stats(nbz_handle h, int64_t stats[16]);
/* ms of [0] preprocess+cut, [1] keys, [2] sort+ranges, [3] composite, [4] step,
 * [5] adc, [6] backward total of the last calls */
int nbz_enable_timing(nbz_handle h, int32_t on);
int nbz_read_timing(nbz_handle h, float ms[8]);
/* Device buffers on the handle's GPU (targets kept resident for
 * nbz_set_backward_target_dev), as nps_dev_alloc / nps_dev_free /
 * nps_dev_upload.  No reference counterpart (WebGPU's createBuffer). */
int nbz_dev_alloc(nbz_handle h, int64_t bytes, void** out);
int nbz_dev_free(nbz_handle h, void* p);
int nbz_dev_upload(nbz_handle h, void* dst, const void* src, int64_t bytes);

/*
 * new GpuBezierForwardPipelineManager({device, numBeziers, bezierBuffer,
 * brushTexture}) (GpuBezierForwardPipelineManager.ts:17-98): the display raster
 * of bezier_forward.wgsl, drawn from the optimizer handle's curve buffer (the
 * reference's manager borrows bezierBuffer).  brush_rgba8: bh rows of bw rgba8
 * texels (the brush texture; only its red channel is sampled), copied at
 * creation.  Errors are reported through nbz_last_error(owner).  After
 * nbz_destroy(curves_owner) every nbf_* call returns NPS_ERR_STATE; nbf_destroy
 * still frees the handle.  A target of a new size starts cleared to 0.
 */
typedef struct nbf_ctx* nbf_handle;
int nbf_create(nbz_handle curves_owner, const uint8_t* brush_rgba8, int32_t bw, int32_t bh, nbf_handle* out);
void nbf_destroy(nbf_handle f);
/* writeVPMatrix(mat) (:100-108) */
int nbf_write_vp(nbf_handle f, const float vp[16]);
/* setTarget(targetView, width, height) (:110-129): the rgba8 target lives in the
 * handle (H*W*4 floats). */
int nbf_set_target(nbf_handle f, int32_t w, int32_t hgt);
/* dispatch(encoder, clear = true) (:131-148): clear sets the target to 0 first;
 * quantize != 0 rounds every channel to k/255 after each blend, as the
 * rgba8unorm target does. */
int nbf_dispatch(nbf_handle f, int32_t clear, int32_t quantize);
int nbf_read_target(nbf_handle f, float* rgba);
int nbf_write_target(nbf_handle f, const float* rgba);

/* ------------------------------------------------------ densification MLP
 * No reference counterpart (the reference's densification is heuristic only).
 * MLP 32 -> 256 -> 256 -> 256 -> 14 with GELU (tanh form) proposing the two
 * children of a densified parent as deltas (3 position, 3 log-scale, 1 opacity
 * logit each) on top of the heuristic children.  BF16 matrix-core products with
 * FP32 accumulation; FP32 master weights and Adam (beta 0.9 / 0.999, eps 1e-8).
 * Parameters (nml_num_params() = 143,630 floats), layer by layer: W [out][in]
 * row-major, then b [out].  Inputs are standardised with (in_mean, in_std) and
 * outputs de-standardised with (out_mean, out_std) (nml_set_norm; default 0/1);
 * the training loss is the MSE of the standardised outputs. */
typedef struct nml_ctx* nml_handle;
int nml_create(int32_t device, int32_t max_batch, nml_handle* out);
void nml_destroy(nml_handle h);
const char* nml_last_error(nml_handle h);
int32_t nml_num_params(void);
int nml_write_params(nml_handle h, const float* params);  /* also resets the Adam state */
int nml_read_params(nml_handle h, float* params);
int nml_set_norm(nml_handle h, const float* in_mean32, const float* in_std32, const float* out_mean14,
                 const float* out_std14);
/* x: n x 32 features, y: n x 14 deltas (host or device pointers) */
int nml_forward(nml_handle h, const float* x, int32_t n, float* y);
int nml_forward_dev(nml_handle h, const float* x_dev, int32_t n, float* y_dev);  /* returns when y_dev is written */
/* gradient of the standardised MSE over one batch (n <= max_batch), host pointers */
int nml_grads(nml_handle h, const float* x, const float* t14, int32_t n, float* grads, double* loss);
/* one Adam step on one batch (device or host pointers) */
int nml_train_dev(nml_handle h, const float* x_dev, const float* t14_dev, int32_t n, float lr, double* loss);
int nml_train(nml_handle h, const float* x, const float* t14, int32_t n, float lr, double* loss);
/* ms (HIP events on the handle's stream) of [0] the last nml_forward_dev launch,
 * [1] the last training step (input through Adam), while timing is on */
int nml_enable_timing(nml_handle h, int32_t on);
int nml_read_timing(nml_handle h, float ms[2]);
/* device buffers on the handle's GPU, as nps_dev_alloc / nps_dev_free / nps_dev_upload */
int nml_dev_alloc(nml_handle h, int64_t bytes, void** out);
int nml_dev_free(nml_handle h, void* p);
int nml_dev_upload(nml_handle h, void* dst, const void* src, int64_t bytes);

/* The viewer's debug display compositor (src/lib/gpu/splat_render.wgsl:24-181,
 * drawn by GpuRunner as the final pass).  textures[8]: host rgba float images in
 * the shader's binding order (target, splat view, target depth, target edges,
 * edge-layer Bezier view, base-colour Bezier view, colour Bezier view, path
 * trace), dims[2 k], dims[2 k + 1] their width and height; enabled[3]: the
 * edge, base-colour and colour Bezier layers (on above 0.5, the uniforms).
 * Writes the W x H canvas (rgba floats), each pixel shaded at its centre.
 * Texel reads clamp to the texture (WGSL leaves an out-of-range read
 * implementation-defined).  Display only: allocates per call. */
int npd_compose(int32_t device, const float* const* textures, const int32_t* dims, const float enabled[3],
                float* canvas, int32_t W, int32_t H);

/* ---------------------------------------------------------------- path trace
 * Progressive path-traced target generator: replaces GpuPathTracePipelineManager
 * (src/lib/gpu/GpuPathTracePipelineManager.ts:8-248) with its BVH builder
 * (src/lib/gpu/bvh.ts:144-223) and its two compute pipelines (path_trace.wgsl,
 * path_trace_resolve.wgsl).  Vertices: 12 floats each [px py pz, nx ny nz,
 * r g b a, u v] (alpha < 0.5 = mirror, else diffuse with albedo rgb); indices:
 * 3 per triangle.  The environment texture (envTexture of the constructor) is
 * ew x eh rgba floats (rgb used; HDR values allowed), sampled equirectangular,
 * Z up, with the sampler restated as FP32 bilinear filtering (repeat in u,
 * clamp in v).  The output texture is W x H rgba floats holding rgba8unorm
 * values (k/255), ready for nps_set_target / npb_blur.
 * The reference's methods return silently when unbound; these return
 * NPS_ERR_STATE (no meshes, no output size, no environment). */
typedef struct npp_ctx* npp_handle;
int npp_create(int32_t device, npp_handle* out);
void npp_destroy(npp_handle p);
const char* npp_last_error(npp_handle p);
/* envTexture (GpuPathTracePipelineManager.ts:42-45) */
int npp_set_env(npp_handle p, const float* rgba, int32_t ew, int32_t eh);
/* setMeshes (GpuPathTracePipelineManager.ts:101-150): the meshes already
 * concatenated (indices offset by the vertex counts before them); builds the
 * SAH BVH on the host (8 bins, at most 4 triangles per leaf, bvh.ts) */
int npp_set_meshes(npp_handle p, const float* vertices, int32_t nverts, const uint32_t* indices, int32_t ntris);
/* setOutputSize (:152-172): a new size clears the accumulator and the frame count */
int npp_set_output_size(npp_handle p, int32_t w, int32_t hgt);
/* reset (:174-180): clears the accumulator and the frame count */
int npp_reset(npp_handle p);
/* writeInvViewProjMat (:182-184): column-major inverse view-projection */
int npp_write_inv_vp(npp_handle p, const float inv_vp[16]);
/* dispatch (:221-247): one sample per pixel added to the accumulator, then the
 * resolve into the output texture; the frame count advances */
int npp_dispatch(npp_handle p);
/* the output texture (W*H*4 floats), the accumulator (W*H*4: radiance sums,
 * sample count), and the output's device address (valid until the next
 * npp_set_output_size; for nps_set_target_dev / npb_blur_dev) */
int npp_read_output(npp_handle p, float* rgba);
int npp_read_accum(npp_handle p, float* accum);
int npp_output_dev(npp_handle p, const float** rgba_dev);
/* the BVH as built: nodes (8 words each: min xyz, data0, max xyz, data1, f32
 * bits / u32) and the reordered triangle indices; counts[0] = nodes,
 * counts[1] = triangles.  Null arrays: counts only. */
int npp_read_bvh(npp_handle p, uint32_t* nodes, uint32_t* tris, int32_t counts[2]);
/* Equirectangular environment background of the raster target (envmap.wgsl,
 * drawn first in GpuRunner's mesh passes): W x H rgba8 values (k/255) of
 * Reinhard(4 env(dir)) per pixel centre, on the handle's environment. */
int npp_env_background(npp_handle p, const float inv_vp[16], float* rgba, int32_t w, int32_t hgt);

/* Mesh raster target: replaces GpuMeshRenderPipelineManager's draw
 * (src/lib/gpu/GpuMeshRenderPipelineManager.ts:246-280, mesh.wgsl:33-74) over
 * the envmap background, on the handle's meshes (npp_set_meshes) with the
 * handle's environment as the matcap texture (GpuRunner passes the same
 * texture to both, GpuRunner.svelte.ts:163-171).  Rasterisation is restated per
 * pixel centre: the closest front-facing (counter-clockwise) triangle along the
 * centre's ray, its varyings blended with the hit's barycentric weights (the
 * perspective-correct weights there); coverage on shared edges and hardware
 * interpolation rounding are parity unpinned.  Writes W x H rgba8 values
 * (k/255): colour = Reinhard(4 c) with c the normal colour (shading < 0.5) or
 * matcap x vertex colour, and depth (d, d, d, 1), d the blended
 * 1 - 0.1 / max(w, 0.1), or the envmap colour and depth 1 where no triangle is.
 * The ground mesh's PBR variant (frag_pbr) needs texture assets the reference
 * does not ship and is not built. */
int npp_raster(npp_handle p, const float vp[16], const float view[16], const float inv_vp[16], float shading_mode,
               float* color, float* depth, int32_t w, int32_t hgt);

/* Device buffers on the handle's GPU for callers without a GPU allocator of
 * their own (Deno, ctypes), e.g. to keep several targets resident in HBM for
 * nps_set_target_dev.  No reference counterpart (WebGPU's createBuffer). */
int nps_dev_alloc(nps_handle h, int64_t bytes, void** out);
int nps_dev_free(nps_handle h, void* p);
int nps_dev_upload(nps_handle h, void* dst, const void* src, int64_t bytes);

/* Deterministic transcendentals used by the kernels, evaluated on the GPU
 * over x[n] (and y[n] for pow): op 0=exp2 1=log2 2=pow 3=exp 4=log 5=sin,
 * 6=float->int32 conversion of the gradient path (result bits in out), 7=cos,
 * 8=atan2(x, y) (x the ordinate), 9=asin.
 * For tests (bit-identity with the CPU oracle). */
int nps_debug_math(int32_t device, int32_t op, const float* x, const float* y, float* out, int32_t n);

/* The EXT depth order on the GPU for tests: keys[n] = depth_bits << 32 | id
 * (depth_bits 0xFFFFFFFF: not drawn) sorted stably on the depth bits, by the
 * 32-bit pass sort (narrow = 0) or the two-pass range-compressed sort of
 * speculative views (narrow = 1; the per-256-key minima of the drawn depths
 * stand in for the preprocess's).  flags_out: [0] any overflow, [1] depth range
 * too wide for the narrow sort (the keys are then not in depth order); n < 2^31.
 * No reference counterpart (EXT's depth-ordered lists; the reference's lists
 * are in id order). */
int nps_debug_depth_sort(int32_t device, const uint64_t* keys, uint32_t n, int32_t narrow, uint64_t* out,
                         uint32_t flags_out[2]);

#ifdef __cplusplus
}
#endif
#endif
