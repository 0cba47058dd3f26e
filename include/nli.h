/*
 * nli.h -- C-ABI of libnli, the MI355X (gfx950) implementation of the inner
 * inpainting loop of Tom-Zheng/non-local-inpainting: PatchMatch NNF search and
 * the non-local image updates (means, medians, Poisson with its conjugate
 * gradient, depth), plus the pyramid, mask and confidence steps that surround
 * them and the whole multiscale process().
 *
 * Conventions (the reference containers, image.hpp:481-485, mask.h):
 *   images  float32 interleaved HWC, index C*(W*y + x) + c
 *   masks   uint8 0/1, index W*y + x (1 = masked / hole / in the set)
 *   NNF     int32 pairs (x, y), index 2*(W*y + x); (-1, -1) = no entry
 * Unless a name ends in _dev, pointer arguments are HOST pointers: the call
 * copies its inputs to the device, computes, and copies its outputs back before
 * returning (the per-call semantics of the reference's host API).  _dev entry
 * points take device pointers and run on the context's stream.
 *
 * Every function returns NLI_OK (0) or a negative NLI_E_* code and never throws;
 * nli_last_error() gives the message.  Functions that fail where the reference
 * throws std::runtime_error report NLI_E_ARGUMENT with the reference's message.
 * A context is used by one host thread at a time.
 */
#ifndef NLI_H_
#define NLI_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLI_OK 0
#define NLI_E_ARGUMENT (-1)     /* invalid argument; reference std::runtime_error cases */
#define NLI_E_DEVICE (-2)       /* HIP runtime error */
#define NLI_E_EMPTY_SOURCE (-3) /* no complete source patch (image_inpainting.cpp:382-384) */
#define NLI_E_INTERNAL (-4)

typedef struct nli_ctx nli_ctx;

/* patch distance kinds (main.cpp:245-263 pairs them with the updaters) */
#define NLI_DIST_L2 0          /* L2NormPatchDistance      l2_norm_patch_distance.h     */
#define NLI_DIST_L1 1          /* L1NormPatchDistance      l1_norm_patch_distance.h     */
#define NLI_DIST_L2COMBINED 2  /* L2CombinedPatchDistance  l2_combined_patch_distance.h */
#define NLI_DIST_L2RGBD 3      /* L2RGBDPatchDistance      l2_rgbd_patch_distance.h     */

/* image updating kinds */
#define NLI_UPD_MEANS 0        /* PatchNonLocalMeans    patch_non_local_means.h   */
#define NLI_UPD_MEDIANS 1      /* PatchNonLocalMedians  patch_non_local_medians.h */
#define NLI_UPD_POISSON 2      /* PatchNonLocalPoisson  patch_non_local_poisson.h */
#define NLI_UPD_DEPTH 3        /* PatchNonLocalDepth    patch_non_local_depth.h   */

/* coarsest-level initialisation, ImageInpainting::InitType order (image_inpainting.h:43) */
#define NLI_INIT_BLACK 0
#define NLI_INIT_AVG 1
#define NLI_INIT_NONE 2
#define NLI_INIT_POISSON 3

/* CG arithmetic */
#define NLI_CG_FP64 0          /* the reference recurrence in double precision */
#define NLI_CG_MIXED 1         /* FP32 vectors, FP64 dot products, periodic FP64 residual */

typedef struct {
    int kind;        /* NLI_DIST_* */
    int patch;       /* square patch side, 1..15 */
    float sigma;     /* Gaussian sigma of the patch weights */
    float lambda;    /* L2Combined lambda / L2RGBD lambda_rgb */
    float lambda_d;  /* L2RGBD lambda_d */
} nli_dist_params;

typedef struct {
    int iterations;  /* sweeps (PatchMatch iteration_count, patch_match.h:41-47) */
    int shots;       /* random shots limit */
    int window;      /* search window size, -1 = max(W, H) */
    int first_step;  /* jump-flood step of the first sweep, 0 = automatic */
} nli_pm_params;

typedef struct {
    int kind;        /* NLI_UPD_* */
    int patch;
    float sigma;
    float lambda;    /* Poisson lambda / Depth lambda_rgb */
    float lambda_d;  /* Depth lambda_d */
    float cg_tol;    /* Poisson: stop when |r|^2 < cg_tol |r0|^2 (Depth uses 1e-6) */
    int cg_limit;    /* Poisson: iteration limit (Depth uses 1000) */
    int cg_precision;      /* NLI_CG_FP64 or NLI_CG_MIXED */
    int cg_refine_every;   /* mixed: FP64 residual replacement period (0 = never) */
} nli_upd_params;

typedef struct {
    int channels;            /* channels solved (3 RGB + 1 depth for Depth) */
    int cg_iterations[4];
    double cg_nr0[4];        /* |r0|^2 per channel */
    double cg_nr[4];         /* final |r|^2 per channel */
} nli_cg_stats;

typedef struct {
    int method;      /* NLI_UPD_*: selects the (distance, updater) pair as main.cpp:245-263 */
    int patch;
    float sigma;
    float lambda;
    float lambda_rgb;
    float lambda_d;
    int iters;       /* inpainting iterations per scale */
    float tol;       /* stop a scale when the total squared update <= tol */
    int scales;
    float coarse;    /* coarsest size ratio; the subsampling rate is exp(log(coarse)/(scales-1)) */
    float conf_t;    /* confidence decay time */
    float conf_a;    /* confidence asymptotic value */
    int init;        /* NLI_INIT_* */
    int pm_iters;
    int pm_shots;
    int pm_window;
    float cg_tol;
    int cg_limit;
    int cg_precision;
    int cg_refine_every;
    uint64_t seed;
} nli_inpaint_params;

typedef struct {
    int scales;
    int iterations[32];      /* inpainting iterations run per scale, finest = 0 */
    int width[32], height[32];
    int hole_pixels[32];
    int target_pixels[32];
    int cg_iterations_total;
    double seconds_total;
    double seconds_patchmatch;
    double seconds_update;
    double seconds_other;
} nli_inpaint_stats;

/* ---- context ---- */
int nli_ctx_create(int device, nli_ctx **out);
int nli_ctx_destroy(nli_ctx *ctx);
const char *nli_last_error(const nli_ctx *ctx);
int nli_synchronize(nli_ctx *ctx);
int nli_version(void);

/* ---- device memory helpers for the _dev entry points ---- */
int nli_dev_alloc(nli_ctx *ctx, uint64_t bytes, void **out);
int nli_dev_free(nli_ctx *ctx, void *p);
int nli_copy_h2d(nli_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int nli_copy_d2h(nli_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* device-to-device copy on the context's stream (returns when it has completed) */
int nli_copy_d2d(nli_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* ---- numerics ---- */
/* replaces Gradient::calculate (gradient.h:25, gradient.cpp:20-76): out W*H*2C */
int nli_gradient(nli_ctx *ctx, const float *img, int W, int H, int C, float *out);

/* replaces DistanceTransform::calculate (distance_transform.h:30) */
int nli_distance_transform(nli_ctx *ctx, const uint8_t *mask, int W, int H, float *out);

/* replaces ImageInpainting::calculate_confidence_mask (image_inpainting.cpp:611-630) */
int nli_confidence(nli_ctx *ctx, const uint8_t *mask, int W, int H, float decay, float asym,
                   float *out);

/* replaces Sampling::_Details::get_sample_size (sampling.cpp:408-411) and
 * ImageInpainting::calculate_subsampling_rate (image_inpainting.cpp:197-204) */
int nli_sample_size(int size, float factor);
float nli_subsampling_rate(float ratio, int scales);

/* replaces Sampling::downsample_with_mask (sampling.h:46, sampling.cpp:114-202);
 * out_img sW*sH*C, out_mask sW*sH with sW = nli_sample_size(W, factor) */
int nli_downsample_with_mask(nli_ctx *ctx, const float *img, const uint8_t *mask, int W, int H,
                             int C, float factor, float threshold, int mask_val, float *out_img,
                             uint8_t *out_mask);
/* replaces Sampling::downsample(image) and Sampling::downsample(mask) (sampling.h:38-44) */
int nli_downsample(nli_ctx *ctx, const float *img, int W, int H, int C, float factor, float *out);
int nli_downsample_mask(nli_ctx *ctx, const uint8_t *mask, int W, int H, float factor,
                        float threshold, uint8_t *out);
/* replaces Sampling::upsample (sampling.h:52, sampling.cpp:205-246) */
int nli_upsample(nli_ctx *ctx, const float *img, int W, int H, int C, int W2, int H2, float *out);

/* ---- per-scale masks ---- */
/* replaces ImageInpainting::get_extended_domain (image_inpainting.cpp:566-600) */
int nli_extended_domain(nli_ctx *ctx, const uint8_t *hole, int W, int H, int patch, uint8_t *out);
/* the mask preparation of inpaint_internal (image_inpainting.cpp:359-384, erode = 1) and of
 * propagate_weights (:496-502, :543-548, erode = 0): extended domain, target, source.
 * Returns NLI_E_EMPTY_SOURCE when erode = 1 and no source pixel survives the erosion. */
int nli_scale_masks(nli_ctx *ctx, const uint8_t *hole, int W, int H, int patch, int erode,
                    uint8_t *ext, uint8_t *target, uint8_t *source);

/* the NNF scaling of ImageInpainting::propagate_weights (image_inpainting.cpp:515-529):
 * out (W*H*2) = (-1,-1) except at round(p * scale) for the lower-level hole points p with
 * an entry, which receive round(nnf(p) * scale); scale = (float)W / lW, (float)H / lH. */
int nli_scale_nnf(nli_ctx *ctx, const int32_t *lower_nnf, const uint8_t *lower_hole, int lW, int lH,
                  int W, int H, int32_t *out);

/* ---- patch distances and PatchMatch ---- */
/* APatchDistance::calculate (a_patch_distance.h:32-35) for every target point of a fixed
 * field: out[W*y+x] = distance(nnf(x,y), (x,y)) on target points, FLT_MAX elsewhere. */
int nli_patch_distance(nli_ctx *ctx, const nli_dist_params *dist, const float *src, int sW,
                       int sH, const float *tgt, int tW, int tH, int C, const int32_t *nnf,
                       const uint8_t *tmask, float *out);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * nli_patch_distance on device pointers (src, tgt, nnf, tmask, out), with the same
 * results, and with no allocation: its scratch (the source and target gradients, the
 * patch weights) is a static array of NLI_NOALLOC_SCRATCH_BYTES per device, whose
 * address nli_ctx_create looks up.  Images needing more than that (256 + 2 C (sW sH +
 * tW tH, the target term only if tgt != src) floats) are refused with NLI_E_ARGUMENT.
 * All calls on one device share that device's array: calls that overlap on one device
 * (contexts on different streams) race. */
#define NLI_NOALLOC_SCRATCH_BYTES (32u << 20)
int nli_patch_distance_dev_noalloc(nli_ctx *ctx, const nli_dist_params *dist, const float *src, int sW,
                                   int sH, const float *tgt, int tW, int tH, int C, const int32_t *nnf,
                                   const uint8_t *tmask, float *out);

/* replaces PatchMatch::calculate (patch_match.h:50-54, patch_match.cpp:57-512).
 * init_nnf = NULL is the reference's empty initial field (random initialisation).
 * out_dist (optional) receives each target's distance (FLT_MAX elsewhere).
 * Returns NLI_E_ARGUMENT on a size mismatch (the reference returns an empty field). */
int nli_patchmatch(nli_ctx *ctx, const nli_dist_params *dist, const nli_pm_params *pm,
                   const float *src, const uint8_t *smask, int sW, int sH, const float *tgt,
                   const uint8_t *tmask, int tW, int tH, int C, const int32_t *init_nnf,
                   uint64_t seed, int32_t *out_nnf, float *out_dist);
int nli_patchmatch_dev(nli_ctx *ctx, const nli_dist_params *dist, const nli_pm_params *pm,
                       const float *src, const uint8_t *smask, int sW, int sH, const float *tgt,
                       const uint8_t *tmask, int tW, int tH, int C, const int32_t *init_nnf,
                       uint64_t seed, int32_t *out_nnf, float *out_dist);

/* Tile arrangement of the PatchMatch sweep, process-wide: 0 = autotuned (the default:
 * for each distance kind, channel count, patch side, row-scan setting and size class, one
 * call times sweeps in each arrangement and keeps the fastest); forced: 1 or 2 target
 * tiles per workgroup, or 3 = one tile per workgroup with two waves per tile row (half
 * the targets per wave; runs as 1 where the row scan is on, and in lockstep batches).
 * Fields and distances do not depend on the arrangement.  No counterpart in the
 * reference. */
int nli_pm_tiles_per_block(int tiles);

/* Row scan of the PatchMatch pool sweep, process-wide: 0 = automatic (the default: on for
 * NLI_DIST_L2RGBD), 1 = on, 2 = off.  After each sweep's propagation and random search,
 * every target of a 16-target tile row repeatedly tries the shifted match of its
 * scan-side neighbour (left on even sweeps, right on odd ones) while that neighbour
 * improves, as the reference's sequential scan does along a row within one sweep
 * (patch_match.cpp:197-232).  No counterpart in the reference as a setting. */
int nli_pm_row_scan(int mode);

/* How the batch entry points (nli_inpaint_batch*) run, process-wide: 0 (the default) or 2 =
 * one per-image driver per worker thread, each on its own stream; 1 = in lockstep groups
 * where the parameters allow it (NL-Poisson / NL-Depth with a candidate-pool PatchMatch,
 * Poisson or no initialisation, every image with the same parameters): `workers` groups,
 * each stepping its images' scales and iterations together on one stream, with one launch
 * per kernel for all of them and the stopping rule applied on the device.  Results do not
 * depend on the mode.  No counterpart in the reference. */
int nli_batch_mode(int mode);

/* ---- image updating ---- */
/* replaces AImageUpdating::update (a_image_updating.h:34-39) of the four updaters:
 * image (W*H*C) is updated in place on the hole pixels (domain); orig is the image the
 * scale started from; ext the extended domain; conf may be NULL (empty confidence).
 * *total_diff receives the returned value of the reference (sum of squared changes). */
int nli_update(nli_ctx *ctx, const nli_upd_params *upd, float *image, const float *orig,
               const uint8_t *domain, const uint8_t *ext, const int32_t *nnf, const float *conf,
               int W, int H, int C, double *total_diff, nli_cg_stats *stats);
int nli_update_dev(nli_ctx *ctx, const nli_upd_params *upd, float *image, const float *orig,
                   const uint8_t *domain, const uint8_t *ext, const int32_t *nnf,
                   const float *conf, int W, int H, int C, double *total_diff,
                   nli_cg_stats *stats);

/* ---- the NL-Poisson patch energy, its gradient and the adjoint aggregation ----
 * For a fixed NNF and confidence, the NL-Poisson update (patch_non_local_poisson.cpp:
 * 36-232) minimises over the hole pixels the patch energy
 *   E(u) = sum_{p in ext} sum_{h: p+h, nnf(p)+h in the image} (double)(conf(p) g(h)) *
 *          sum_c [ lambda (u(p+h) - u(s))^2 + (1-lambda) |grad u(p+h) - grad u(s)|^2 ],
 *   s = nnf(p) + h, grad = forward differences (0 on the last column / row), lambda and
 *   1-lambda in float as the reference's coefficients take them,
 * whose terms are those the reference accumulates into a1, a2, f1, F.  Its gradient is
 * twice the residual of the reference's linear system on the hole pixels.  No
 * counterpart exists in the reference: these serve differentiation (energy gradient
 * with respect to the pixels) and adjoint checks. */
typedef struct {
    int patch;       /* odd, <= 15 */
    float sigma;     /* Gaussian sigma of the intra-patch weights */
    float lambda;
} nli_energy_params;
/* E(u); with px >= 0 only the terms that read pixel (px, py) channel pc, evaluated with
 * that value increased by delta (FP64; for finite differences).  conf may be NULL. */
int nli_patch_energy(nli_ctx *ctx, const nli_energy_params *params, const float *image,
                     const uint8_t *ext, const int32_t *nnf, const float *conf, int W, int H, int C,
                     int px, int py, int pc, double delta, double *energy);
/* dE/du for every pixel and channel (W*H*C doubles, HWC): target-side terms by gather,
 * source-side terms by scatter, NNF re-read instead of stored per-patch weights */
int nli_energy_gradient(nli_ctx *ctx, const nli_energy_params *params, const float *image,
                        const uint8_t *ext, const int32_t *nnf, const float *conf, int W, int H,
                        int C, double *grad);
/* patch extraction: for the extended-domain centres p (row-major, *n_centres of them),
 * patches[((e*P*P) + h)*C + c] = u(p+h) (side 0) or u(nnf(p)+h) (side 1), 0 where p+h
 * or nnf(p)+h leaves the image; capacity = doubles available in patches */
int nli_patch_extract(nli_ctx *ctx, int patch, const double *u, const uint8_t *ext,
                      const int32_t *nnf, int W, int H, int C, int side, double *patches,
                      uint64_t capacity, int *n_centres);
/* the adjoint of nli_patch_extract (the aggregation of patches back onto pixels):
 * side 0 gathers at each pixel the entries of the centres p = q - h, side 1 scatters to
 * nnf(p) + h; out W*H*C doubles */
int nli_adjoint_aggregate(nli_ctx *ctx, int patch, const double *patches, int n_centres,
                          const uint8_t *ext, const int32_t *nnf, int W, int H, int C, int side,
                          double *out);

/* ---- coarsest-level initialisation ---- */
/* replaces poisson_solver_separable_simplest (3rdparty/simpois/simpois.c:281-291) as
 * called by ImageInpainting::process (image_inpainting.cpp:125-153): fills the masked
 * pixels of img (W*H*C, HWC) in place. Host computation (coarsest level only). */
int nli_poisson_fill(float *img, const uint8_t *mask, int W, int H, int C);
/* the same fill on the device (img, mask device pointers), bit-identical to
 * nli_poisson_fill: one workgroup per channel, the in-place Gauss-Seidel sweeps run as
 * anti-diagonal wavefronts */
int nli_poisson_fill_dev(nli_ctx *ctx, float *img, const uint8_t *mask, int W, int H, int C);

/* ---- the whole multiscale inpainting ---- */
/* replaces ImageInpainting::process (image_inpainting.h:54, image_inpainting.cpp:68-190)
 * with the plugins of main.cpp:245-266; runs device-resident. out: W*H*C. */
int nli_inpaint(nli_ctx *ctx, const nli_inpaint_params *params, const float *in,
                const uint8_t *mask, int W, int H, int C, float *out, nli_inpaint_stats *stats);
int nli_inpaint_dev(nli_ctx *ctx, const nli_inpaint_params *params, const float *in,
                    const uint8_t *mask, int W, int H, int C, float *out,
                    nli_inpaint_stats *stats);

/* ---- kernel timing ----
 * With timing on, HIP events recorded on the context's stream bracket the launches of
 * the kernels below; nli_profile_read synchronises and returns, for one kernel, the
 * launches, their total duration, their algorithmic work (flops for PatchMatch sweeps,
 * bytes for CG iterations) and the elements processed.  `kernels` is a bitmask of
 * (1 << NLI_KERNEL_*) selecting the kernels to time (NLI_PROFILE_ALL: all of them,
 * 0: timing off); every call resets the totals.  Each timed launch adds an event pair
 * to its stream, so time only the kernels of interest in throughput runs. */
#define NLI_KERNEL_PM_SWEEP 0
#define NLI_KERNEL_PM_INIT 1
#define NLI_KERNEL_CG_ITERATION 2
#define NLI_KERNEL_POISSON_COEF 3
#define NLI_KERNEL_NL_GATHER 4
#define NLI_KERNEL_COLOR 6       /* sRGB <-> Lab conversions; work = bytes */
#define NLI_PROFILE_ALL (-1)
int nli_profile_enable(nli_ctx *ctx, int kernels);
/* as nli_profile_enable, timing one launch in `period` of each selected kernel (counted
 * in issue order over all streams), so that throughput runs carry few event pairs;
 * nli_profile_read then reports the sampled launches only */
int nli_profile_enable_sampled(nli_ctx *ctx, int kernels, int period);
int nli_profile_read(nli_ctx *ctx, int kernel, int *launches, double *total_ms, double *work,
                     double *units);

/* ---- batches of images ----
 * Inpaints n independent images (sizes may differ; all have C channels) with the same
 * parameters.  The images are taken from a shared queue by `workers` concurrent
 * device-resident drivers, each with its own stream and buffers (0 = one per hardware
 * queue, 4).  Per-image results are identical to nli_inpaint_dev with the same seed;
 * stats (optional) receives n records.  _dev: in/mask/out are device pointers. */
int nli_inpaint_batch(nli_ctx *ctx, const nli_inpaint_params *params, int n, const float *const *in,
                      const uint8_t *const *mask, const int *W, const int *H, int C, float *const *out,
                      int workers, nli_inpaint_stats *stats);
int nli_inpaint_batch_dev(nli_ctx *ctx, const nli_inpaint_params *params, int n,
                          const float *const *in, const uint8_t *const *mask, const int *W,
                          const int *H, int C, float *const *out, int workers,
                          nli_inpaint_stats *stats);
/* as nli_inpaint_batch(_dev) with one parameter record per image (params[n]): methods,
 * patch sides, scales and seeds may differ between images.  Images are run largest
 * first; results do not depend on the order. */
int nli_inpaint_batch_each(nli_ctx *ctx, const nli_inpaint_params *params, int n,
                           const float *const *in, const uint8_t *const *mask, const int *W,
                           const int *H, int C, float *const *out, int workers,
                           nli_inpaint_stats *stats);
int nli_inpaint_batch_each_dev(nli_ctx *ctx, const nli_inpaint_params *params, int n,
                               const float *const *in, const uint8_t *const *mask, const int *W,
                               const int *H, int C, float *const *out, int workers,
                               nli_inpaint_stats *stats);

/* fills params with the reference CLI defaults (main.cpp:68-82, 220-233, 254, 266) */
void nli_inpaint_params_default(nli_inpaint_params *p);

/* ---- the colour path around process() ----
 * The reference's command-line program converts its 0-255 RGB input to CIE L*a*b*
 * before ImageInpainting::process and the result back afterwards (main.cpp:117, 307).
 * Conversions run one pixel per thread with the reference's float/double operation
 * order; each of its powf calls gives the correctly rounded power (equal to glibc's
 * powf except for a few inputs in ten thousand, where the two differ by one unit in the
 * last place). */
/* replaces IOUtility::rgb_to_lab (io_utility.h:51, io_utility.cpp:268-289): for C = 3,
 * lab receives W*H*3 L*a*b* values; any other C is returned unchanged (lab receives a
 * copy of the W*H*C input), as the reference returns its argument. */
int nli_rgb_to_lab(nli_ctx *ctx, const float *rgb, int W, int H, int C, float *lab);
/* replaces IOUtility::lab_to_rgb (io_utility.h:52, io_utility.cpp:292-318): for C = 3 or
 * 4, rgb receives W*H*3 values converted from the first three channels; any other C is
 * returned unchanged (a copy of the W*H*C input).  For C = 4 the reference reads the
 * colour planes through a pointer into a local image that is released when its branch
 * ends (io_utility.cpp:303-308), so its result there is undefined; here it is the
 * conversion of channels 0-2. */
int nli_lab_to_rgb(nli_ctx *ctx, const float *lab, int W, int H, int C, float *rgb);
/* device-pointer forms of the two (return when the conversion has completed) */
int nli_rgb_to_lab_dev(nli_ctx *ctx, const float *rgb, int W, int H, int C, float *lab);
int nli_lab_to_rgb_dev(nli_ctx *ctx, const float *lab, int W, int H, int C, float *rgb);
/* main.cpp:115-309 without file I/O and argument parsing: rgb (W*H*3, 0-255) is
 * converted to Lab; depth (W*H, optional) is appended as channel 3 (IOUtility::cat,
 * io_utility.cpp:453-470); the image is inpainted as by nli_inpaint; the result's colour
 * channels are converted back into rgb_out (W*H*3) and, with depth, channel 3 goes to
 * depth_out (W*H; IOUtility::get_depth, :506-513).  mask: W*H, 1 = hole (main.cpp:126-133
 * sets it where the mask image is positive).  As main.cpp:166-203 does: params->coarse = 0
 * selects min(1.5 * patch / max distance to the hole boundary, 1), a coarsest rate
 * within 1e-5 of 1 runs one scale, and conf_t must be positive.  NLI_UPD_DEPTH needs
 * depth (main.cpp:256-257). */
int nli_inpaint_rgb(nli_ctx *ctx, const nli_inpaint_params *params, const float *rgb, const float *depth,
                    const uint8_t *mask, int W, int H, float *rgb_out, float *depth_out,
                    nli_inpaint_stats *stats);
/* the same with device pointers (depth and depth_out may be NULL) */
int nli_inpaint_rgb_dev(nli_ctx *ctx, const nli_inpaint_params *params, const float *rgb,
                        const float *depth, const uint8_t *mask, int W, int H, float *rgb_out,
                        float *depth_out, nli_inpaint_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* NLI_H_ */
