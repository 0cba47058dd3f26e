// nli_internal.h -- shared declarations of libnli (device buffers, launch
// wrappers).  Everything here operates on device pointers and a HIP stream;
// the C-ABI (nli_capi.hip) and the device-resident driver are built on it.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <memory>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace nli {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};

#define NLI_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::nli::HipError(e_, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define NLI_LAUNCH_CHECK() NLI_HIP(hipGetLastError())

// Grow-only device buffer owned by a context.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T *ensure(size_t count)
    {
        if (count == 0) count = 1;
        if (count > n) {
            release();
            NLI_HIP(hipMalloc(&p, count * sizeof(T)));
            n = count;
        }
        return p;
    }
};

// Distance kinds and updater kinds; numbering shared with the public header.
enum DistKind { DIST_L2 = 0, DIST_L1 = 1, DIST_L2COMBINED = 2, DIST_L2RGBD = 3 };
enum UpdKind { UPD_MEANS = 0, UPD_MEDIANS = 1, UPD_POISSON = 2, UPD_DEPTH = 3 };

struct DistDesc {
    int kind;
    int patch;
    float sigma;
    float lambda;
    float lambda_d;
};

// ---- kernel timing (HIP events on the launching stream), off unless enabled ----
enum TimedKernel {
    TK_PM_SWEEP = 0,    // k_pm_sweep: one PatchMatch sweep
    TK_PM_INIT = 1,     // k_pm_init
    TK_CG = 2,          // one CG iteration (phase A + phase B kernels)
    TK_POISSON_COEF = 3,
    TK_NL_GATHER = 4,   // NL-means / NL-medians gathers
    TK_PYRAMID = 5,
    TK_COLOR = 6,       // sRGB <-> Lab conversions
    TK_COUNT = 7
};
namespace prof {
// kernels: bitmask over TimedKernel ids (0 = timing off); one launch in `period` of each
// kernel is timed (launches counted in issue order over all streams)
void enable(unsigned kernels, unsigned period = 1);
// brackets launches on stream s; `work` is the algorithmic flops or bytes of the
// bracketed launches, `units` the elements they process, `launches` their number (a
// bracket of several back-to-back launches of one kernel: one event pair for all)
void begin(int kernel, hipStream_t s);
void end(int kernel, hipStream_t s, double work, double units, int launches = 1);
// synchronises and returns launches, total milliseconds, total work, total units
void read(int kernel, int *launches, double *ms, double *work, double *units);
}  // namespace prof

// ---- patch weights / filters (host side, double libm as the reference) ----
void gaussian_weights(int size_x, int size_y, float sigma_x, float sigma_y, float *out);
float subsampling_rate(float ratio, int scales);
int sample_size(int size, float factor);
// coarsest-level Poisson fill of the masked pixels (host, in place, HWC)
void poisson_fill(float *img, const uint8_t *mask, int W, int H, int C);
// the same on the device (k_simpois.hip): img/mask device pointers, scratch of
// poisson_fill_scratch(W, H, C) floats; bit-identical to poisson_fill
size_t poisson_fill_scratch(int W, int H, int C);
void poisson_fill_device(float *img, const uint8_t *mask, int W, int H, int C, float *scratch, hipStream_t s);

// ---- generic device helpers (k_masks.hip) ----
// Order-preserving compaction: writes the indices i < n with flag[i] != 0 to out,
// returns the count (synchronises the stream to read it).
int compact_indices(const uint8_t *flag, int n, int *out, int *scratch_counts, int *d_total,
                    hipStream_t s);
void fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
void fill_i32(int32_t *p, int32_t v, size_t n, hipStream_t s);
void fill_f32(float *p, float v, size_t n, hipStream_t s);
void fill_f64(double *p, double v, size_t n, hipStream_t s);
// NNF fields n0 (and n1, if not null) of n pixels to (-1, -1), distances d0 (d1) to FLT_MAX
void fill_fields(int32_t *n0, float *d0, int32_t *n1, float *d1, size_t n, hipStream_t s);

// ---- gradient (k_patchmatch.hip) ----
void gradient(const float *img, int W, int H, int C, float *grad, hipStream_t s);

// ---- sRGB <-> Lab (k_color.hip) ----
// rgb: n x 3 (0-255); lab: n x out_c, out_c = 3, or 4 with depth[n] as channel 3
void rgb_to_lab(const float *rgb, const float *depth, long n, int out_c, float *lab, hipStream_t s);
// lab: n x in_c (3 or 4; channel 3 is not converted); rgb: n x 3; depth (optional,
// in_c = 4) receives channel 3
void lab_to_rgb(const float *lab, long n, int in_c, float *rgb, float *depth, hipStream_t s);

// ---- masks (k_masks.hip) ----
void extended_domain(const uint8_t *hole, int W, int H, int P, uint8_t *ext, uint8_t *scratch,
                     hipStream_t s);
// target/source masks of a scale; erode = the right/bottom erosion of inpaint_internal.
// *d_nonempty (device int) is set to 1 if the eroded source (before the margin) is non-empty.
void scale_masks(const uint8_t *ext, int W, int H, int margin, int erode, uint8_t *target,
                 uint8_t *source, int *d_nonempty, hipStream_t s);
void scale_nnf(const int32_t *lower_nnf, const int *lower_hole_pts, int n_lower, int lW, int lH,
               int uW, int uH, int32_t *upper_nnf, hipStream_t s);
void copy_into_hole(const float *src, float *dst, const int *hole_pts, int n, int C, hipStream_t s);
void distance_transform(const uint8_t *mask, int W, int H, float *out, float *g_scratch,
                        int *st_scratch, hipStream_t s);
void confidence(const uint8_t *mask, const float *dt, int W, int H, float decay, float asym,
                float *out, hipStream_t s);

// ---- pyramid (k_pyramid.hip) ----
void downsample_with_mask(const float *img, const uint8_t *mask, int W, int H, int C, float factor,
                          float threshold, int mask_val, float *out_img, uint8_t *out_mask,
                          float *scratch, hipStream_t s);
void downsample(const float *img, int W, int H, int C, float factor, float *out, float *scratch,
                hipStream_t s);
void downsample_mask(const uint8_t *mask, int W, int H, float factor, float threshold,
                     uint8_t *out, float *scratch, hipStream_t s);
void upsample(const float *in, int W, int H, int C, int W2, int H2, float *out, hipStream_t s);

// ---- PatchMatch (k_patchmatch.hip) ----
struct PMConfig {
    int iterations;
    int shots;
    int window;
    int first_step;   // jump-flood step of the first sweep; 0 = automatic
};

// A table of kernel arguments for batched launches: filled on the host in pinned memory
// and copied to the device in stream order.  The host side is refilled only after the
// previous copy from it has completed (an event after each copy), and the device side is
// overwritten only in stream order, after the kernels that read it.
struct ArgTable {
    char *host = nullptr;
    size_t cap = 0;
    DBuf<char> dev;
    hipEvent_t copied = nullptr;
    ArgTable() = default;
    ArgTable(const ArgTable &) = delete;
    ArgTable &operator=(const ArgTable &) = delete;
    ~ArgTable()
    {
        if (host) (void)hipHostFree(host);
        if (copied) (void)hipEventDestroy(copied);
    }
    // host storage for `bytes`, safe to write
    void *fill(size_t bytes)
    {
        if (copied) NLI_HIP(hipEventSynchronize(copied));
        if (bytes > cap) {
            if (host) NLI_HIP(hipHostFree(host));
            host = nullptr;
            cap = std::max<size_t>(bytes, 4096);
            NLI_HIP(hipHostMalloc((void **)&host, cap, hipHostMallocDefault));
        }
        return host;
    }
    // copies the first `bytes` to the device; returns the device address
    const void *upload(size_t bytes, hipStream_t s)
    {
        dev.ensure(bytes);
        if (!copied) NLI_HIP(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
        NLI_HIP(hipMemcpyAsync(dev.p, host, bytes, hipMemcpyHostToDevice, s));
        NLI_HIP(hipEventRecord(copied, s));
        return dev.p;
    }
};

struct PMWork {          // per-call scratch, grown on demand
    DBuf<float> sgrad, tgrad, weights, dist0, dist1;
    DBuf<int32_t> nnf1;
    DBuf<int> tiles, counts, total;
    DBuf<uint8_t> tflags;
    // cached patch weights (uploaded for this patch side and sigma)
    int w_patch = 0;
    float w_sigma = 0.0f;
    // cached list of target tiles, valid for (mask pointer, size, epoch)
    const uint8_t *tiles_mask = nullptr;
    int tiles_W = 0, tiles_H = 0, ntiles = 0, targets = 0;
    uint64_t tiles_epoch = ~0ull, epoch = 0;
};

// Tile arrangement of the candidate-pool sweep: 0 = autotuned (per process, device,
// distance kind, channels, patch side and size class), 1 or 2 tiles per block = forced.
// The arrangement does not change results.
void set_pm_tiles_per_block(int tpb);
int pm_tiles_per_block();
void set_pm_row_scan(int mode);
int pm_row_scan();
int pm_row_scan_for(int kind);

// The target mask behind a PatchMatch call changed (contents may change behind the
// same pointer): the cached tile list is rebuilt on the next call.
inline void pm_masks_changed(PMWork &w) { w.epoch++; }

// Full PatchMatch::calculate on device.  init_nnf may be null (random init).
// out_nnf receives the field (W x H x 2, (-1,-1) outside the targets); out_dist (optional)
// the distance of every target to its entry.
void patchmatch(const DistDesc &d, const PMConfig &cfg, const float *src, const uint8_t *smask,
                int sW, int sH, const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                const int32_t *init_nnf, uint64_t seed, int32_t *out_nnf, float *out_dist,
                PMWork &w, hipStream_t s);

// The same without allocating: scratch in a per-device static array of
// kNoAllocScratchFloats floats, whose address noalloc_scratch_init() looks up (once per
// device, with the device current).
constexpr size_t kNoAllocScratchFloats = (size_t)8 << 20;
void noalloc_scratch_init();
void patch_distance_field_noalloc(const DistDesc &d, const float *src, int sW, int sH, const float *tgt, int tW,
                                  int tH, int C, const int32_t *nnf, const uint8_t *tmask, float *out,
                                  hipStream_t s);

// Several PatchMatch calls in lockstep: one launch per initialisation and per sweep
// covers the target tiles of every job.  The jobs share the distance kind, channel count
// and patch side, and pm_batch_supported() holds for each; results equal patchmatch()'s
// on each job, except that a job whose *active is 0 is skipped by the search kernels.
struct PMJob {
    DistDesc d;
    PMConfig cfg;
    const float *src;
    const uint8_t *smask;
    int sW, sH;
    const float *tgt;
    const uint8_t *tmask;
    int tW, tH, C;
    const int32_t *init_nnf;
    uint64_t seed;
    int32_t *out_nnf;
    float *out_dist;
    PMWork *w;
    const int *active;   // device flag, or null (always run)
};
struct PMBatchWork {
    ArgTable args, starts;
};
bool pm_batch_supported(const DistDesc &d, int C, int sW, int sH, const PMConfig &cfg);
void patchmatch_batch(const PMJob *jobs, int n, PMBatchWork &bw, hipStream_t s);

// Distances of every target to a fixed NNF (FLT_MAX elsewhere).
void patch_distance_field(const DistDesc &d, const float *src, int sW, int sH, const float *tgt,
                          int tW, int tH, int C, const int32_t *nnf, const uint8_t *tmask,
                          float *out, PMWork &w, hipStream_t s);

// ---- image updating (k_update.hip) ----
struct UpdWork {
    int w_patch = 0;          // cached patch weights
    float w_sigma = 0.0f;
    int *done_host = nullptr; // pinned: pipelined CG convergence polls (flags + stamp per slot)
    int *done_dev = nullptr;  // the same memory as the kernels address it
    unsigned poll_seq = 1;    // next unused poll stamp (stamps are never reused)
    double *sum_host = nullptr;   // pinned: the update's summary, then its stamp
    double *sum_dev = nullptr;
    unsigned sum_seq = 0;         // stamp of the last summary asked for
    // the list-index map (lidx) is kept while the hole point list is unchanged: callers
    // call points_changed() whenever they rewrite a point list
    unsigned long long pts_version = 1, lidx_version = 0;
    const int *lidx_of = nullptr, *lidx_pts = nullptr;
    int lidx_n = -1;
    size_t lidx_np = 0;
    UpdWork() = default;
    UpdWork(const UpdWork &) = delete;
    UpdWork &operator=(const UpdWork &) = delete;
    ~UpdWork()
    {
        if (done_host) (void)hipHostFree(done_host);
        if (sum_host) (void)hipHostFree(sum_host);
    }
    DBuf<float> weights, grad, conf_one;
    DBuf<double> a1, a2, f1, f2, F1, F2, partials, scalars;
    DBuf<double> x0, x1, r, d0, d1, b, u0, stencil;
    DBuf<float> xs0, xs1, rs, ds0, ds1, bs;   // mixed-precision CG vectors
    DBuf<int> pts, nbr, lidx, counts, total, flag;
    DBuf<uint8_t> pflag;
};

struct CGParams {
    float tol;
    int limit;
    int precision;       // 0 = FP64 (reference recurrence), 1 = FP32 vectors + FP64 reductions
    int refine_every;    // mixed precision: recompute r = b - A x in FP64 every k iterations
};

struct CGStats {
    int iterations[4];
    double nr0[4];
    double nr[4];
};

// AImageUpdating::update for the four updaters, on device.  image (W x H x Cs, channel
// stride Cs) is updated in place on channels [c0, c0+nc); orig is u0 (same stride).
// hole_pts (n_hole, row-major) are the inpainting-domain points.  Returns the total
// difference (synchronises the stream).
// a caller's hole point list has been rewritten: the next update rebuilds its index map
void points_changed(UpdWork &w);
double update(int kind, int patch, float sigma, float lambda, const CGParams &cg, float *image,
              const float *orig, int Cs, int c0, int nc, const uint8_t *hole, const int *hole_pts,
              int n_hole, const uint8_t *ext, const int32_t *nnf, const float *conf, int W, int H,
              UpdWork &w, CGStats *stats, hipStream_t s);

// ---- updates in lockstep (k_update.hip) ----
// The stopping rule of the inpainting loop (image_inpainting.cpp:389), kept on the device:
// after the update, *active = total > tol && ++*count < iters.  part: 0 = one update per
// iteration; 1 and 2 = the first and second of two updates whose totals add (NL-Depth):
// the first stores its total in *partial, the second adds it and applies the rule.  The
// new flag also goes to host[0] (pinned, mapped), followed by `stamp` in host[1].
struct StopRule {
    int *active = nullptr;   // null: no rule (the update's total is not used)
    int *count = nullptr;
    double *total = nullptr;
    double *partial = nullptr;
    int *cg_total = nullptr;   // CG iterations summed over channels and updates
    int iters = 0;
    float tol = 0.0f;
    int part = 0;
    int *host = nullptr;
    unsigned stamp = 0;
};
struct UpdJob {
    int kind, patch;
    float sigma, lambda;
    CGParams cg;
    float *image;
    const float *orig;
    int Cs, c0, nc;
    const uint8_t *hole;
    const int *hole_pts;
    int n_hole;
    const uint8_t *ext;
    const int32_t *nnf;
    const float *conf;
    int W, H;
    UpdWork *w;
    const int *active;   // device flag (0: skip this image), or null
    StopRule stop;
};
struct UpdBatchWork {
    ArgTable jobs, starts;
    unsigned poll_seq = 1;
};
// Several updates in lockstep: one launch per kernel covers every job.  The jobs share
// the kind, patch side, channel count and CG precision.  Each job's result equals
// update()'s.
void update_batch(const UpdJob *jobs, int n, UpdBatchWork &bw, hipStream_t s);

// ---- the multiscale driver (nli_driver.cpp) ----
enum InitKind { INIT_BLACK = 0, INIT_AVG = 1, INIT_NONE = 2, INIT_POISSON = 3 };

struct EmptySource : std::runtime_error {
    EmptySource()
        : std::runtime_error("ERROR: Empty source mask (no complete patches to copy from. This may happen due to "
                             "a too big inpainting domain, too big patch, or too much downscaling)")
    {
    }
};

struct InpaintConfig {
    int method, patch;
    float sigma, lambda, lambda_rgb, lambda_d;
    int iters;
    float tol;
    int scales;
    float coarse, conf_t, conf_a;
    int init, pm_iters, pm_shots, pm_window;
    float cg_tol;
    int cg_limit, cg_precision, cg_refine_every;
    uint64_t seed;
};

struct InpaintReport {
    int scales;
    int iterations[32];
    int width[32], height[32];
    int hole_pixels[32], target_pixels[32];
    double last_total[32];
    int cg_iterations_total;
    int pm_calls;
    double seconds_total, seconds_patchmatch, seconds_update;
};

struct DriverWork {
    std::vector<DBuf<float>> img;
    std::vector<DBuf<uint8_t>> mask;
    DBuf<float> scratch, conf, dtg, up, orig;
    DBuf<int> dtst, hole_pts, low_pts, counts, total, nonempty;
    DBuf<uint8_t> ext, tgt, src, bscr;
    DBuf<int32_t> nnfA, nnfB, nnfC;
    DBuf<float> simpois;   // coarsest-level Poisson fill scratch
    // timing events (PatchMatch / update spans), reused across runs
    std::vector<hipEvent_t> events;
    DriverWork() = default;
    DriverWork(const DriverWork &) = delete;
    DriverWork &operator=(const DriverWork &) = delete;
    ~DriverWork()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

void inpaint(const InpaintConfig &cfg, const float *d_in, const uint8_t *d_mask, int W, int H, int C,
             float *d_out, InpaintReport &rep, DriverWork &dw, PMWork &pmw, UpdWork &uw, hipStream_t s);

// Several images in lockstep on one stream (same parameters, any sizes): every scale's
// PatchMatch calls and updates of all images run as single launches (patchmatch_batch,
// update_batch), and the stopping rule is applied on the device: an image whose
// iterations have ended is skipped by the kernels while the others continue, and the host
// learns it one iteration later without waiting.  Each image's output and report equal
// inpaint()'s.  lockstep_supported() tells whether the parameters and sizes allow it.
struct LockstepImage {
    DriverWork dw;
    PMWork pm;
    UpdWork uw;
};
struct LockstepWork {
    std::vector<std::unique_ptr<LockstepImage>> img;
    PMBatchWork pmb;
    UpdBatchWork ub;
    DBuf<int> state;       // per image: active, count, CG iterations
    DBuf<double> totals;   // per image: total, partial
    int *host = nullptr;   // pinned, mapped: per image [active flag, stamp]
    int *host_dev = nullptr;
    size_t host_slots = 0;
    unsigned stamp = 0;
    LockstepWork() = default;
    LockstepWork(const LockstepWork &) = delete;
    LockstepWork &operator=(const LockstepWork &) = delete;
    ~LockstepWork()
    {
        if (host) (void)hipHostFree(host);
    }
};
bool lockstep_supported(const InpaintConfig &cfg, int W, int H, int C);
void inpaint_lockstep(const InpaintConfig &cfg, int n, const float *const *d_in, const uint8_t *const *d_mask,
                      const int *W, const int *H, int C, float *const *d_out, InpaintReport *rep, LockstepWork &lw,
                      hipStream_t s);

// ---- patch energy, its gradient, extraction/aggregation pair (k_adjoint.hip) ----
struct EnergyDesc {
    int patch;
    float sigma;
    float lambda;
};
struct AdjWork {
    DBuf<float> weights;
    int w_patch = -1;
    float w_sigma = 0.0f;
    DBuf<int> pts, lidx, counts, total;
    DBuf<double> partials, S, Rx, Ry;
};
// energy over the extended-domain centres; with tx >= 0, only the terms that read pixel
// (tx, ty) channel tc, whose value is taken as img + delta (FP64 finite differences)
double patch_energy(const EnergyDesc &d, const float *img, const uint8_t *ext, const int32_t *nnf,
                    const float *conf, int W, int H, int C, int tx, int ty, int tc, double delta, AdjWork &w,
                    hipStream_t s);
// dE/du for every pixel and channel (HWC, FP64)
void energy_gradient(const EnergyDesc &d, const float *img, const uint8_t *ext, const int32_t *nnf,
                     const float *conf, int W, int H, int C, double *grad, AdjWork &w, hipStream_t s);
// side 0: u(p + h), side 1: u(nnf(p) + h) for the extended-domain centres p in row-major
// order; returns the number of centres
int patch_extract(int P, const double *u, const uint8_t *ext, const int32_t *nnf, int W, int H, int C, int side,
                  double *out, size_t out_capacity, AdjWork &w, hipStream_t s);
// the adjoint of patch_extract (same side)
void adjoint_aggregate(int P, const double *patches, const uint8_t *ext, const int32_t *nnf, int W, int H, int C,
                       int side, double *out, AdjWork &w, hipStream_t s);

}  // namespace nli
