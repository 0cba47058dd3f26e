// nli_capi.cpp -- the extern "C" boundary of libnli (declared in include/nli.h).
//
// Host-pointer entry points stage their inputs into context-owned device buffers,
// run on the context's stream and copy the outputs back before returning; _dev
// entry points take device pointers.  No exception crosses the boundary: errors
// become NLI_E_* codes with the message kept in the context.
#include "nli.h"
#include "nli_internal.h"

#include <algorithm>

#include <atomic>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

// One batch worker: a stream and the driver's buffers.
struct nli_worker {
    hipStream_t stream = nullptr;
    nli::PMWork pm;
    nli::UpdWork upd;
    nli::DriverWork drv;
    nli::DBuf<float> fin, fout;
    nli::DBuf<uint8_t> min;
    nli::LockstepWork lock;                                     // lockstep groups
    std::vector<nli::DBuf<float>> lin, lout;                    // their staged images
    std::vector<nli::DBuf<uint8_t>> lmask;
    ~nli_worker()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct nli_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    nli::PMWork pm;
    nli::UpdWork upd;
    nli::DriverWork drv;
    nli::DBuf<float> f0, f1, f2;
    nli::DBuf<uint8_t> b0, b1, b2, b3;
    nli::DBuf<int32_t> n0, n1;
    nli::DBuf<int> pts, counts, total, nonempty;
    nli::DBuf<double> d0, d1;
    nli::DBuf<float> c0, c1, c2, c3;   // colour path: Lab in/out, RGB, depth
    nli::AdjWork adj;
    std::vector<std::unique_ptr<nli_worker>> workers;   // batch workers, created on demand
};

namespace {

// batches: 0 = concurrent per-image drivers (the default: faster on the measured workloads,
// where lockstep groups wait for their slowest image at every scale), 1 = lockstep groups
// where supported, 2 = concurrent per-image drivers
std::atomic<int> g_batch_mode{0};

struct ArgError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

template <class F>
int guard(nli_ctx *ctx, F &&f)
{
    if (!ctx) return NLI_E_ARGUMENT;
    try {
        NLI_HIP(hipSetDevice(ctx->device));
        f();
        ctx->err.clear();
        return NLI_OK;
    } catch (const nli::HipError &e) {
        ctx->err = e.what();
        return NLI_E_DEVICE;
    } catch (const nli::EmptySource &e) {
        ctx->err = e.what();
        return NLI_E_EMPTY_SOURCE;
    } catch (const std::invalid_argument &e) {
        ctx->err = e.what();
        return NLI_E_ARGUMENT;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return NLI_E_INTERNAL;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return NLI_E_INTERNAL;
    }
}

void need(bool cond, const char *msg)
{
    if (!cond) throw ArgError(msg);
}

template <class T>
T *stage_in(nli_ctx *c, nli::DBuf<T> &b, const T *h, size_t n)
{
    b.ensure(n);
    if (h && n) NLI_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return b.p;
}

template <class T>
void stage_out(nli_ctx *c, T *h, const T *d, size_t n)
{
    if (h && n) NLI_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
}

void sync(nli_ctx *c) { NLI_HIP(hipStreamSynchronize(c->stream)); }

nli::DistDesc desc(const nli_dist_params *d)
{
    need(d != nullptr, "null distance parameters");
    need(d->kind >= NLI_DIST_L2 && d->kind <= NLI_DIST_L2RGBD, "unknown patch distance kind");
    need(d->patch >= 1 && d->patch <= 15, "patch side must be in [1, 15]");
    return nli::DistDesc{d->kind, d->patch, d->sigma, d->lambda, d->lambda_d};
}

nli::PMConfig pmconfig(const nli_pm_params *pm)
{
    need(pm != nullptr, "null PatchMatch parameters");
    need(pm->iterations >= 0 && pm->shots >= 0, "negative PatchMatch iterations or shots");
    need(pm->window == -1 || pm->window >= 1, "search window must be -1 or positive");
    return nli::PMConfig{pm->iterations, pm->shots, pm->window, pm->first_step};
}

double do_update(nli_ctx *ctx, const nli_upd_params *u, float *image, const float *orig, const uint8_t *domain,
                 const uint8_t *ext, const int32_t *nnf, const float *conf, int W, int H, int C,
                 nli_cg_stats *stats)
{
    need(u != nullptr, "null update parameters");
    need(W > 0 && H > 0 && C >= 1 && C <= 4, "image must be W x H x C with 1 <= C <= 4");
    const size_t n = (size_t)W * H;
    ctx->pts.ensure(n);
    ctx->counts.ensure(n / 1024 + 2);
    ctx->total.ensure(2);
    const int nh = nli::compact_indices(domain, (int)n, ctx->pts.p, ctx->counts.p, ctx->total.p, ctx->stream);
    nli::points_changed(ctx->upd);
    nli::CGStats st{};
    double total = 0.0;
    int channels = 0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    switch (u->kind) {
    case NLI_UPD_MEANS:
    case NLI_UPD_MEDIANS:
        total = nli::update(u->kind, u->patch, u->sigma, 0.0f, nli::CGParams{0, 0, 0, 0}, image, orig, C, 0, C,
                            domain, ctx->pts.p, nh, ext, nnf, conf, W, H, ctx->upd, nullptr, ctx->stream);
        break;
    case NLI_UPD_POISSON: {
        nli::CGParams cg{u->cg_tol, u->cg_limit, u->cg_precision, u->cg_refine_every};
        need(cg.limit >= 0, "negative CG iteration limit");
        total = nli::update(NLI_UPD_POISSON, u->patch, u->sigma, u->lambda, cg, image, orig, C, 0, C, domain,
                            ctx->pts.p, nh, ext, nnf, conf, W, H, ctx->upd, &st, ctx->stream);
        channels = C;
        if (stats)
            for (int c = 0; c < C; c++) {
                stats->cg_iterations[c] = st.iterations[c];
                stats->cg_nr0[c] = st.nr0[c];
                stats->cg_nr[c] = st.nr[c];
            }
        break;
    }
    case NLI_UPD_DEPTH: {
        need(C == 4, "Separate: Channels do not match.");
        // patch_non_local_depth.cpp:24-55: two NL-Poisson updates with (1e-6, 1000); the
        // "original" of both is the current image (:48-49)
        nli::CGParams cg{0.000001f, 1000, u->cg_precision, u->cg_refine_every};
        total += nli::update(NLI_UPD_POISSON, u->patch, u->sigma, u->lambda, cg, image, image, 4, 0, 3, domain,
                             ctx->pts.p, nh, ext, nnf, conf, W, H, ctx->upd, &st, ctx->stream);
        if (stats)
            for (int c = 0; c < 3; c++) {
                stats->cg_iterations[c] = st.iterations[c];
                stats->cg_nr0[c] = st.nr0[c];
                stats->cg_nr[c] = st.nr[c];
            }
        total += nli::update(NLI_UPD_POISSON, u->patch, u->sigma, u->lambda_d, cg, image, image, 4, 3, 1, domain,
                             ctx->pts.p, nh, ext, nnf, conf, W, H, ctx->upd, &st, ctx->stream);
        if (stats) {
            stats->cg_iterations[3] = st.iterations[0];
            stats->cg_nr0[3] = st.nr0[0];
            stats->cg_nr[3] = st.nr[0];
        }
        channels = 4;
        break;
    }
    default: throw ArgError("unknown image updating kind");
    }
    if (stats) stats->channels = channels;
    return total;
}

nli::InpaintConfig inpaint_config(const nli_inpaint_params *p)
{
    need(p != nullptr, "null inpainting parameters");
    need(p->patch > 0, "ERROR: patch_side needs to be a positive integer.");
    need(p->patch <= 15, "patch side must be at most 15");
    need(!(p->coarse > 1 || p->coarse < 0), "ERROR: coarsest rate should be between 0 and 1.");
    need(p->iters >= 0, "ERROR: inpainting_iterations cannot be negative.");
    need(!(p->lambda > 1.f || p->lambda < 0.f), "ERROR: lambda should be between 0 and 1.");
    need(p->scales >= 1 && p->scales <= 32, "ERROR: number of scales must be at least 1.");
    need(p->init >= NLI_INIT_BLACK && p->init <= NLI_INIT_POISSON, "ERROR: Unknown initialization type");
    nli::InpaintConfig c;
    c.method = p->method;
    c.patch = p->patch;
    c.sigma = p->sigma;
    c.lambda = p->lambda;
    c.lambda_rgb = p->lambda_rgb;
    c.lambda_d = p->lambda_d;
    c.iters = p->iters;
    c.tol = p->tol;
    c.scales = p->scales;
    c.coarse = p->coarse;
    c.conf_t = p->conf_t;
    c.conf_a = p->conf_a;
    c.init = p->init;
    c.pm_iters = p->pm_iters;
    c.pm_shots = p->pm_shots;
    c.pm_window = p->pm_window;
    c.cg_tol = p->cg_tol;
    c.cg_limit = p->cg_limit;
    c.cg_precision = p->cg_precision;
    c.cg_refine_every = p->cg_refine_every;
    c.seed = p->seed;
    return c;
}

void report(const nli::InpaintReport &r, nli_inpaint_stats *st)
{
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->scales = r.scales;
    for (int i = 0; i < 32; i++) {
        st->iterations[i] = r.iterations[i];
        st->width[i] = r.width[i];
        st->height[i] = r.height[i];
        st->hole_pixels[i] = r.hole_pixels[i];
        st->target_pixels[i] = r.target_pixels[i];
    }
    st->cg_iterations_total = r.cg_iterations_total;
    st->seconds_total = r.seconds_total;
    st->seconds_patchmatch = r.seconds_patchmatch;
    st->seconds_update = r.seconds_update;
    st->seconds_other = r.seconds_total - r.seconds_patchmatch - r.seconds_update;
}

}  // namespace

namespace {

// Runs the batch on `workers` host threads, each driving its own stream; images are
// taken from a shared counter.  staged = false: in/mask/out are device pointers.
// params: one record for all images (per_image = false) or one per image
void run_batch(nli_ctx *ctx, const nli_inpaint_params *params, bool per_image, int n, const float *const *in,
               const uint8_t *const *mask, const int *W, const int *H, int C, float *const *out, int workers,
               nli_inpaint_stats *stats, bool staged)
{
    need(n >= 0 && in && mask && W && H && out && params, "bad batch arguments");
    need(C >= 1 && C <= 4, "image must have 1 to 4 channels");
    for (int i = 0; i < n; i++) need(in[i] && mask[i] && out[i] && W[i] > 0 && H[i] > 0, "bad batch image");
    if (n == 0) return;
    std::vector<nli::InpaintConfig> cfgs;
    for (int i = 0; i < (per_image ? n : 1); i++) cfgs.push_back(inpaint_config(params + i));
    // largest images first (longest-processing-time order): the batch ends sooner, and
    // each image's result does not depend on when or where it runs
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const nli::InpaintConfig &ca = cfgs[per_image ? a : 0], &cb = cfgs[per_image ? b : 0];
        return (double)W[a] * H[a] * ca.scales * ca.patch > (double)W[b] * H[b] * cb.scales * cb.patch;
    });
    if (workers <= 0) workers = 4;
    workers = std::min(workers, n);
    while ((int)ctx->workers.size() < workers) {
        auto wk = std::make_unique<nli_worker>();
        NLI_HIP(hipStreamCreateWithFlags(&wk->stream, hipStreamNonBlocking));
        ctx->workers.push_back(std::move(wk));
    }
    // lockstep groups: every image has the same parameters and the lockstep driver
    // supports them; the images (largest first) are dealt round-robin to `workers`
    // groups, one per host thread and stream
    bool lockstep = g_batch_mode.load() == 1;
    for (int i = 0; i < n && lockstep; i++) {
        const nli::InpaintConfig &c = cfgs[per_image ? i : 0];
        lockstep = nli::lockstep_supported(c, W[i], H[i], C) &&
                   (!per_image || std::memcmp(&c, &cfgs[0], sizeof(c)) == 0);
    }
    std::vector<std::vector<int>> groups(workers);
    for (int k = 0; k < n; k++) groups[k % workers].push_back(order[k]);
    std::atomic<int> next(0);
    std::mutex mu;
    std::string first_error;
    int first_code = 0;
    auto body = [&](nli_worker *wk, int group) {
        try {
            NLI_HIP(hipSetDevice(ctx->device));
            if (lockstep && !groups[group].empty()) {
                const std::vector<int> &g = groups[group];
                const int m = (int)g.size();
                std::vector<const float *> gin(m);
                std::vector<const uint8_t *> gmask(m);
                std::vector<float *> gout(m);
                std::vector<int> gw(m), gh(m);
                std::vector<nli::InpaintReport> reps(m);
                if (staged) {
                    wk->lin.resize(std::max<size_t>(wk->lin.size(), m));
                    wk->lout.resize(std::max<size_t>(wk->lout.size(), m));
                    wk->lmask.resize(std::max<size_t>(wk->lmask.size(), m));
                }
                for (int j = 0; j < m; j++) {
                    const int i = g[j];
                    const size_t np = (size_t)W[i] * H[i];
                    gw[j] = W[i];
                    gh[j] = H[i];
                    gin[j] = in[i];
                    gmask[j] = mask[i];
                    gout[j] = out[i];
                    if (staged) {
                        gin[j] = wk->lin[j].ensure(np * C);
                        gmask[j] = wk->lmask[j].ensure(np);
                        gout[j] = wk->lout[j].ensure(np * C);
                        NLI_HIP(hipMemcpyAsync((void *)gin[j], in[i], np * C * sizeof(float), hipMemcpyHostToDevice,
                                               wk->stream));
                        NLI_HIP(hipMemcpyAsync((void *)gmask[j], mask[i], np, hipMemcpyHostToDevice, wk->stream));
                    }
                }
                nli::inpaint_lockstep(cfgs[0], m, gin.data(), gmask.data(), gw.data(), gh.data(), C, gout.data(),
                                      reps.data(), wk->lock, wk->stream);
                for (int j = 0; j < m; j++) {
                    const int i = g[j];
                    if (staged)
                        NLI_HIP(hipMemcpyAsync(out[i], gout[j], (size_t)W[i] * H[i] * C * sizeof(float),
                                               hipMemcpyDeviceToHost, wk->stream));
                    if (stats) report(reps[j], &stats[i]);
                }
                NLI_HIP(hipStreamSynchronize(wk->stream));
            }
            for (int k = next++; k < n && !lockstep; k = next++) {
                const int i = order[k];
                const nli::InpaintConfig &cfg = cfgs[per_image ? i : 0];
                const size_t np = (size_t)W[i] * H[i];
                const float *d_in = in[i];
                const uint8_t *d_m = mask[i];
                float *d_out = out[i];
                if (staged) {
                    d_in = wk->fin.ensure(np * C);
                    d_m = wk->min.ensure(np);
                    d_out = wk->fout.ensure(np * C);
                    NLI_HIP(hipMemcpyAsync((void *)d_in, in[i], np * C * sizeof(float), hipMemcpyHostToDevice,
                                           wk->stream));
                    NLI_HIP(hipMemcpyAsync((void *)d_m, mask[i], np, hipMemcpyHostToDevice, wk->stream));
                }
                nli::InpaintReport rep{};
                nli::inpaint(cfg, d_in, d_m, W[i], H[i], C, d_out, rep, wk->drv, wk->pm, wk->upd, wk->stream);
                if (staged) {
                    NLI_HIP(hipMemcpyAsync(out[i], d_out, np * C * sizeof(float), hipMemcpyDeviceToHost, wk->stream));
                    NLI_HIP(hipStreamSynchronize(wk->stream));
                }
                if (stats) report(rep, &stats[i]);
            }
        } catch (const nli::EmptySource &e) {
            std::lock_guard<std::mutex> g(mu);
            if (!first_code) { first_code = NLI_E_EMPTY_SOURCE; first_error = e.what(); }
        } catch (const nli::HipError &e) {
            std::lock_guard<std::mutex> g(mu);
            if (!first_code) { first_code = NLI_E_DEVICE; first_error = e.what(); }
        } catch (const std::invalid_argument &e) {
            std::lock_guard<std::mutex> g(mu);
            if (!first_code) { first_code = NLI_E_ARGUMENT; first_error = e.what(); }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> g(mu);
            if (!first_code) { first_code = NLI_E_INTERNAL; first_error = e.what(); }
        }
        next = n;   // stop the other workers after a failure or at the end
    };
    std::vector<std::thread> threads;
    for (int k = 1; k < workers; k++) threads.emplace_back(body, ctx->workers[k].get(), k);
    body(ctx->workers[0].get(), 0);
    for (auto &t : threads) t.join();
    if (first_code == NLI_E_EMPTY_SOURCE) throw nli::EmptySource();
    if (first_code == NLI_E_ARGUMENT) throw ArgError(first_error);
    if (first_code == NLI_E_DEVICE) throw nli::HipError(hipErrorUnknown, first_error);
    if (first_code) throw std::runtime_error(first_error);
}

}  // namespace

extern "C" {

int nli_version(void) { return 1; }

int nli_ctx_create(int device, nli_ctx **out)
{
    if (!out) return NLI_E_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return NLI_E_DEVICE;
    nli_ctx *c = new (std::nothrow) nli_ctx();
    if (!c) return NLI_E_INTERNAL;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return NLI_E_DEVICE;
    }
    try {
        nli::noalloc_scratch_init();   // the device address of nli_patch_distance_dev_noalloc's scratch
    } catch (const std::exception &) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return NLI_E_DEVICE;
    }
    *out = c;
    return NLI_OK;
}

int nli_ctx_destroy(nli_ctx *ctx)
{
    if (!ctx) return NLI_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NLI_OK;
}

const char *nli_last_error(const nli_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int nli_synchronize(nli_ctx *ctx)
{
    return guard(ctx, [&] { sync(ctx); });
}

int nli_dev_alloc(nli_ctx *ctx, uint64_t bytes, void **out)
{
    return guard(ctx, [&] {
        need(out != nullptr, "null output pointer");
        NLI_HIP(hipMalloc(out, bytes ? bytes : 1));
    });
}

int nli_dev_free(nli_ctx *ctx, void *p)
{
    return guard(ctx, [&] {
        if (p) NLI_HIP(hipFree(p));
    });
}

int nli_copy_h2d(nli_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return guard(ctx, [&] {
        NLI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        sync(ctx);
    });
}

int nli_copy_d2h(nli_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return guard(ctx, [&] {
        NLI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    });
}

int nli_pm_tiles_per_block(int tiles)
{
    if (tiles < 0 || tiles > 3) return NLI_E_ARGUMENT;
    nli::set_pm_tiles_per_block(tiles);
    return NLI_OK;
}

int nli_batch_mode(int mode)
{
    if (mode < 0 || mode > 2) return NLI_E_ARGUMENT;
    g_batch_mode = mode;
    return NLI_OK;
}

int nli_pm_row_scan(int mode)
{
    if (mode < 0 || mode > 2) return NLI_E_ARGUMENT;
    nli::set_pm_row_scan(mode);
    return NLI_OK;
}

int nli_copy_d2d(nli_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return guard(ctx, [&] {
        NLI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        sync(ctx);
    });
}

int nli_profile_enable(nli_ctx *ctx, int kernels)
{
    return guard(ctx, [&] {
        sync(ctx);
        nli::prof::enable((unsigned)kernels);
    });
}

int nli_profile_enable_sampled(nli_ctx *ctx, int kernels, int period)
{
    return guard(ctx, [&] {
        need(period >= 1, "sampling period must be at least 1");
        sync(ctx);
        nli::prof::enable((unsigned)kernels, (unsigned)period);
    });
}

int nli_profile_read(nli_ctx *ctx, int kernel, int *launches, double *total_ms, double *work, double *units)
{
    return guard(ctx, [&] {
        need(kernel >= 0 && kernel < nli::TK_COUNT, "unknown kernel id");
        sync(ctx);
        nli::prof::read(kernel, launches, total_ms, work, units);
    });
}

int nli_sample_size(int size, float factor) { return nli::sample_size(size, factor); }
float nli_subsampling_rate(float ratio, int scales) { return nli::subsampling_rate(ratio, scales); }

int nli_gradient(nli_ctx *ctx, const float *img, int W, int H, int C, float *out)
{
    return guard(ctx, [&] {
        need(img && out && W > 0 && H > 0 && C > 0, "bad gradient arguments");
        const size_t n = (size_t)W * H;
        const float *d_img = stage_in(ctx, ctx->f0, img, n * C);
        float *d_out = ctx->f1.ensure(n * C * 2);
        nli::gradient(d_img, W, H, C, d_out, ctx->stream);
        stage_out(ctx, out, d_out, n * C * 2);
        sync(ctx);
    });
}

int nli_distance_transform(nli_ctx *ctx, const uint8_t *mask, int W, int H, float *out)
{
    return guard(ctx, [&] {
        need(mask && out && W > 0 && H > 0, "bad distance transform arguments");
        const size_t n = (size_t)W * H;
        const uint8_t *d_m = stage_in(ctx, ctx->b0, mask, n);
        float *d_out = ctx->f0.ensure(n);
        float *g = ctx->f1.ensure(n);
        int *st = ctx->pts.ensure(2 * n);
        nli::distance_transform(d_m, W, H, d_out, g, st, ctx->stream);
        stage_out(ctx, out, d_out, n);
        sync(ctx);
    });
}

int nli_confidence(nli_ctx *ctx, const uint8_t *mask, int W, int H, float decay, float asym, float *out)
{
    return guard(ctx, [&] {
        need(mask && out && W > 0 && H > 0, "bad confidence arguments");
        const size_t n = (size_t)W * H;
        const uint8_t *d_m = stage_in(ctx, ctx->b0, mask, n);
        float *dt = ctx->f0.ensure(n);
        float *g = ctx->f1.ensure(n);
        float *d_out = ctx->f2.ensure(n);
        int *st = ctx->pts.ensure(2 * n);
        if (decay > 0) nli::distance_transform(d_m, W, H, dt, g, st, ctx->stream);
        nli::confidence(d_m, dt, W, H, decay, asym, d_out, ctx->stream);
        stage_out(ctx, out, d_out, n);
        sync(ctx);
    });
}

int nli_downsample_with_mask(nli_ctx *ctx, const float *img, const uint8_t *mask, int W, int H, int C,
                             float factor, float threshold, int mask_val, float *out_img, uint8_t *out_mask)
{
    return guard(ctx, [&] {
        need(img && mask && out_img && out_mask && W > 0 && H > 0 && C >= 1 && C <= 4, "bad downsample arguments");
        need(factor > 0 && factor < 1.0, "downsampling factor must be in (0, 1)");
        const size_t n = (size_t)W * H;
        const int sW = nli::sample_size(W, factor), sH = nli::sample_size(H, factor);
        need(sW > 0 && sH > 0, "downsampled image is empty");
        const float *d_img = stage_in(ctx, ctx->f0, img, n * C);
        const uint8_t *d_m = stage_in(ctx, ctx->b0, mask, n);
        float *scratch = ctx->f1.ensure(n * (C + 1));
        float *d_oi = ctx->f2.ensure((size_t)sW * sH * C);
        uint8_t *d_om = ctx->b1.ensure((size_t)sW * sH);
        nli::downsample_with_mask(d_img, d_m, W, H, C, factor, threshold, mask_val, d_oi, d_om, scratch, ctx->stream);
        stage_out(ctx, out_img, d_oi, (size_t)sW * sH * C);
        stage_out(ctx, out_mask, d_om, (size_t)sW * sH);
        sync(ctx);
    });
}

int nli_downsample(nli_ctx *ctx, const float *img, int W, int H, int C, float factor, float *out)
{
    return guard(ctx, [&] {
        need(img && out && W > 0 && H > 0 && C >= 1 && C <= 4, "bad downsample arguments");
        need(factor > 0 && factor < 1.0, "downsampling factor must be in (0, 1)");
        const size_t n = (size_t)W * H;
        const int sW = nli::sample_size(W, factor), sH = nli::sample_size(H, factor);
        need(sW > 0 && sH > 0, "downsampled image is empty");
        const float *d_img = stage_in(ctx, ctx->f0, img, n * C);
        float *scratch = ctx->f1.ensure(n * (C + 1));
        float *d_o = ctx->f2.ensure((size_t)sW * sH * C);
        nli::downsample(d_img, W, H, C, factor, d_o, scratch, ctx->stream);
        stage_out(ctx, out, d_o, (size_t)sW * sH * C);
        sync(ctx);
    });
}

int nli_downsample_mask(nli_ctx *ctx, const uint8_t *mask, int W, int H, float factor, float threshold, uint8_t *out)
{
    return guard(ctx, [&] {
        need(mask && out && W > 0 && H > 0, "bad downsample arguments");
        need(factor > 0 && factor < 1.0, "downsampling factor must be in (0, 1)");
        const size_t n = (size_t)W * H;
        const int sW = nli::sample_size(W, factor), sH = nli::sample_size(H, factor);
        need(sW > 0 && sH > 0, "downsampled mask is empty");
        const uint8_t *d_m = stage_in(ctx, ctx->b0, mask, n);
        float *scratch = ctx->f1.ensure(n);
        uint8_t *d_o = ctx->b1.ensure((size_t)sW * sH);
        nli::downsample_mask(d_m, W, H, factor, threshold, d_o, scratch, ctx->stream);
        stage_out(ctx, out, d_o, (size_t)sW * sH);
        sync(ctx);
    });
}

int nli_upsample(nli_ctx *ctx, const float *img, int W, int H, int C, int W2, int H2, float *out)
{
    return guard(ctx, [&] {
        need(img && out && W > 0 && H > 0 && W2 > 0 && H2 > 0 && C > 0, "bad upsample arguments");
        const float *d_in = stage_in(ctx, ctx->f0, img, (size_t)W * H * C);
        float *d_o = ctx->f1.ensure((size_t)W2 * H2 * C);
        nli::upsample(d_in, W, H, C, W2, H2, d_o, ctx->stream);
        stage_out(ctx, out, d_o, (size_t)W2 * H2 * C);
        sync(ctx);
    });
}

int nli_extended_domain(nli_ctx *ctx, const uint8_t *hole, int W, int H, int patch, uint8_t *out)
{
    return guard(ctx, [&] {
        need(hole && out && W > 0 && H > 0 && patch > 0, "bad extended domain arguments");
        const size_t n = (size_t)W * H;
        const uint8_t *d_h = stage_in(ctx, ctx->b0, hole, n);
        uint8_t *d_e = ctx->b1.ensure(n);
        uint8_t *scr = ctx->b2.ensure(n);
        nli::extended_domain(d_h, W, H, patch, d_e, scr, ctx->stream);
        stage_out(ctx, out, d_e, n);
        sync(ctx);
    });
}

int nli_scale_nnf(nli_ctx *ctx, const int32_t *lower_nnf, const uint8_t *lower_hole, int lW, int lH, int W, int H,
                  int32_t *out)
{
    return guard(ctx, [&] {
        need(lower_nnf && lower_hole && out && lW > 0 && lH > 0 && W > 0 && H > 0, "bad NNF scaling arguments");
        const size_t ln = (size_t)lW * lH, n = (size_t)W * H;
        const int32_t *d_l = stage_in(ctx, ctx->n0, lower_nnf, 2 * ln);
        const uint8_t *d_h = stage_in(ctx, ctx->b0, lower_hole, ln);
        int *pts = ctx->pts.ensure(ln);
        int *counts = ctx->counts.ensure(ln / 1024 + 2);
        int *total = ctx->total.ensure(2);
        int32_t *d_o = ctx->n1.ensure(2 * n);
        const int nl = nli::compact_indices(d_h, (int)ln, pts, counts, total, ctx->stream);
        nli::scale_nnf(d_l, pts, nl, lW, lH, W, H, d_o, ctx->stream);
        stage_out(ctx, out, d_o, 2 * n);
        sync(ctx);
    });
}

int nli_scale_masks(nli_ctx *ctx, const uint8_t *hole, int W, int H, int patch, int erode, uint8_t *ext,
                    uint8_t *target, uint8_t *source)
{
    int nonempty = 1;
    int rc = guard(ctx, [&] {
        need(hole && ext && target && source && W > 0 && H > 0 && patch > 0, "bad mask arguments");
        const size_t n = (size_t)W * H;
        const uint8_t *d_h = stage_in(ctx, ctx->b0, hole, n);
        uint8_t *d_e = ctx->b1.ensure(n);
        uint8_t *scr = ctx->b2.ensure(n);
        uint8_t *d_t = ctx->b3.ensure(n);
        uint8_t *d_s = ctx->b2.p;   // scratch is free again after the dilation
        int *d_ne = ctx->nonempty.ensure(1);
        nli::extended_domain(d_h, W, H, patch, d_e, scr, ctx->stream);
        sync(ctx);
        nli::scale_masks(d_e, W, H, patch / 2, erode ? 1 : 0, d_t, d_s, d_ne, ctx->stream);
        stage_out(ctx, ext, d_e, n);
        stage_out(ctx, target, d_t, n);
        stage_out(ctx, source, d_s, n);
        NLI_HIP(hipMemcpyAsync(&nonempty, d_ne, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    });
    if (rc == NLI_OK && erode && !nonempty) {
        ctx->err = nli::EmptySource().what();
        return NLI_E_EMPTY_SOURCE;
    }
    return rc;
}

namespace {

nli::EnergyDesc edesc(const nli_energy_params *e)
{
    need(e != nullptr, "null energy parameters");
    need(e->patch >= 1 && e->patch <= 15 && e->patch % 2 == 1, "patch side must be odd and at most 15");
    return nli::EnergyDesc{e->patch, e->sigma, e->lambda};
}

}  // namespace

int nli_patch_energy(nli_ctx *ctx, const nli_energy_params *params, const float *image, const uint8_t *ext,
                     const int32_t *nnf, const float *conf, int W, int H, int C, int px, int py, int pc, double delta,
                     double *energy)
{
    return guard(ctx, [&] {
        const nli::EnergyDesc d = edesc(params);
        need(image && ext && nnf && energy && W > 0 && H > 0 && C >= 1 && C <= 4, "bad energy arguments");
        const size_t n = (size_t)W * H;
        const float *d_i = stage_in(ctx, ctx->f0, image, n * C);
        const uint8_t *d_e = stage_in(ctx, ctx->b0, ext, n);
        const int32_t *d_n = stage_in(ctx, ctx->n0, nnf, 2 * n);
        const float *d_c = conf ? stage_in(ctx, ctx->f1, conf, n) : nullptr;
        *energy = nli::patch_energy(d, d_i, d_e, d_n, d_c, W, H, C, px, py, pc, delta, ctx->adj, ctx->stream);
    });
}

int nli_energy_gradient(nli_ctx *ctx, const nli_energy_params *params, const float *image, const uint8_t *ext,
                        const int32_t *nnf, const float *conf, int W, int H, int C, double *grad)
{
    return guard(ctx, [&] {
        const nli::EnergyDesc d = edesc(params);
        need(image && ext && nnf && grad && W > 0 && H > 0 && C >= 1 && C <= 4, "bad gradient arguments");
        const size_t n = (size_t)W * H;
        const float *d_i = stage_in(ctx, ctx->f0, image, n * C);
        const uint8_t *d_e = stage_in(ctx, ctx->b0, ext, n);
        const int32_t *d_n = stage_in(ctx, ctx->n0, nnf, 2 * n);
        const float *d_c = conf ? stage_in(ctx, ctx->f1, conf, n) : nullptr;
        double *d_g = ctx->d0.ensure(n * C);
        nli::energy_gradient(d, d_i, d_e, d_n, d_c, W, H, C, d_g, ctx->adj, ctx->stream);
        stage_out(ctx, grad, d_g, n * C);
        sync(ctx);
    });
}

int nli_patch_extract(nli_ctx *ctx, int patch, const double *u, const uint8_t *ext, const int32_t *nnf, int W,
                      int H, int C, int side, double *patches, uint64_t capacity, int *n_centres)
{
    return guard(ctx, [&] {
        need(u && ext && nnf && patches && n_centres && W > 0 && H > 0 && C >= 1 && C <= 4, "bad extract arguments");
        need(side == 0 || side == 1, "side must be 0 (target) or 1 (source)");
        const size_t n = (size_t)W * H;
        const double *d_u = stage_in(ctx, ctx->d0, u, n * C);
        const uint8_t *d_e = stage_in(ctx, ctx->b0, ext, n);
        const int32_t *d_n = stage_in(ctx, ctx->n0, nnf, 2 * n);
        double *d_p = ctx->d1.ensure(std::max<size_t>((size_t)capacity, 1));
        const int nc = nli::patch_extract(patch, d_u, d_e, d_n, W, H, C, side, d_p, (size_t)capacity, ctx->adj,
                                          ctx->stream);
        stage_out(ctx, patches, d_p, (size_t)nc * patch * patch * C);
        sync(ctx);
        *n_centres = nc;
    });
}

int nli_adjoint_aggregate(nli_ctx *ctx, int patch, const double *patches, int n_centres, const uint8_t *ext,
                          const int32_t *nnf, int W, int H, int C, int side, double *out)
{
    return guard(ctx, [&] {
        need(patches && ext && nnf && out && W > 0 && H > 0 && C >= 1 && C <= 4 && n_centres >= 0,
             "bad aggregate arguments");
        need(side == 0 || side == 1, "side must be 0 (target) or 1 (source)");
        const size_t n = (size_t)W * H;
        const uint8_t *d_e = stage_in(ctx, ctx->b0, ext, n);
        const int32_t *d_n = stage_in(ctx, ctx->n0, nnf, 2 * n);
        size_t ne = 0;
        for (size_t i = 0; i < n; i++) ne += ext[i] != 0;
        need(ne == (size_t)n_centres, "n_centres differs from the number of extended-domain points");
        const double *d_p = stage_in(ctx, ctx->d1, patches, (size_t)n_centres * patch * patch * C);
        double *d_o = ctx->d0.ensure(n * C);
        nli::adjoint_aggregate(patch, d_p, d_e, d_n, W, H, C, side, d_o, ctx->adj, ctx->stream);
        stage_out(ctx, out, d_o, n * C);
        sync(ctx);
    });
}

int nli_patch_distance(nli_ctx *ctx, const nli_dist_params *dist, const float *src, int sW, int sH, const float *tgt,
                       int tW, int tH, int C, const int32_t *nnf, const uint8_t *tmask, float *out)
{
    return guard(ctx, [&] {
        nli::DistDesc d = desc(dist);
        need(src && tgt && nnf && tmask && out && sW > 0 && sH > 0 && tW > 0 && tH > 0, "bad distance arguments");
        need(C >= 1 && C <= 4, "1 to 4 channels supported");
        need(d.kind != NLI_DIST_L2RGBD || C == 4, "L2 RGBD Distance: Channels do not match.");
        const size_t sn = (size_t)sW * sH, tn = (size_t)tW * tH;
        const float *d_s = stage_in(ctx, ctx->f0, src, sn * C);
        const float *d_t = (tgt == src) ? d_s : stage_in(ctx, ctx->f1, tgt, tn * C);
        const int32_t *d_n = stage_in(ctx, ctx->n0, nnf, tn * 2);
        const uint8_t *d_tm = stage_in(ctx, ctx->b0, tmask, tn);
        float *d_o = ctx->f2.ensure(tn);
        nli::patch_distance_field(d, d_s, sW, sH, d_t, tW, tH, C, d_n, d_tm, d_o, ctx->pm, ctx->stream);
        stage_out(ctx, out, d_o, tn);
        sync(ctx);
    });
}

int nli_patch_distance_dev_noalloc(nli_ctx *ctx, const nli_dist_params *dist, const float *src, int sW, int sH,
                                   const float *tgt, int tW, int tH, int C, const int32_t *nnf, const uint8_t *tmask,
                                   float *out)
{
    return guard(ctx, [&] {
        nli::DistDesc d = desc(dist);
        need(src && tgt && nnf && tmask && out && sW > 0 && sH > 0 && tW > 0 && tH > 0, "bad distance arguments");
        need(C >= 1 && C <= 4, "1 to 4 channels supported");
        need(d.kind != NLI_DIST_L2RGBD || C == 4, "L2 RGBD Distance: Channels do not match.");
        nli::patch_distance_field_noalloc(d, src, sW, sH, tgt, tW, tH, C, nnf, tmask, out, ctx->stream);
        sync(ctx);
    });
}

int nli_patchmatch_dev(nli_ctx *ctx, const nli_dist_params *dist, const nli_pm_params *pm, const float *src,
                       const uint8_t *smask, int sW, int sH, const float *tgt, const uint8_t *tmask, int tW, int tH,
                       int C, const int32_t *init_nnf, uint64_t seed, int32_t *out_nnf, float *out_dist)
{
    return guard(ctx, [&] {
        nli::DistDesc d = desc(dist);
        nli::PMConfig cfg = pmconfig(pm);
        need(src && smask && tgt && tmask && out_nnf && sW > 0 && sH > 0 && tW > 0 && tH > 0,
             "PatchMatch: size mismatch or missing input (the reference returns an empty field)");
        need(C >= 1 && C <= 4, "1 to 4 channels supported");
        need(d.kind != NLI_DIST_L2RGBD || C == 4, "L2 RGBD Distance: Channels do not match.");
        need(init_nnf != out_nnf, "initial field and output must not alias");
        nli::pm_masks_changed(ctx->pm);
        nli::patchmatch(d, cfg, src, smask, sW, sH, tgt, tmask, tW, tH, C, init_nnf, seed, out_nnf, out_dist, ctx->pm,
                        ctx->stream);
        sync(ctx);
    });
}

int nli_patchmatch(nli_ctx *ctx, const nli_dist_params *dist, const nli_pm_params *pm, const float *src,
                   const uint8_t *smask, int sW, int sH, const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                   const int32_t *init_nnf, uint64_t seed, int32_t *out_nnf, float *out_dist)
{
    return guard(ctx, [&] {
        nli::DistDesc d = desc(dist);
        nli::PMConfig cfg = pmconfig(pm);
        need(src && smask && tgt && tmask && out_nnf && sW > 0 && sH > 0 && tW > 0 && tH > 0,
             "PatchMatch: size mismatch or missing input (the reference returns an empty field)");
        need(C >= 1 && C <= 4, "1 to 4 channels supported");
        need(d.kind != NLI_DIST_L2RGBD || C == 4, "L2 RGBD Distance: Channels do not match.");
        const size_t sn = (size_t)sW * sH, tn = (size_t)tW * tH;
        const float *d_s = stage_in(ctx, ctx->f0, src, sn * C);
        const float *d_t = (tgt == src) ? d_s : stage_in(ctx, ctx->f1, tgt, tn * C);
        const uint8_t *d_sm = stage_in(ctx, ctx->b0, smask, sn);
        const uint8_t *d_tm = (tmask == smask) ? d_sm : stage_in(ctx, ctx->b1, tmask, tn);
        const int32_t *d_init = init_nnf ? stage_in(ctx, ctx->n1, init_nnf, tn * 2) : nullptr;
        int32_t *d_out = ctx->n0.ensure(tn * 2);
        float *d_dist = out_dist ? ctx->f2.ensure(tn) : nullptr;
        nli::pm_masks_changed(ctx->pm);
        nli::patchmatch(d, cfg, d_s, d_sm, sW, sH, d_t, d_tm, tW, tH, C, d_init, seed, d_out, d_dist, ctx->pm,
                        ctx->stream);
        stage_out(ctx, out_nnf, d_out, tn * 2);
        if (out_dist) stage_out(ctx, out_dist, d_dist, tn);
        sync(ctx);
    });
}

int nli_update_dev(nli_ctx *ctx, const nli_upd_params *upd, float *image, const float *orig, const uint8_t *domain,
                   const uint8_t *ext, const int32_t *nnf, const float *conf, int W, int H, int C, double *total_diff,
                   nli_cg_stats *stats)
{
    return guard(ctx, [&] {
        need(image && orig && domain && ext && nnf, "bad update arguments");
        double t = do_update(ctx, upd, image, orig, domain, ext, nnf, conf, W, H, C, stats);
        if (total_diff) *total_diff = t;
    });
}

int nli_update(nli_ctx *ctx, const nli_upd_params *upd, float *image, const float *orig, const uint8_t *domain,
               const uint8_t *ext, const int32_t *nnf, const float *conf, int W, int H, int C, double *total_diff,
               nli_cg_stats *stats)
{
    return guard(ctx, [&] {
        need(image && orig && domain && ext && nnf, "bad update arguments");
        need(W > 0 && H > 0 && C >= 1 && C <= 4, "image must be W x H x C with 1 <= C <= 4");
        const size_t n = (size_t)W * H;
        float *d_img = stage_in(ctx, ctx->f0, (const float *)image, n * C);
        const float *d_orig = (orig == image) ? d_img : stage_in(ctx, ctx->f1, orig, n * C);
        const uint8_t *d_dom = stage_in(ctx, ctx->b0, domain, n);
        const uint8_t *d_ext = stage_in(ctx, ctx->b1, ext, n);
        const int32_t *d_nnf = stage_in(ctx, ctx->n0, nnf, n * 2);
        const float *d_conf = conf ? stage_in(ctx, ctx->f2, conf, n) : nullptr;
        double t = do_update(ctx, upd, d_img, d_orig, d_dom, d_ext, d_nnf, d_conf, W, H, C, stats);
        stage_out(ctx, image, (const float *)d_img, n * C);
        sync(ctx);
        if (total_diff) *total_diff = t;
    });
}

int nli_poisson_fill_dev(nli_ctx *ctx, float *img, const uint8_t *mask, int W, int H, int C)
{
    return guard(ctx, [&] {
        need(img && mask && W > 0 && H > 0 && C >= 1 && C <= 4, "bad poisson fill arguments");
        float *scr = ctx->f2.ensure(nli::poisson_fill_scratch(W, H, C));
        nli::poisson_fill_device(img, mask, W, H, C, scr, ctx->stream);
        sync(ctx);
    });
}

int nli_poisson_fill(float *img, const uint8_t *mask, int W, int H, int C)
{
    if (!img || !mask || W <= 0 || H <= 0 || C <= 0) return NLI_E_ARGUMENT;
    try {
        nli::poisson_fill(img, mask, W, H, C);
    } catch (...) {
        return NLI_E_INTERNAL;
    }
    return NLI_OK;
}

void nli_inpaint_params_default(nli_inpaint_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->method = NLI_UPD_MEANS;
    p->patch = 7;
    p->sigma = 10000.0f;
    p->lambda = 0.05f;
    p->lambda_rgb = 0.05f;
    p->lambda_d = 0.05f;
    p->iters = 200;
    p->tol = 0.1f;
    p->scales = 11;
    p->coarse = 0.1f;
    p->conf_t = 5.0f;
    p->conf_a = 0.1f;
    p->init = NLI_INIT_POISSON;
    p->pm_iters = 10;
    p->pm_shots = 30;
    p->pm_window = -1;
    p->cg_tol = 0.000001f;
    p->cg_limit = 1000;
    p->cg_precision = NLI_CG_FP64;
    p->cg_refine_every = 0;
    p->seed = 1;
}

int nli_inpaint_dev(nli_ctx *ctx, const nli_inpaint_params *params, const float *in, const uint8_t *mask, int W,
                    int H, int C, float *out, nli_inpaint_stats *stats)
{
    return guard(ctx, [&] {
        nli::InpaintConfig cfg = inpaint_config(params);
        need(in && mask && out, "bad inpainting arguments");
        nli::InpaintReport rep{};
        nli::inpaint(cfg, in, mask, W, H, C, out, rep, ctx->drv, ctx->pm, ctx->upd, ctx->stream);
        report(rep, stats);
    });
}

int nli_inpaint_batch(nli_ctx *ctx, const nli_inpaint_params *params, int n, const float *const *in,
                      const uint8_t *const *mask, const int *W, const int *H, int C, float *const *out, int workers,
                      nli_inpaint_stats *stats)
{
    return guard(ctx, [&] { run_batch(ctx, params, false, n, in, mask, W, H, C, out, workers, stats, true); });
}

int nli_inpaint_batch_each(nli_ctx *ctx, const nli_inpaint_params *params, int n, const float *const *in,
                           const uint8_t *const *mask, const int *W, const int *H, int C, float *const *out,
                           int workers, nli_inpaint_stats *stats)
{
    return guard(ctx, [&] { run_batch(ctx, params, true, n, in, mask, W, H, C, out, workers, stats, true); });
}

int nli_inpaint_batch_each_dev(nli_ctx *ctx, const nli_inpaint_params *params, int n, const float *const *in,
                               const uint8_t *const *mask, const int *W, const int *H, int C, float *const *out,
                               int workers, nli_inpaint_stats *stats)
{
    return guard(ctx, [&] { run_batch(ctx, params, true, n, in, mask, W, H, C, out, workers, stats, false); });
}

int nli_inpaint_batch_dev(nli_ctx *ctx, const nli_inpaint_params *params, int n, const float *const *in,
                          const uint8_t *const *mask, const int *W, const int *H, int C, float *const *out,
                          int workers, nli_inpaint_stats *stats)
{
    return guard(ctx, [&] { run_batch(ctx, params, false, n, in, mask, W, H, C, out, workers, stats, false); });
}

int nli_inpaint(nli_ctx *ctx, const nli_inpaint_params *params, const float *in, const uint8_t *mask, int W, int H,
                int C, float *out, nli_inpaint_stats *stats)
{
    return guard(ctx, [&] {
        nli::InpaintConfig cfg = inpaint_config(params);
        need(in && mask && out && W > 0 && H > 0 && C >= 1 && C <= 4, "bad inpainting arguments");
        const size_t n = (size_t)W * H;
        bool empty = true;
        for (size_t k = 0; k < n && empty; k++) empty = mask[k] == 0;
        need(!empty, "ERROR: Mask should not be empty (all black)");
        const float *d_in = stage_in(ctx, ctx->f0, in, n * C);
        const uint8_t *d_m = stage_in(ctx, ctx->b0, mask, n);
        float *d_out = ctx->f1.ensure(n * C);
        nli::InpaintReport rep{};
        nli::inpaint(cfg, d_in, d_m, W, H, C, d_out, rep, ctx->drv, ctx->pm, ctx->upd, ctx->stream);
        stage_out(ctx, out, (const float *)d_out, n * C);
        sync(ctx);
        report(rep, stats);
    });
}

}  // extern "C"

namespace {

// one colour conversion on the context's stream.  IOUtility returns images it does not
// convert unchanged (io_utility.cpp:270-272, 294-296); those are copied through.
int color_out_channels(bool to_lab, int C) { return to_lab ? C : ((C == 3 || C == 4) ? 3 : C); }

void color_dev(nli_ctx *ctx, bool to_lab, const float *in, int W, int H, int C, float *out)
{
    need(in && out && W > 0 && H > 0 && C >= 1, "bad colour conversion arguments");
    const long n = (long)W * H;
    if (to_lab && C == 3)
        nli::rgb_to_lab(in, nullptr, n, 3, out, ctx->stream);
    else if (!to_lab && (C == 3 || C == 4))
        nli::lab_to_rgb(in, n, C, out, nullptr, ctx->stream);
    else
        NLI_HIP(hipMemcpyAsync(out, in, sizeof(float) * (size_t)n * C, hipMemcpyDeviceToDevice, ctx->stream));
}

void color_host(nli_ctx *ctx, bool to_lab, const float *in, int W, int H, int C, float *out)
{
    need(in && out && W > 0 && H > 0 && C >= 1, "bad colour conversion arguments");
    const size_t n = (size_t)W * H;
    const float *d_in = stage_in(ctx, ctx->c0, in, n * C);
    float *d_out = ctx->c1.ensure(n * color_out_channels(to_lab, C));
    color_dev(ctx, to_lab, d_in, W, H, C, d_out);
    stage_out(ctx, out, (const float *)d_out, n * color_out_channels(to_lab, C));
    sync(ctx);
}

// main.cpp:115-309 around ImageInpainting::process, on device buffers: rgb (n x 3) and
// depth (n, optional) -> Lab (+ depth) in lab_in -> inpainting into lab_out -> rgb_out
// (n x 3) and depth_out (n, optional).  lab_in and lab_out hold n x (3 or 4) floats.
void inpaint_rgb(nli_ctx *ctx, const nli_inpaint_params *params, const float *rgb, const float *depth,
                 const uint8_t *mask, int W, int H, float *lab_in, float *lab_out, float *rgb_out,
                 float *depth_out, nli_inpaint_stats *stats)
{
    nli::InpaintConfig cfg = inpaint_config(params);
    need(rgb && mask && rgb_out && W > 0 && H > 0, "bad inpainting arguments");
    need(cfg.method != NLI_UPD_DEPTH || depth, "Specify depth input using -depth [filename]");
    const size_t n = (size_t)W * H;
    if (cfg.coarse == 0) {
        // main.cpp:166-178: the rule of thumb "patch side no smaller than twice the size of
        // the coarsest inpainting domain", from the largest distance to the hole boundary
        float *dt = ctx->f0.ensure(n);
        float *g = ctx->f1.ensure(n);
        int *st = ctx->pts.ensure(2 * n);
        nli::distance_transform(mask, W, H, dt, g, st, ctx->stream);
        std::vector<float> h(n);
        NLI_HIP(hipMemcpyAsync(h.data(), dt, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        float max_dist = 0;
        for (size_t k = 0; k < n; k++) max_dist = (h[k] > max_dist) ? h[k] : max_dist;
        cfg.coarse = std::min(1.5f * cfg.patch / max_dist, 1.f);
    }
    need(params->conf_t > 0.f, "ERROR: confidence_decay_time needs to be positive.");
    if (std::abs(cfg.coarse - 1.0f) < 0.00001f) cfg.scales = 1;   // main.cpp:200-203
    const int C = depth ? 4 : 3;
    nli::rgb_to_lab(rgb, depth, (long)n, C, lab_in, ctx->stream);
    nli::InpaintReport rep{};
    nli::inpaint(cfg, lab_in, mask, W, H, C, lab_out, rep, ctx->drv, ctx->pm, ctx->upd, ctx->stream);
    nli::lab_to_rgb(lab_out, (long)n, C, rgb_out, depth ? depth_out : nullptr, ctx->stream);
    report(rep, stats);
}

}  // namespace

extern "C" {

int nli_rgb_to_lab(nli_ctx *ctx, const float *rgb, int W, int H, int C, float *lab)
{
    return guard(ctx, [&] { color_host(ctx, true, rgb, W, H, C, lab); });
}

int nli_lab_to_rgb(nli_ctx *ctx, const float *lab, int W, int H, int C, float *rgb)
{
    return guard(ctx, [&] { color_host(ctx, false, lab, W, H, C, rgb); });
}

int nli_rgb_to_lab_dev(nli_ctx *ctx, const float *rgb, int W, int H, int C, float *lab)
{
    return guard(ctx, [&] {
        color_dev(ctx, true, rgb, W, H, C, lab);
        sync(ctx);
    });
}

int nli_lab_to_rgb_dev(nli_ctx *ctx, const float *lab, int W, int H, int C, float *rgb)
{
    return guard(ctx, [&] {
        color_dev(ctx, false, lab, W, H, C, rgb);
        sync(ctx);
    });
}

int nli_inpaint_rgb(nli_ctx *ctx, const nli_inpaint_params *params, const float *rgb, const float *depth,
                    const uint8_t *mask, int W, int H, float *rgb_out, float *depth_out, nli_inpaint_stats *stats)
{
    return guard(ctx, [&] {
        need(rgb && mask && rgb_out && W > 0 && H > 0 && (!depth || depth_out), "bad inpainting arguments");
        const size_t n = (size_t)W * H;
        bool empty = true;
        for (size_t k = 0; k < n && empty; k++) empty = mask[k] == 0;
        need(!empty, "ERROR: Mask should not be empty (all black)");
        const int C = depth ? 4 : 3;
        const float *d_rgb = stage_in(ctx, ctx->c2, rgb, n * 3);
        const float *d_depth = depth ? stage_in(ctx, ctx->c3, depth, n) : nullptr;
        const uint8_t *d_m = stage_in(ctx, ctx->b0, mask, n);
        float *lab_in = ctx->c0.ensure(n * C);
        float *lab_out = ctx->c1.ensure(n * C);
        // the RGB input and depth planes are read before the outputs are written
        inpaint_rgb(ctx, params, d_rgb, d_depth, d_m, W, H, lab_in, lab_out, ctx->c2.p, depth ? ctx->c3.p : nullptr,
                    stats);
        stage_out(ctx, rgb_out, (const float *)ctx->c2.p, n * 3);
        if (depth) stage_out(ctx, depth_out, (const float *)ctx->c3.p, n);
        sync(ctx);
    });
}

int nli_inpaint_rgb_dev(nli_ctx *ctx, const nli_inpaint_params *params, const float *rgb, const float *depth,
                        const uint8_t *mask, int W, int H, float *rgb_out, float *depth_out,
                        nli_inpaint_stats *stats)
{
    return guard(ctx, [&] {
        need(rgb && mask && rgb_out && W > 0 && H > 0 && (!depth || depth_out), "bad inpainting arguments");
        const size_t n = (size_t)W * H;
        const int C = depth ? 4 : 3;
        float *lab_in = ctx->c0.ensure(n * C);
        float *lab_out = ctx->c1.ensure(n * C);
        inpaint_rgb(ctx, params, rgb, depth, mask, W, H, lab_in, lab_out, rgb_out, depth_out, stats);
        sync(ctx);
    });
}

}  // extern "C"

static_assert(nli::kNoAllocScratchFloats * sizeof(float) == NLI_NOALLOC_SCRATCH_BYTES,
              "the header's scratch size is the library's");
