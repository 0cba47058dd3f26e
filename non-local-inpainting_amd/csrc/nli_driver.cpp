// nli_driver.cpp -- device-resident ImageInpainting::process.
//
// Reference: image_inpainting.cpp:68-190 (process), :349-417 (inpaint_internal),
// :492-557 (propagate_weights), main.cpp:245-266 (method -> distance/updater pairs).
// The image pyramid, the per-scale masks, the confidence mask, every PatchMatch and
// every update run on the device; the host keeps the control flow and reads back one
// scalar per inpainting iteration (the total difference of the stopping rule).  The
// coarsest-level Poisson fill runs on the device too; the black and average-colour
// initialisations run on the host.
#include "nli_internal.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <limits>
#include <vector>

namespace nli {

namespace {
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void inpaint(const InpaintConfig &cfg, const float *d_in, const uint8_t *d_mask, int W, int H, int C, float *d_out,
             InpaintReport &rep, DriverWork &dw, PMWork &pmw, UpdWork &uw, hipStream_t s)
{
    const double t_start = now_s();
    double t_pm = 0, t_upd = 0;
    rep = InpaintReport{};
    if (W < 1 || H < 1 || C < 1 || C > 4) throw std::invalid_argument("image must be W x H x C with 1 <= C <= 4");
    if (cfg.scales < 1 || cfg.scales > 32) throw std::invalid_argument("ERROR: number of scales must be at least 1.");
    if (cfg.patch <= 0) throw std::invalid_argument("ERROR: patch_side needs to be a positive integer.");

    // main.cpp:245-263 pairing
    DistDesc dd{DIST_L2, cfg.patch, cfg.sigma, 0.0f, 0.0f};
    switch (cfg.method) {
    case UPD_MEANS: dd.kind = DIST_L2; break;
    case UPD_MEDIANS: dd.kind = DIST_L1; break;
    case UPD_POISSON: dd.kind = DIST_L2COMBINED; dd.lambda = cfg.lambda; break;
    case UPD_DEPTH:
        if (C != 4) throw std::invalid_argument("L2 RGBD Distance: Channels do not match.");
        dd.kind = DIST_L2RGBD;
        dd.lambda = cfg.lambda_rgb;
        dd.lambda_d = cfg.lambda_d;
        break;
    default: throw std::invalid_argument("ERROR: Unknown method name");
    }
    const PMConfig pmc{cfg.pm_iters, cfg.pm_shots, cfg.pm_window, 0};
    const int S = cfg.scales;
    const int P = cfg.patch;
    const float rate = subsampling_rate(cfg.coarse, S);

    std::vector<int> ws(S), hs(S);
    ws[0] = W;
    hs[0] = H;
    for (int i = 1; i < S; i++) {
        ws[i] = sample_size(ws[i - 1], rate);
        hs[i] = sample_size(hs[i - 1], rate);
        if (ws[i] < 1 || hs[i] < 1) throw std::invalid_argument("pyramid level of zero size");
    }
    if ((int)dw.img.size() < S) {
        dw.img.resize(S);
        dw.mask.resize(S);
    }
    for (int i = 0; i < S; i++) {
        dw.img[i].ensure((size_t)ws[i] * hs[i] * C);
        dw.mask[i].ensure((size_t)ws[i] * hs[i]);
    }
    const size_t n0 = (size_t)W * H;
    NLI_HIP(hipMemcpyAsync(dw.img[0].p, d_in, n0 * C * sizeof(float), hipMemcpyDeviceToDevice, s));
    NLI_HIP(hipMemcpyAsync(dw.mask[0].p, d_mask, n0, hipMemcpyDeviceToDevice, s));
    dw.scratch.ensure(n0 * (C + 1));
    dw.conf.ensure(n0);
    dw.dtg.ensure(n0);
    dw.dtst.ensure(2 * n0);
    dw.up.ensure(n0 * C);
    dw.orig.ensure(n0 * C);
    dw.ext.ensure(n0);
    dw.tgt.ensure(n0);
    dw.src.ensure(n0);
    dw.bscr.ensure(n0);
    dw.hole_pts.ensure(n0);
    dw.low_pts.ensure(n0);
    dw.counts.ensure(n0 / 1024 + 2);
    dw.total.ensure(2);
    dw.nonempty.ensure(1);
    for (auto *b : {&dw.nnfA, &dw.nnfB, &dw.nnfC}) b->ensure(2 * n0);

    // ---- pyramid (:80-100) ----
    for (int i = 1; i < S; i++) {
        if (cfg.init != INIT_NONE) {
            downsample_with_mask(dw.img[i - 1].p, dw.mask[i - 1].p, ws[i - 1], hs[i - 1], C, rate, 0.4f, 0,
                                 dw.img[i].p, dw.mask[i].p, dw.scratch.p, s);
        } else {
            downsample(dw.img[i - 1].p, ws[i - 1], hs[i - 1], C, rate, dw.img[i].p, dw.scratch.p, s);
            downsample_mask(dw.mask[i - 1].p, ws[i - 1], hs[i - 1], rate, 0.4f, dw.mask[i].p, dw.scratch.p, s);
        }
    }

    // ---- coarsest-level initialisation (:104-154) ----
    if (cfg.init == INIT_POISSON) {
        const int w = ws[S - 1], h = hs[S - 1];
        dw.simpois.ensure(poisson_fill_scratch(w, h, C));
        poisson_fill_device(dw.img[S - 1].p, dw.mask[S - 1].p, w, h, C, dw.simpois.p, s);
    } else if (cfg.init != INIT_NONE) {   // black / average colour (host, as the reference)
        const int w = ws[S - 1], h = hs[S - 1];
        const size_t n = (size_t)w * h;
        std::vector<float> img(n * C);
        std::vector<uint8_t> m(n);
        NLI_HIP(hipMemcpyAsync(img.data(), dw.img[S - 1].p, n * C * sizeof(float), hipMemcpyDeviceToHost, s));
        NLI_HIP(hipMemcpyAsync(m.data(), dw.mask[S - 1].p, n, hipMemcpyDeviceToHost, s));
        NLI_HIP(hipStreamSynchronize(s));
        std::vector<float> color(C, 0.0f);
        if (cfg.init == INIT_AVG)
            for (int c = 0; c < C; c++) {
                float sum = 0.0f;
                int count = 0;
                for (size_t k = 0; k < n; k++)
                    if (!m[k]) {
                        sum += img[k * C + c];
                        count++;
                    }
                if (count > 0) color[c] = sum / (float)count;
            }
        for (size_t k = 0; k < n; k++)
            if (m[k])
                for (int c = 0; c < C; c++) img[k * C + c] = color[c];
        NLI_HIP(hipMemcpyAsync(dw.img[S - 1].p, img.data(), n * C * sizeof(float), hipMemcpyHostToDevice, s));
        NLI_HIP(hipStreamSynchronize(s));
    }

    // PatchMatch and update spans are timed with events on the stream (no host waits);
    // spans[k] = (start event, end event, is PatchMatch)
    struct Span {
        int a, b;
        bool pm;
    };
    std::vector<Span> spans;
    size_t nev = 0;
    auto event = [&]() {
        if (nev == dw.events.size()) {
            hipEvent_t e;
            NLI_HIP(hipEventCreate(&e));
            dw.events.push_back(e);
        }
        NLI_HIP(hipEventRecord(dw.events[nev], s));
        return (int)nev++;
    };
    uint64_t pm_calls = 0;
    // `start`: an event recorded just before (the end of the previous span, with nothing
    // enqueued since), shared as this span's start so that a span costs one marker
    auto run_pm = [&](const float *img, const uint8_t *sm, const uint8_t *tm, int w, int h, const int32_t *init,
                      int32_t *out, int start = -1) {
        const int e0 = start >= 0 ? start : event();
        patchmatch(dd, pmc, img, sm, w, h, img, tm, w, h, C, init, cfg.seed * 0x9E3779B97F4A7C15ull + (pm_calls++),
                   out, nullptr, pmw, s);
        spans.push_back({e0, event(), true});
        return spans.back().b;
    };
    auto run_update = [&](float *img, const float *orig, const uint8_t *hole, const int *pts, int npts,
                          const uint8_t *ext, const int32_t *nnf, const float *conf, int w, int h,
                          int start = -1) {
        const int e0 = start >= 0 ? start : event();
        double total = 0;
        CGStats st{};
        if (cfg.method == UPD_DEPTH) {
            const CGParams cgd{0.000001f, 1000, cfg.cg_precision, cfg.cg_refine_every};
            total += update(UPD_POISSON, P, cfg.sigma, cfg.lambda_rgb, cgd, img, img, C, 0, 3, hole, pts, npts, ext,
                            nnf, conf, w, h, uw, &st, s);
            for (int c = 0; c < 3; c++) rep.cg_iterations_total += st.iterations[c];
            total += update(UPD_POISSON, P, cfg.sigma, cfg.lambda_d, cgd, img, img, C, 3, 1, hole, pts, npts, ext,
                            nnf, conf, w, h, uw, &st, s);
            rep.cg_iterations_total += st.iterations[0];
        } else {
            const CGParams cg{cfg.cg_tol, cfg.cg_limit, cfg.cg_precision, cfg.cg_refine_every};
            total = update(cfg.method, P, cfg.sigma, cfg.lambda, cg, img, orig, C, 0, C, hole, pts, npts, ext, nnf,
                           conf, w, h, uw, &st, s);
            if (cfg.method == UPD_POISSON)
                for (int c = 0; c < C; c++) rep.cg_iterations_total += st.iterations[c];
        }
        spans.push_back({e0, event(), false});
        return total;
    };

    int32_t *nnf_cur = nullptr;   // NNF carried between iterations/scales
    int32_t *bufs[3] = {dw.nnfA.p, dw.nnfB.p, dw.nnfC.p};
    auto next_buf = [&](const int32_t *avoid1, const int32_t *avoid2) {
        for (int32_t *b : bufs)
            if (b != avoid1 && b != avoid2) return b;
        return bufs[0];
    };

    for (int L = S - 1; L >= 0; L--) {
        const int w = ws[L], h = hs[L];
        const size_t n = (size_t)w * h;
        float *img = dw.img[L].p;
        const uint8_t *hole = dw.mask[L].p;
        const int nh = compact_indices(hole, (int)n, dw.hole_pts.p, dw.counts.p, dw.total.p, s);
        points_changed(uw);
        rep.width[L] = w;
        rep.height[L] = h;
        rep.hole_pixels[L] = nh;

        // confidence mask (:161, :174, :611-630); the EDT is only needed for a positive decay
        if (cfg.conf_t > 0) distance_transform(hole, w, h, dw.dtg.p, dw.scratch.p, dw.dtst.p, s);
        confidence(hole, dw.dtg.p, w, h, cfg.conf_t, cfg.conf_a, dw.conf.p, s);

        if (L < S - 1) {
            // ---- propagate_weights (:492-557) ----
            const int lw = ws[L + 1], lh = hs[L + 1];
            const size_t ln = (size_t)lw * lh;
            const float *limg = dw.img[L + 1].p;
            const uint8_t *lhole = dw.mask[L + 1].p;
            extended_domain(lhole, lw, lh, P, dw.ext.p, dw.bscr.p, s);
            scale_masks(dw.ext.p, lw, lh, P / 2, 0, dw.tgt.p, dw.src.p, dw.nonempty.p, s);
            pm_masks_changed(pmw);
            int32_t *lnnf = next_buf(nullptr, nullptr);
            run_pm(limg, dw.src.p, dw.tgt.p, lw, lh, nullptr, lnnf);
            const int nl = compact_indices(lhole, (int)ln, dw.low_pts.p, dw.counts.p, dw.total.p, s);
            int32_t *snnf = next_buf(lnnf, nullptr);
            scale_nnf(lnnf, dw.low_pts.p, nl, lw, lh, w, h, snnf, s);
            upsample(limg, lw, lh, C, w, h, dw.up.p, s);
            copy_into_hole(dw.up.p, img, dw.hole_pts.p, nh, C, s);
            extended_domain(hole, w, h, P, dw.ext.p, dw.bscr.p, s);
            scale_masks(dw.ext.p, w, h, P / 2, 0, dw.tgt.p, dw.src.p, dw.nonempty.p, s);
            pm_masks_changed(pmw);
            int32_t *rnnf = next_buf(snnf, nullptr);
            run_pm(dw.up.p, dw.src.p, dw.tgt.p, w, h, snnf, rnnf);
            run_update(img, img, hole, dw.hole_pts.p, nh, dw.ext.p, rnnf, dw.conf.p, w, h);
            nnf_cur = rnnf;
        }

        // ---- inpaint_internal (:349-417) ----
        extended_domain(hole, w, h, P, dw.ext.p, dw.bscr.p, s);
        scale_masks(dw.ext.p, w, h, P / 2, 1, dw.tgt.p, dw.src.p, dw.nonempty.p, s);
        pm_masks_changed(pmw);
        int nonempty = 0;
        NLI_HIP(hipMemcpyAsync(&nonempty, dw.nonempty.p, sizeof(int), hipMemcpyDeviceToHost, s));
        NLI_HIP(hipStreamSynchronize(s));
        if (!nonempty) throw EmptySource();
        rep.target_pixels[L] = compact_indices(dw.tgt.p, (int)n, dw.low_pts.p, dw.counts.p, dw.total.p, s);
        NLI_HIP(hipMemcpyAsync(dw.orig.p, img, n * C * sizeof(float), hipMemcpyDeviceToDevice, s));
        double total = std::numeric_limits<double>::max();
        int it = 0;
        int mark = -1;   // end event of the previous span of this loop
        for (it = 0; it < cfg.iters && total > (double)cfg.tol; it++) {
            int32_t *out = next_buf(nnf_cur, nullptr);
            mark = run_pm(img, dw.src.p, dw.tgt.p, w, h, nnf_cur, out, mark);
            nnf_cur = out;
            total = run_update(img, dw.orig.p, hole, dw.hole_pts.p, nh, dw.ext.p, nnf_cur, dw.conf.p, w, h, mark);
            mark = spans.back().b;
        }
        rep.iterations[L] = it;
        rep.last_total[L] = total;
    }
    NLI_HIP(hipMemcpyAsync(d_out, dw.img[0].p, n0 * C * sizeof(float), hipMemcpyDeviceToDevice, s));
    NLI_HIP(hipStreamSynchronize(s));
    rep.scales = S;
    for (const Span &sp : spans) {
        float ms = 0.0f;
        NLI_HIP(hipEventElapsedTime(&ms, dw.events[sp.a], dw.events[sp.b]));
        (sp.pm ? t_pm : t_upd) += ms * 1e-3;
    }
    rep.seconds_total = now_s() - t_start;
    rep.seconds_patchmatch = t_pm;
    rep.seconds_update = t_upd;
    rep.pm_calls = (int)pm_calls;
}

}  // namespace nli

// ======================= several images in lockstep =======================
namespace nli {
namespace {

// main.cpp:245-263 pairing of method and patch distance
DistDesc distance_for(const InpaintConfig &cfg, int C)
{
    DistDesc dd{DIST_L2, cfg.patch, cfg.sigma, 0.0f, 0.0f};
    switch (cfg.method) {
    case UPD_MEANS: dd.kind = DIST_L2; break;
    case UPD_MEDIANS: dd.kind = DIST_L1; break;
    case UPD_POISSON: dd.kind = DIST_L2COMBINED; dd.lambda = cfg.lambda; break;
    case UPD_DEPTH:
        if (C != 4) throw std::invalid_argument("L2 RGBD Distance: Channels do not match.");
        dd.kind = DIST_L2RGBD;
        dd.lambda = cfg.lambda_rgb;
        dd.lambda_d = cfg.lambda_d;
        break;
    default: throw std::invalid_argument("ERROR: Unknown method name");
    }
    return dd;
}

// per image, host side
struct Img {
    int W, H;
    std::vector<int> ws, hs;
    uint64_t pm_calls = 0;
    int nh = 0;                      // hole points of the current scale
    int32_t *nnf_cur = nullptr;      // the field carried into the iterations
    int *active, *count, *cg_total;  // device state
    double *total, *partial;
    int *host;                       // pinned [active, stamp]
};

}  // namespace

bool lockstep_supported(const InpaintConfig &cfg, int W, int H, int C)
{
    if (cfg.method != UPD_POISSON && cfg.method != UPD_DEPTH) return false;
    if (cfg.method == UPD_DEPTH && C != 4) return false;
    if (cfg.init == INIT_BLACK || cfg.init == INIT_AVG) return false;
    const DistDesc dd = distance_for(cfg, C);
    const PMConfig pmc{cfg.pm_iters, cfg.pm_shots, cfg.pm_window, 0};
    return pm_batch_supported(dd, C, W, H, pmc);
}

void inpaint_lockstep(const InpaintConfig &cfg, int n, const float *const *d_in, const uint8_t *const *d_mask,
                      const int *Wv, const int *Hv, int C, float *const *d_out, InpaintReport *rep, LockstepWork &lw,
                      hipStream_t s)
{
    const double t_start = now_s();
    if (n <= 0) return;
    if (C < 1 || C > 4) throw std::invalid_argument("image must be W x H x C with 1 <= C <= 4");
    if (cfg.scales < 1 || cfg.scales > 32) throw std::invalid_argument("ERROR: number of scales must be at least 1.");
    if (cfg.patch <= 0) throw std::invalid_argument("ERROR: patch_side needs to be a positive integer.");
    for (int i = 0; i < n; i++)
        if (!lockstep_supported(cfg, Wv[i], Hv[i], C))
            throw std::invalid_argument("lockstep batch: unsupported parameters or size");
    const DistDesc dd = distance_for(cfg, C);
    const PMConfig pmc{cfg.pm_iters, cfg.pm_shots, cfg.pm_window, 0};
    const int S = cfg.scales, P = cfg.patch;
    const float rate = subsampling_rate(cfg.coarse, S);
    while ((int)lw.img.size() < n) lw.img.push_back(std::make_unique<LockstepImage>());
    lw.state.ensure((size_t)4 * n);
    lw.totals.ensure((size_t)2 * n);
    if (lw.host_slots < (size_t)n) {   // pinned [active, stamp] per image, grown with the group
        NLI_HIP(hipStreamSynchronize(s));
        if (lw.host) NLI_HIP(hipHostFree(lw.host));
        lw.host = nullptr;
        NLI_HIP(hipHostMalloc((void **)&lw.host, sizeof(int) * 2 * n, hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(lw.host, 0, sizeof(int) * 2 * n);
        NLI_HIP(hipHostGetDevicePointer((void **)&lw.host_dev, lw.host, 0));
        lw.host_slots = (size_t)n;
    }
    fill_i32(lw.state.p, 0, (size_t)4 * n, s);
    std::vector<Img> im(n);
    for (int i = 0; i < n; i++) {
        Img &g = im[i];
        rep[i] = InpaintReport{};
        g.W = Wv[i];
        g.H = Hv[i];
        g.ws.assign(S, 0);
        g.hs.assign(S, 0);
        g.ws[0] = g.W;
        g.hs[0] = g.H;
        for (int l = 1; l < S; l++) {
            g.ws[l] = sample_size(g.ws[l - 1], rate);
            g.hs[l] = sample_size(g.hs[l - 1], rate);
            if (g.ws[l] < 1 || g.hs[l] < 1) throw std::invalid_argument("pyramid level of zero size");
        }
        g.active = lw.state.p + 4 * i;
        g.count = g.active + 1;
        g.cg_total = g.active + 2;
        g.total = lw.totals.p + 2 * i;
        g.partial = g.total + 1;
        g.host = lw.host_dev + 2 * i;
        DriverWork &dw = lw.img[i]->dw;
        if ((int)dw.img.size() < S) {
            dw.img.resize(S);
            dw.mask.resize(S);
        }
        for (int l = 0; l < S; l++) {
            dw.img[l].ensure((size_t)g.ws[l] * g.hs[l] * C);
            dw.mask[l].ensure((size_t)g.ws[l] * g.hs[l]);
        }
        const size_t n0 = (size_t)g.W * g.H;
        NLI_HIP(hipMemcpyAsync(dw.img[0].p, d_in[i], n0 * C * sizeof(float), hipMemcpyDeviceToDevice, s));
        NLI_HIP(hipMemcpyAsync(dw.mask[0].p, d_mask[i], n0, hipMemcpyDeviceToDevice, s));
        dw.scratch.ensure(n0 * (C + 1));
        dw.conf.ensure(n0);
        dw.dtg.ensure(n0);
        dw.dtst.ensure(2 * n0);
        dw.up.ensure(n0 * C);
        dw.orig.ensure(n0 * C);
        dw.ext.ensure(n0);
        dw.tgt.ensure(n0);
        dw.src.ensure(n0);
        dw.bscr.ensure(n0);
        dw.hole_pts.ensure(n0);
        dw.low_pts.ensure(n0);
        dw.counts.ensure(n0 / 1024 + 2);
        dw.total.ensure(2);
        dw.nonempty.ensure(1);
        for (auto *b : {&dw.nnfA, &dw.nnfB, &dw.nnfC}) b->ensure(2 * n0);
        // pyramid (image_inpainting.cpp:80-100) and coarsest-level initialisation (:104-154)
        for (int l = 1; l < S; l++) {
            if (cfg.init != INIT_NONE) {
                downsample_with_mask(dw.img[l - 1].p, dw.mask[l - 1].p, g.ws[l - 1], g.hs[l - 1], C, rate, 0.4f, 0,
                                     dw.img[l].p, dw.mask[l].p, dw.scratch.p, s);
            } else {
                downsample(dw.img[l - 1].p, g.ws[l - 1], g.hs[l - 1], C, rate, dw.img[l].p, dw.scratch.p, s);
                downsample_mask(dw.mask[l - 1].p, g.ws[l - 1], g.hs[l - 1], rate, 0.4f, dw.mask[l].p, dw.scratch.p,
                                s);
            }
        }
        if (cfg.init == INIT_POISSON) {
            const int w = g.ws[S - 1], h = g.hs[S - 1];
            dw.simpois.ensure(poisson_fill_scratch(w, h, C));
            poisson_fill_device(dw.img[S - 1].p, dw.mask[S - 1].p, w, h, C, dw.simpois.p, s);
        }
    }
    auto next_buf = [](DriverWork &dw, const int32_t *avoid1, const int32_t *avoid2) {
        for (int32_t *b : {dw.nnfA.p, dw.nnfB.p, dw.nnfC.p})
            if (b != avoid1 && b != avoid2) return b;
        return dw.nnfA.p;
    };
    std::vector<PMJob> pj(n);
    auto pm_job = [&](int i, const float *img, const uint8_t *sm, const uint8_t *tm, int w, int h,
                      const int32_t *init, int32_t *out, const int *active) {
        Img &g = im[i];
        pj[i] = PMJob{dd, pmc, img, sm, w, h, img, tm, w, h, C, init,
                      cfg.seed * 0x9E3779B97F4A7C15ull + (g.pm_calls++), out, nullptr, &lw.img[i]->pm, active};
    };
    // the update jobs of one iteration (one, or two for NL-Depth: RGB then depth)
    std::vector<UpdJob> uj(n), uj2(n);
    auto upd_jobs = [&](int i, float *img, const float *orig, const int *pts, int npts, const uint8_t *ext,
                        const int32_t *nnf, const float *conf, int w, int h, bool loop) {
        Img &g = im[i];
        DriverWork &dw = lw.img[i]->dw;
        StopRule st;
        st.cg_total = g.cg_total;
        if (loop) {
            st.active = g.active;
            st.count = g.count;
            st.total = g.total;
            st.partial = g.partial;
            st.iters = cfg.iters;
            st.tol = cfg.tol;
            st.host = g.host;
            st.stamp = lw.stamp;
        }
        const int *act = loop ? g.active : nullptr;
        if (cfg.method == UPD_DEPTH) {
            const CGParams cgd{0.000001f, 1000, cfg.cg_precision, cfg.cg_refine_every};
            StopRule s1 = st, s2 = st;
            s1.part = loop ? 1 : 0;
            s1.host = nullptr;
            s2.part = loop ? 2 : 0;
            uj[i] = UpdJob{UPD_POISSON, P, cfg.sigma, cfg.lambda_rgb, cgd, img, img, C, 0, 3, dw.mask[0].p, pts, npts,
                           ext, nnf, conf, w, h, &lw.img[i]->uw, act, s1};
            uj2[i] = UpdJob{UPD_POISSON, P, cfg.sigma, cfg.lambda_d, cgd, img, img, C, 3, 1, dw.mask[0].p, pts, npts,
                            ext, nnf, conf, w, h, &lw.img[i]->uw, act, s2};
        } else {
            const CGParams cg{cfg.cg_tol, cfg.cg_limit, cfg.cg_precision, cfg.cg_refine_every};
            uj[i] = UpdJob{cfg.method, P, cfg.sigma, cfg.lambda, cg, img, orig, C, 0, C, dw.mask[0].p, pts, npts, ext,
                           nnf, conf, w, h, &lw.img[i]->uw, act, st};
        }
    };
    auto run_updates = [&]() {
        update_batch(uj.data(), n, lw.ub, s);
        if (cfg.method == UPD_DEPTH) update_batch(uj2.data(), n, lw.ub, s);
    };

    for (int L = S - 1; L >= 0; L--) {
        for (int i = 0; i < n; i++) {
            Img &g = im[i];
            DriverWork &dw = lw.img[i]->dw;
            const int w = g.ws[L], h = g.hs[L];
            const size_t np = (size_t)w * h;
            g.nh = compact_indices(dw.mask[L].p, (int)np, dw.hole_pts.p, dw.counts.p, dw.total.p, s);
            points_changed(lw.img[i]->uw);
            rep[i].width[L] = w;
            rep[i].height[L] = h;
            rep[i].hole_pixels[L] = g.nh;
            // confidence mask (:161, :174, :611-630)
            if (cfg.conf_t > 0) distance_transform(dw.mask[L].p, w, h, dw.dtg.p, dw.scratch.p, dw.dtst.p, s);
            confidence(dw.mask[L].p, dw.dtg.p, w, h, cfg.conf_t, cfg.conf_a, dw.conf.p, s);
        }
        if (L < S - 1) {
            // ---- propagate_weights (:492-557) ----
            for (int i = 0; i < n; i++) {   // PatchMatch on the lower level from a random field (:509)
                Img &g = im[i];
                DriverWork &dw = lw.img[i]->dw;
                const int lw_ = g.ws[L + 1], lh = g.hs[L + 1];
                extended_domain(dw.mask[L + 1].p, lw_, lh, P, dw.ext.p, dw.bscr.p, s);
                scale_masks(dw.ext.p, lw_, lh, P / 2, 0, dw.tgt.p, dw.src.p, dw.nonempty.p, s);
                pm_masks_changed(lw.img[i]->pm);
                pm_job(i, dw.img[L + 1].p, dw.src.p, dw.tgt.p, lw_, lh, nullptr, next_buf(dw, nullptr, nullptr),
                       nullptr);
            }
            patchmatch_batch(pj.data(), n, lw.pmb, s);
            for (int i = 0; i < n; i++) {   // scaled field, upsampled image, refine on the upper level (:512-551)
                Img &g = im[i];
                DriverWork &dw = lw.img[i]->dw;
                const int w = g.ws[L], h = g.hs[L], lw_ = g.ws[L + 1], lh = g.hs[L + 1];
                int32_t *lnnf = pj[i].out_nnf;
                const int nl = compact_indices(dw.mask[L + 1].p, lw_ * lh, dw.low_pts.p, dw.counts.p, dw.total.p, s);
                int32_t *snnf = next_buf(dw, lnnf, nullptr);
                scale_nnf(lnnf, dw.low_pts.p, nl, lw_, lh, w, h, snnf, s);
                upsample(dw.img[L + 1].p, lw_, lh, C, w, h, dw.up.p, s);
                copy_into_hole(dw.up.p, dw.img[L].p, dw.hole_pts.p, g.nh, C, s);
                extended_domain(dw.mask[L].p, w, h, P, dw.ext.p, dw.bscr.p, s);
                scale_masks(dw.ext.p, w, h, P / 2, 0, dw.tgt.p, dw.src.p, dw.nonempty.p, s);
                pm_masks_changed(lw.img[i]->pm);
                pm_job(i, dw.up.p, dw.src.p, dw.tgt.p, w, h, snnf, next_buf(dw, snnf, nullptr), nullptr);
            }
            patchmatch_batch(pj.data(), n, lw.pmb, s);
            for (int i = 0; i < n; i++) {   // one update (:554)
                Img &g = im[i];
                DriverWork &dw = lw.img[i]->dw;
                g.nnf_cur = pj[i].out_nnf;
                upd_jobs(i, dw.img[L].p, dw.img[L].p, dw.hole_pts.p, g.nh, dw.ext.p, g.nnf_cur, dw.conf.p, g.ws[L],
                         g.hs[L], false);
            }
            run_updates();
        } else {
            for (int i = 0; i < n; i++) im[i].nnf_cur = nullptr;
        }

        // ---- inpaint_internal (:349-417) ----
        for (int i = 0; i < n; i++) {
            Img &g = im[i];
            DriverWork &dw = lw.img[i]->dw;
            const int w = g.ws[L], h = g.hs[L];
            const size_t np = (size_t)w * h;
            extended_domain(dw.mask[L].p, w, h, P, dw.ext.p, dw.bscr.p, s);
            scale_masks(dw.ext.p, w, h, P / 2, 1, dw.tgt.p, dw.src.p, dw.nonempty.p, s);
            pm_masks_changed(lw.img[i]->pm);
            int nonempty = 0;
            NLI_HIP(hipMemcpyAsync(&nonempty, dw.nonempty.p, sizeof(int), hipMemcpyDeviceToHost, s));
            NLI_HIP(hipStreamSynchronize(s));
            if (!nonempty) throw EmptySource();
            rep[i].target_pixels[L] = compact_indices(dw.tgt.p, (int)np, dw.low_pts.p, dw.counts.p, dw.total.p, s);
            NLI_HIP(hipMemcpyAsync(dw.orig.p, dw.img[L].p, np * C * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        // every image starts the scale's iterations (the first always runs)
        for (int i = 0; i < n; i++) {
            fill_i32(im[i].active, cfg.iters > 0 ? 1 : 0, 1, s);
            fill_i32(im[i].count, 0, 1, s);
        }
        std::vector<uint64_t> pm_base(n);
        for (int i = 0; i < n; i++) pm_base[i] = im[i].pm_calls;
        const unsigned stamp0 = lw.stamp + 1;
        int it = 0;
        for (it = 0; it < cfg.iters; it++) {
            // the device's stopping rule, read without waiting: after iteration it - 2 at
            // the latest (the host is at most one iteration ahead of the GPU here)
            if (it >= 2) {
                bool any = false;
                for (int i = 0; i < n && !any; i++) {
                    const volatile int *h = lw.host + 2 * i;
                    const unsigned want = stamp0 + (unsigned)(it - 2);
                    const auto t0 = std::chrono::steady_clock::now();
                    while ((unsigned)h[1] - want > 0x7FFFFFFFu) {   // stamp older than want
                        __builtin_ia32_pause();
                        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
                            NLI_HIP(hipStreamSynchronize(s));
                            if ((unsigned)h[1] - want > 0x7FFFFFFFu)
                                throw std::runtime_error("stopping rule: the device did not write its flag");
                        }
                    }
                    std::atomic_thread_fence(std::memory_order_acquire);
                    any = h[0] != 0;
                }
                if (!any) break;
            }
            lw.stamp = stamp0 + (unsigned)it;
            for (int i = 0; i < n; i++) {
                Img &g = im[i];
                DriverWork &dw = lw.img[i]->dw;
                int32_t *out = next_buf(dw, g.nnf_cur, nullptr);
                pm_job(i, dw.img[L].p, dw.src.p, dw.tgt.p, g.ws[L], g.hs[L], g.nnf_cur, out, g.active);
                g.pm_calls = pm_base[i] + (uint64_t)it + 1;
                pj[i].seed = cfg.seed * 0x9E3779B97F4A7C15ull + pm_base[i] + (uint64_t)it;
            }
            patchmatch_batch(pj.data(), n, lw.pmb, s);
            for (int i = 0; i < n; i++) {
                Img &g = im[i];
                DriverWork &dw = lw.img[i]->dw;
                g.nnf_cur = pj[i].out_nnf;
                upd_jobs(i, dw.img[L].p, dw.orig.p, dw.hole_pts.p, g.nh, dw.ext.p, g.nnf_cur, dw.conf.p, g.ws[L],
                         g.hs[L], true);
            }
            run_updates();
        }
        lw.stamp = stamp0 + (unsigned)it;
        // the scale's report (one readback per scale): iterations run, last total
        std::vector<int> st(4 * n);
        std::vector<double> tt(2 * n);
        NLI_HIP(hipMemcpyAsync(st.data(), lw.state.p, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, s));
        NLI_HIP(hipMemcpyAsync(tt.data(), lw.totals.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
        NLI_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < n; i++) {
            rep[i].iterations[L] = st[4 * i + 1];
            rep[i].last_total[L] = st[4 * i + 1] > 0 ? tt[2 * i] : std::numeric_limits<double>::max();
            // the PatchMatch calls actually made: one per iteration run
            im[i].pm_calls = pm_base[i] + (uint64_t)st[4 * i + 1];
        }
    }
    std::vector<int> st(4 * n);
    NLI_HIP(hipMemcpyAsync(st.data(), lw.state.p, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, s));
    for (int i = 0; i < n; i++) {
        const size_t n0 = (size_t)Wv[i] * Hv[i];
        NLI_HIP(hipMemcpyAsync(d_out[i], lw.img[i]->dw.img[0].p, n0 * C * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    NLI_HIP(hipStreamSynchronize(s));
    const double secs = now_s() - t_start;
    for (int i = 0; i < n; i++) {
        rep[i].scales = S;
        rep[i].cg_iterations_total = st[4 * i + 2];
        rep[i].pm_calls = (int)im[i].pm_calls;
        rep[i].seconds_total = secs;
    }
}

}  // namespace nli
