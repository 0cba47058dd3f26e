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
