// nli_plain.cpp -- the reference's main.cpp composition over the C++ plugin classes
// (include/nli_plugins.hpp): the plain drop-in path, in which every PatchMatch, update,
// pyramid and mask step is one C-ABI call on host buffers.
//
// Reference: main.cpp:220-284 (parameters, method -> distance/updater pairs, PatchMatch
// construction, set_weights_updating / set_image_updating, process), and main.cpp:115-309
// around it for 0-255 RGB input (colour conversion, mask, depth, output).
#include "nli_plugins.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

namespace {

thread_local std::string g_err;

using namespace nli;

// main.cpp:220-284: the (distance, updater) pair, PatchMatch, ImageInpainting, process
Image<float> run_process(Context &c, const nli_inpaint_params *p, int scales, float coarse, Image<float> input,
                         const Mask &m, nli_inpaint_stats *stats, std::chrono::steady_clock::time_point t0)
{
    const Shape patch((unsigned)p->patch, (unsigned)p->patch);
    // main.cpp:245-263
    std::unique_ptr<APatchDistance> distance;
    std::unique_ptr<AImageUpdating> updating;
    switch (p->method) {
    case NLI_UPD_MEANS:
        updating.reset(new HipPatchNonLocalMeans(c, patch, p->sigma));
        distance.reset(new L2NormPatchDistance(patch, p->sigma));
        break;
    case NLI_UPD_MEDIANS:
        updating.reset(new HipPatchNonLocalMedians(c, patch, p->sigma));
        distance.reset(new L1NormPatchDistance(patch, p->sigma));
        break;
    case NLI_UPD_POISSON:
        updating.reset(new HipPatchNonLocalPoisson(c, patch, p->sigma, p->lambda, p->cg_tol, p->cg_limit,
                                                   p->cg_precision, p->cg_refine_every));
        distance.reset(new L2CombinedPatchDistance(p->lambda, patch, p->sigma));
        break;
    case NLI_UPD_DEPTH:
        updating.reset(new HipPatchNonLocalDepth(c, patch, p->sigma, p->lambda_rgb, p->lambda_d, 0.000001f, 1000,
                                                 p->cg_precision, p->cg_refine_every));
        distance.reset(new L2RGBDPatchDistance(p->lambda_rgb, p->lambda_d, patch, p->sigma));
        break;
    default: throw std::invalid_argument("ERROR: Unknown method name");
    }
    HipPatchMatch patch_match(c, distance.get(), p->pm_iters, p->pm_shots, p->pm_window, p->seed);
    const float rate = ImageInpainting::calculate_subsampling_rate(coarse, (unsigned)scales);
    ImageInpainting inpainting(c, p->iters, p->tol, scales, rate, p->conf_t, p->conf_a,
                               (ImageInpainting::InitType)p->init);
    inpainting.set_weights_updating(&patch_match);
    inpainting.set_image_updating(updating.get());
    Image<float> result = inpainting.process(input, m);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        const auto &its = inpainting.get_iterations_per_scale();
        stats->scales = (int)its.size();
        for (size_t i = 0; i < its.size() && i < 32; i++) stats->iterations[i] = its[i];
        stats->seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return result;
}

template <class F>
int plain_guard(F &&f)
{
    try {
        f();
        return NLI_OK;
    } catch (const EmptySourceError &e) {
        g_err = e.what();
        return NLI_E_EMPTY_SOURCE;
    } catch (const std::invalid_argument &e) {
        g_err = e.what();
        return NLI_E_ARGUMENT;
    } catch (const std::exception &e) {
        g_err = e.what();
        return NLI_E_INTERNAL;
    }
}

}  // namespace

extern "C" int nli_plain_inpaint(nli_ctx *ctx, const nli_inpaint_params *p, const float *in, const uint8_t *mask,
                                 int W, int H, int C, float *out, nli_inpaint_stats *stats)
{
    using namespace nli;
    if (!ctx || !p || !in || !mask || !out || W < 1 || H < 1 || C < 1) return NLI_E_ARGUMENT;
    return plain_guard([&] {
        const auto t0 = std::chrono::steady_clock::now();
        Context c(ctx);
        Image<float> input((unsigned)W, (unsigned)H, (unsigned)C);
        std::memcpy(input.raw(), in, sizeof(float) * (size_t)W * H * C);
        Mask m((unsigned)W, (unsigned)H);
        std::memcpy(m.raw(), mask, (size_t)W * H);
        Image<float> result = run_process(c, p, p->scales, p->coarse, input, m, stats, t0);
        std::memcpy(out, result.raw(), sizeof(float) * (size_t)W * H * C);
    });
}

// main.cpp:115-309 without file I/O and argument parsing: the 0-255 RGB input converted
// to Lab (IOUtility::rgb_to_lab), the mask taken where positive, the depth plane appended
// (IOUtility::cat), the parameter rules of main.cpp:166-203, process, and the result
// converted back (IOUtility::lab_to_rgb, IOUtility::get_depth).
extern "C" int nli_plain_inpaint_rgb(nli_ctx *ctx, const nli_inpaint_params *p, const float *rgb,
                                     const float *depth, const uint8_t *mask, int W, int H, float *rgb_out,
                                     float *depth_out, nli_inpaint_stats *stats)
{
    using namespace nli;
    if (!ctx || !p || !rgb || !mask || !rgb_out || W < 1 || H < 1 || (depth && !depth_out)) return NLI_E_ARGUMENT;
    return plain_guard([&] {
        const auto t0 = std::chrono::steady_clock::now();
        Context c(ctx);
        Image<float> rgb_img((unsigned)W, (unsigned)H, 3u);
        std::memcpy(rgb_img.raw(), rgb, sizeof(float) * (size_t)W * H * 3);
        Image<float> input = IOUtility::rgb_to_lab(c, rgb_img);   // main.cpp:117
        bool is_mask_empty = true;                                 // main.cpp:125-138
        Mask m((unsigned)W, (unsigned)H);
        for (unsigned y = 0; y < (unsigned)H; y++)
            for (unsigned x = 0; x < (unsigned)W; x++)
                if (mask[(size_t)W * y + x] > 0) {
                    m.mask(x, y);
                    is_mask_empty = false;
                }
        if (is_mask_empty) throw std::invalid_argument("ERROR: Mask should not be empty (all black)");
        if (depth) {                                               // main.cpp:140-147
            Image<float> d((unsigned)W, (unsigned)H);
            std::memcpy(d.raw(), depth, sizeof(float) * (size_t)W * H);
            input = IOUtility::cat(input, d);
        }
        float coarse = p->coarse;
        int scales = p->scales;
        if (coarse == 0) {                                         // main.cpp:166-178
            std::vector<float> dt((size_t)W * H);
            c.check(nli_distance_transform(ctx, m.raw(), W, H, dt.data()), "distance transform");
            float max_dist = 0;
            for (float v : dt) max_dist = (v > max_dist) ? v : max_dist;
            coarse = std::min(1.5f * p->patch / max_dist, 1.f);
        }
        if (p->conf_t <= 0.f) throw std::invalid_argument("ERROR: confidence_decay_time needs to be positive.");
        if (std::abs(coarse - 1.0f) < 0.00001f) scales = 1;        // main.cpp:200-203
        if (p->method == NLI_UPD_DEPTH && !depth)                  // main.cpp:256-257
            throw std::invalid_argument("Specify depth input using -depth [filename]");
        Image<float> output = run_process(c, p, scales, coarse, input, m, stats, t0);
        Image<float> out_rgb = IOUtility::lab_to_rgb(c, output);  // main.cpp:307
        std::memcpy(rgb_out, out_rgb.raw(), sizeof(float) * (size_t)W * H * 3);
        if (depth) {                                               // main.cpp:308-309
            Image<float> out_d = IOUtility::get_depth(output);
            std::memcpy(depth_out, out_d.raw(), sizeof(float) * (size_t)W * H);
        }
    });
}

extern "C" const char *nli_plain_last_error() { return g_err.c_str(); }
