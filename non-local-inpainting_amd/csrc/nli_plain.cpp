// nli_plain.cpp -- the reference's main.cpp composition over the C++ plugin classes
// (include/nli_plugins.hpp): the plain drop-in path, in which every PatchMatch, update,
// pyramid and mask step is one C-ABI call on host buffers.
//
// Reference: main.cpp:220-284 (parameters, method -> distance/updater pairs, PatchMatch
// construction, set_weights_updating / set_image_updating, process).
#include "nli_plugins.hpp"

#include <chrono>
#include <cstring>
#include <memory>
#include <string>

namespace {

thread_local std::string g_err;

}  // namespace

extern "C" int nli_plain_inpaint(nli_ctx *ctx, const nli_inpaint_params *p, const float *in, const uint8_t *mask,
                                 int W, int H, int C, float *out, nli_inpaint_stats *stats)
{
    using namespace nli;
    if (!ctx || !p || !in || !mask || !out || W < 1 || H < 1 || C < 1) return NLI_E_ARGUMENT;
    try {
        const auto t0 = std::chrono::steady_clock::now();
        Context c(ctx);
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
        const float rate = ImageInpainting::calculate_subsampling_rate(p->coarse, (unsigned)p->scales);
        ImageInpainting inpainting(c, p->iters, p->tol, p->scales, rate, p->conf_t, p->conf_a,
                                   (ImageInpainting::InitType)p->init);
        inpainting.set_weights_updating(&patch_match);
        inpainting.set_image_updating(updating.get());

        Image<float> input((unsigned)W, (unsigned)H, (unsigned)C);
        std::memcpy(input.raw(), in, sizeof(float) * (size_t)W * H * C);
        Mask m((unsigned)W, (unsigned)H);
        std::memcpy(m.raw(), mask, (size_t)W * H);
        Image<float> result = inpainting.process(input, m);
        std::memcpy(out, result.raw(), sizeof(float) * (size_t)W * H * C);
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            const auto &its = inpainting.get_iterations_per_scale();
            stats->scales = (int)its.size();
            for (size_t i = 0; i < its.size() && i < 32; i++) stats->iterations[i] = its[i];
            stats->seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
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

extern "C" const char *nli_plain_last_error() { return g_err.c_str(); }
