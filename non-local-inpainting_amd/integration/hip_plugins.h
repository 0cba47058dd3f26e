// hip_plugins.h -- adapters that plug libnli into the reference program's own classes.
//
// A maintainer-side file: it is compiled inside the reference's source tree, against its
// headers (patch_match.h, a_image_updating.h, image.h, mask.h, point.h, shape.h) and
// include/nli.h, and linked with libnli.so.  It needs one edit of the reference:
// PatchMatch::calculate (patch_match.h:50-54) declared virtual, so that
// ImageInpainting's calls through its PatchMatch pointer (image_inpainting.cpp:391, 509,
// 551) reach the adapter.  Then ImageInpainting::set_weights_updating() /
// set_image_updating() (image_inpainting.cpp:285-300) take these objects in place of
// `new PatchMatch(...)` and the updater of main.cpp:245-266, and the rest of the reference
// runs unchanged.
//
// The reference containers have the C-ABI's layout (float HWC raw(), one byte per mask
// pixel, Point as two int32), so raw() pointers pass straight through.
#ifndef NLI_HIP_PLUGINS_H_
#define NLI_HIP_PLUGINS_H_

#include <cstdint>
#include <stdexcept>
#include <string>

#include "a_image_updating.h"
#include "image.h"
#include "mask.h"
#include "nli.h"
#include "patch_match.h"
#include "point.h"
#include "shape.h"

// PatchMatch::calculate on the GPU.  The distance (kind, patch, sigma, lambdas) is given
// as nli_dist_params because the reference's APatchDistance keeps lambda private
// (l2_combined_patch_distance.h:40-43).  Call k draws from seed * 0x9E3779B97F4A7C15 + k,
// the sequence libnli's own drivers use, so that runs are reproducible.
class HipPatchMatchAdapter : public PatchMatch {
public:
    HipPatchMatchAdapter(nli_ctx *ctx, nli_dist_params d, int iteration_count, int random_shots_limit = 20,
                         int search_window_size = -1, uint64_t seed = 1)
        : PatchMatch(iteration_count, random_shots_limit, search_window_size), ctx_(ctx), d_(d), seed_(seed)
    {
    }

    Image<Point> calculate(FixedImage<float> source, FixedMask source_mask, FixedImage<float> target,
                           FixedMask target_mask, Image<Point> initial_field = Image<Point>()) override
    {
        // patch_match.cpp:71-76: inconsistent sizes give an empty field
        if ((!initial_field.is_empty() && initial_field.get_size() != target.get_size()) ||
            source.get_size() != source_mask.get_size() || target.get_size() != target_mask.get_size())
            return Image<Point>();
        nli_pm_params pm = {get_iteration_count(), get_random_shots_limit(), get_search_window_size(), 0};
        Image<Point> out(target.get_size_x(), target.get_size_y(), Point(-1, -1));
        const int rc = nli_patchmatch(
            ctx_, &d_, &pm, source.raw(), reinterpret_cast<const uint8_t *>(source_mask.raw()),
            (int)source.get_size_x(), (int)source.get_size_y(), target.raw(),
            reinterpret_cast<const uint8_t *>(target_mask.raw()), (int)target.get_size_x(),
            (int)target.get_size_y(), (int)target.get_number_of_channels(),
            initial_field.is_empty() ? nullptr : reinterpret_cast<const int32_t *>(initial_field.raw()),
            seed_ * 0x9E3779B97F4A7C15ull + calls_++, reinterpret_cast<int32_t *>(out.raw()), nullptr);
        if (rc != NLI_OK) throw std::runtime_error(std::string("PatchMatch: ") + nli_last_error(ctx_));
        return out;
    }

private:
    nli_ctx *ctx_;
    nli_dist_params d_;
    uint64_t seed_, calls_ = 0;
};

// AImageUpdating::update on the GPU for the four updaters (u.kind: NLI_UPD_MEANS,
// _MEDIANS, _POISSON, _DEPTH; lambda / lambda_d and the CG settings as the reference's
// constructors take them, main.cpp:245-263).
class HipUpdatingAdapter : public AImageUpdating {
public:
    HipUpdatingAdapter(nli_ctx *ctx, nli_upd_params u)
        : AImageUpdating(Shape((uint)u.patch, (uint)u.patch), u.sigma), ctx_(ctx), u_(u)
    {
    }

    double update(Image<float> image, Image<float> orig_image, FixedMask inpainting_domain,
                  FixedMask extended_inpainting_domain, FixedImage<Point> nnf,
                  FixedImage<float> confidence_mask) override
    {
        double total = 0.0;
        const int rc = nli_update(
            ctx_, &u_, image.raw(), orig_image.raw(), reinterpret_cast<const uint8_t *>(inpainting_domain.raw()),
            reinterpret_cast<const uint8_t *>(extended_inpainting_domain.raw()),
            reinterpret_cast<const int32_t *>(nnf.raw()), confidence_mask.is_empty() ? nullptr : confidence_mask.raw(),
            (int)image.get_size_x(), (int)image.get_size_y(), (int)image.get_number_of_channels(), &total, nullptr);
        if (rc != NLI_OK) throw std::runtime_error(std::string("update: ") + nli_last_error(ctx_));
        return total;
    }

private:
    nli_ctx *ctx_;
    nli_upd_params u_;
};

#endif
