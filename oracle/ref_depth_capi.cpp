/*
 * ref_depth_capi.cpp -- a C entry point over the reference's own NL-Depth updater
 * (PatchNonLocalDepth, patch_non_local_depth.cpp:24-55), compiled with the reference's
 * sources where they lie (see oracle/Makefile) into oracle/_ref/libnli_ref_depth.so.
 * TEST INFRASTRUCTURE ONLY: pins the oracle's NL-Depth (two NL-Poisson solves on
 * channel subsets, merged back) against the reference; never part of the product.
 *
 * The updater splits and merges channels with IOUtility::separate / cat, which live in
 * io_utility.cpp beside the iio readers and writers (iio needs libpng/jpeg/tiff, absent
 * here).  As for the colour conversions, the objects are compiled with one section per
 * function and hidden visibility and linked with section garbage collection, so only
 * what the updater calls is kept; nothing stands in for iio.
 */
#include <cstdint>
#include <cstring>

#include "image.h"
#include "mask.h"
#include "point.h"
#include "shape.h"
#include "patch_non_local_depth.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
Image<float> to_image(const float *p, int W, int H, int C)
{
    Image<float> im((uint)W, (uint)H, (uint)C);
    std::memcpy(im.raw(), p, sizeof(float) * (size_t)W * H * C);
    return im;
}
Mask to_mask(const uint8_t *m, int W, int H)
{
    Mask mk((uint)W, (uint)H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            if (m[(size_t)W * y + x]) mk.mask((uint)x, (uint)y);
    return mk;
}
Image<Point> to_nnf(const int32_t *n, int W, int H)
{
    Image<Point> f((uint)W, (uint)H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t i = (size_t)W * y + x;
            f((uint)x, (uint)y) = Point(n[2 * i], n[2 * i + 1]);
        }
    return f;
}
}  // namespace

/* PatchNonLocalDepth(Shape(P, P), sigma, lambda_rgb, lambda_d).update on a 4-channel image,
 * updated in place; orig is passed as the reference's original_image (the updater does not
 * read it, patch_non_local_depth.cpp:48-49).  Returns the total difference. */
REF_API double ref_depth_update(int P, float sigma, float lambda_rgb, float lambda_d, float *image,
                                const float *orig, const uint8_t *domain, const uint8_t *ext,
                                const int32_t *nnf, const float *conf, int W, int H)
{
    PatchNonLocalDepth u(Shape((uint)P, (uint)P), sigma, lambda_rgb, lambda_d);
    Image<float> img = to_image(image, W, H, 4);
    Image<float> org = to_image(orig, W, H, 4);
    FixedImage<float> cf = conf ? FixedImage<float>(to_image(conf, W, H, 1)) : FixedImage<float>();
    double r = u.update(img, org, to_mask(domain, W, H), to_mask(ext, W, H), to_nnf(nnf, W, H), cf);
    std::memcpy(image, img.raw(), sizeof(float) * (size_t)W * H * 4);
    return r;
}
