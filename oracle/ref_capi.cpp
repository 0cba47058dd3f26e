/*
 * ref_capi.cpp -- C entry points over the reference's own classes, compiled
 * together with the reference sources where they lie (see oracle/Makefile) into
 * oracle/_ref/.  TEST INFRASTRUCTURE ONLY: used to pin the oracle against the
 * reference and as the "reference" CPU baseline; never part of the product.
 *
 * The reference's patch_non_local_depth.cpp is not compiled into this library (it
 * links io_utility.cpp, whose iio image I/O needs libpng/jpeg/tiff that are absent);
 * the NL-Depth row is composed here from the reference's own PatchNonLocalPoisson as
 * patch_non_local_depth.cpp:36-55 does, and the real updater is built separately
 * into libnli_ref_depth.so (ref_depth_capi.cpp).
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <limits>
#include <stdexcept>
#include <utility>
#include <vector>

/* expose private helpers of ImageInpainting (confidence, extended domain) to this wrapper */
#define private public
#define protected public
#include "image.h"
#include "mask.h"
#include "point.h"
#include "shape.h"
#include "gradient.h"
#include "gaussian_weights.h"
#include "distance_transform.h"
#include "sampling.h"
#include "a_patch_distance.h"
#include "l2_norm_patch_distance.h"
#include "l1_norm_patch_distance.h"
#include "l2_combined_patch_distance.h"
#include "l2_rgbd_patch_distance.h"
#include "patch_match.h"
#include "a_image_updating.h"
#include "patch_non_local_means.h"
#include "patch_non_local_medians.h"
#include "patch_non_local_poisson.h"
#include "image_inpainting.h"
#undef private
#undef protected

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

void from_nnf(const FixedImage<Point> &f, int32_t *n)
{
    int W = (int)f.get_size_x(), H = (int)f.get_size_y();
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t i = (size_t)W * y + x;
            Point p = f((uint)x, (uint)y);
            n[2 * i] = p.x;
            n[2 * i + 1] = p.y;
        }
}

void from_mask(const FixedMask &m, uint8_t *out)
{
    int W = (int)m.get_size_x(), H = (int)m.get_size_y();
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) out[(size_t)W * y + x] = m.get((uint)x, (uint)y) ? 1 : 0;
}

APatchDistance *make_distance(int kind, int P, float sigma, float lambda, float lambda_d)
{
    Shape ps((uint)P, (uint)P);
    switch (kind) {
    case 0: return new L2NormPatchDistance(ps, sigma);
    case 1: return new L1NormPatchDistance(ps, sigma);
    case 2: return new L2CombinedPatchDistance(lambda, ps, sigma);
    case 3: return new L2RGBDPatchDistance(lambda, lambda_d, ps, sigma);
    }
    return nullptr;
}

/* patch_non_local_depth.cpp:24-55, composed from the reference's NL-Poisson */
class DepthComposition : public AImageUpdating {
public:
    DepthComposition(Shape ps, float sigma, float lr, float ld)
        : AImageUpdating(ps, sigma), rgb_(ps, sigma, lr, 0.000001f, 1000), dep_(ps, sigma, ld, 0.000001f, 1000) {}
    double update(Image<float> image, Image<float>, FixedMask dom, FixedMask ext,
                  FixedImage<Point> nnf, FixedImage<float> conf) override
    {
        uint W = image.get_size_x(), H = image.get_size_y();
        Image<float> rgb(W, H, 3u), rgb_o(W, H, 3u), d(W, H), d_o(W, H);
        for (uint y = 0; y < H; y++)
            for (uint x = 0; x < W; x++) {
                for (uint c = 0; c < 3; c++) rgb(x, y, c) = rgb_o(x, y, c) = image(x, y, c);
                d(x, y) = d_o(x, y) = image(x, y, 3u);
            }
        double total = 0;
        total += rgb_.update(rgb, rgb_o, dom, ext, nnf, conf);
        total += dep_.update(d, d_o, dom, ext, nnf, conf);
        for (uint y = 0; y < H; y++)
            for (uint x = 0; x < W; x++) {
                for (uint c = 0; c < 3; c++) image(x, y, c) = rgb(x, y, c);
                image(x, y, 3u) = d(x, y);
            }
        return total;
    }
private:
    PatchNonLocalPoisson rgb_, dep_;
};

AImageUpdating *make_updater(int kind, int P, float sigma, float lambda, float lambda_d, float tol, int limit)
{
    Shape ps((uint)P, (uint)P);
    switch (kind) {
    case 0: return new PatchNonLocalMeans(ps, sigma);
    case 1: return new PatchNonLocalMedians(ps, sigma);
    case 2: return new PatchNonLocalPoisson(ps, sigma, lambda, tol, limit);
    case 3: return new DepthComposition(ps, sigma, lambda, lambda_d);
    }
    return nullptr;
}

}  // namespace

extern "C" {

int ref_is_openmp(void)
{
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}

void ref_gaussian_weights(int sx, int sy, float sigx, float sigy, float *out)
{
    Image<float> w = GaussianWeights::calculate((uint)sx, (uint)sy, sigx, sigy);
    std::memcpy(out, w.raw(), sizeof(float) * (size_t)sx * sy);
}

void ref_gradient(const float *img, int W, int H, int C, float *out)
{
    Image<float> g = Gradient::calculate(to_image(img, W, H, C));
    std::memcpy(out, g.raw(), sizeof(float) * (size_t)W * H * C * 2);
}

void ref_distance_transform(const uint8_t *m, int W, int H, float *out)
{
    Image<float> d = DistanceTransform::calculate(to_mask(m, W, H));
    std::memcpy(out, d.raw(), sizeof(float) * (size_t)W * H);
}

void ref_confidence(const uint8_t *m, int W, int H, float decay, float asym, float *out)
{
    ImageInpainting ii;
    Image<float> c = ii.calculate_confidence_mask(to_mask(m, W, H), decay, asym);
    std::memcpy(out, c.raw(), sizeof(float) * (size_t)W * H);
}

float ref_subsampling_rate(float ratio, int scales)
{
    return ImageInpainting::calculate_subsampling_rate(ratio, (unsigned)scales);
}

int ref_sample_size(int size, float factor)
{
    return Sampling::_Details::get_sample_size((uint)size, factor);
}

int ref_downsample_with_mask(const float *img, const uint8_t *m, int W, int H, int C, float factor,
                             float threshold, int mask_val, float *out_img, uint8_t *out_mask)
{
    std::pair<Image<float>, Mask> r =
        Sampling::downsample_with_mask(to_image(img, W, H, C), to_mask(m, W, H), factor, threshold, mask_val != 0);
    if (r.first.is_empty()) return -1;
    std::memcpy(out_img, r.first.raw(), sizeof(float) * r.first.get_size_x() * r.first.get_size_y() * C);
    from_mask(r.second, out_mask);
    return 0;
}

int ref_downsample(const float *img, int W, int H, int C, float factor, float *out)
{
    Image<float> r = Sampling::downsample(to_image(img, W, H, C), factor);
    if (r.is_empty()) return -1;
    std::memcpy(out, r.raw(), sizeof(float) * r.get_size_x() * r.get_size_y() * C);
    return 0;
}

int ref_downsample_mask(const uint8_t *m, int W, int H, float factor, float threshold, uint8_t *out)
{
    Mask r = Sampling::downsample(FixedMask(to_mask(m, W, H)), factor, threshold);
    if (r.is_empty()) return -1;
    from_mask(r, out);
    return 0;
}

void ref_upsample(const float *in, int W, int H, int C, int W2, int H2, float *out)
{
    Image<float> r = Sampling::upsample(to_image(in, W, H, C), Shape((uint)W2, (uint)H2));
    std::memcpy(out, r.raw(), sizeof(float) * (size_t)W2 * H2 * C);
}

void ref_extended_domain(const uint8_t *hole, int W, int H, int P, uint8_t *out)
{
    ImageInpainting ii;
    Mask e = ii.get_extended_domain(to_mask(hole, W, H), Shape((uint)P, (uint)P));
    from_mask(e, out);
}

/* Distances at a fixed NNF: out[i] for every target-mask point (FLT_MAX elsewhere).
 * Returns -1 when the distance throws (RGBD with C != 4). */
int ref_patch_distance_field(int kind, int P, float sigma, float lambda, float lambda_d,
                             const float *src, int sW, int sH, const float *tgt, int tW, int tH, int C,
                             const int32_t *nnf, const uint8_t *tmask, float *out)
{
    APatchDistance *d = make_distance(kind, P, sigma, lambda, lambda_d);
    d->initialize(to_image(src, sW, sH, C), to_image(tgt, tW, tH, C));
    int rc = 0;
    try {
        for (int y = 0; y < tH; y++)
            for (int x = 0; x < tW; x++) {
                size_t i = (size_t)tW * y + x;
                out[i] = std::numeric_limits<float>::max();
                if (!tmask[i] || nnf[2 * i] < 0) continue;
                out[i] = d->calculate(Point(nnf[2 * i], nnf[2 * i + 1]), Point(x, y));
            }
    } catch (const std::exception &) {
        rc = -1;
    }
    delete d;
    return rc;
}

/* PatchMatch::calculate with srand(seed); init_nnf NULL = empty initial field. */
int ref_patchmatch(int kind, int P, float sigma, float lambda, float lambda_d, int iters, int shots,
                   int window, const float *src, const uint8_t *smask, int sW, int sH, const float *tgt,
                   const uint8_t *tmask, int tW, int tH, int C, const int32_t *init_nnf, unsigned seed,
                   int32_t *out_nnf)
{
    APatchDistance *d = make_distance(kind, P, sigma, lambda, lambda_d);
    PatchMatch pm(d, iters, shots, window);
    srand(seed);
    Image<Point> init = init_nnf ? to_nnf(init_nnf, tW, tH) : Image<Point>();
    Image<Point> r = pm.calculate(to_image(src, sW, sH, C), to_mask(smask, sW, sH), to_image(tgt, tW, tH, C),
                                  to_mask(tmask, tW, tH), init);
    delete d;
    if (r.is_empty()) return -1;
    from_nnf(r, out_nnf);
    return 0;
}

/* AImageUpdating::update; image is updated in place, returns the total difference. */
double ref_update(int kind, int P, float sigma, float lambda, float lambda_d, float cg_tol, int cg_limit,
                  float *image, const float *orig, const uint8_t *domain, const uint8_t *ext,
                  const int32_t *nnf, const float *conf, int W, int H, int C)
{
    AImageUpdating *u = make_updater(kind, P, sigma, lambda, lambda_d, cg_tol, cg_limit);
    Image<float> img = to_image(image, W, H, C);
    Image<float> org = (orig == image) ? img : to_image(orig, W, H, C);
    FixedImage<float> cf = conf ? FixedImage<float>(to_image(conf, W, H, 1)) : FixedImage<float>();
    double r = u->update(img, org, to_mask(domain, W, H), to_mask(ext, W, H), to_nnf(nnf, W, H), cf);
    std::memcpy(image, img.raw(), sizeof(float) * (size_t)W * H * C);
    delete u;
    return r;
}

/* ImageInpainting::process with the main.cpp:245-270 pairing; srand(seed) once
 * (main.cpp:273).  method: 0 nlmeans, 1 nlmedians, 2 nlpoisson, 3 nldepth. */
int ref_process(int method, int P, float sigma, float lambda, float lambda_rgb, float lambda_d, int iters,
                float tol, int scales, float coarse, float conf_t, float conf_a, int init, int pm_iters,
                int pm_shots, int pm_window, float cg_tol, int cg_limit, unsigned seed,
                const float *in, const uint8_t *mask, int W, int H, int C, float *out)
{
    float rate = ImageInpainting::calculate_subsampling_rate(coarse, (unsigned)scales);
    ImageInpainting ii(iters, tol, scales, rate, conf_t, conf_a, (ImageInpainting::InitType)init);
    APatchDistance *d = nullptr;
    AImageUpdating *u = nullptr;
    switch (method) {
    case 0: d = make_distance(0, P, sigma, 0, 0); u = make_updater(0, P, sigma, 0, 0, 0, 0); break;
    case 1: d = make_distance(1, P, sigma, 0, 0); u = make_updater(1, P, sigma, 0, 0, 0, 0); break;
    case 2: d = make_distance(2, P, sigma, lambda, 0); u = make_updater(2, P, sigma, lambda, 0, cg_tol, cg_limit); break;
    case 3: d = make_distance(3, P, sigma, lambda_rgb, lambda_d); u = make_updater(3, P, sigma, lambda_rgb, lambda_d, 0, 0); break;
    default: return -1;
    }
    PatchMatch pm(d, pm_iters, pm_shots, pm_window);
    ii.set_weights_updating(&pm);
    ii.set_image_updating(u);
    srand(seed);
    int rc = 0;
    try {
        Image<float> r = ii.process(to_image(in, W, H, C), to_mask(mask, W, H));
        std::memcpy(out, r.raw(), sizeof(float) * (size_t)W * H * C);
    } catch (const std::exception &) {
        rc = -2;
    }
    delete d;
    delete u;
    return rc;
}

}  /* extern "C" */
