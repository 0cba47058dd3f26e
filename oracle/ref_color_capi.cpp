/*
 * ref_color_capi.cpp -- C entry points over the reference's colour conversion
 * (IOUtility::rgb_to_lab / lab_to_rgb, io_utility.cpp:254-303) and its PGM reader
 * and writer (io_utility.cpp:26-117), compiled with the
 * reference's io_utility.cpp where it lies (see oracle/Makefile) into
 * oracle/_ref/libnli_ref_color.so.  TEST INFRASTRUCTURE ONLY: pins the oracle's
 * colour restatement; never part of the product.
 *
 * io_utility.cpp also holds the iio image readers and writers, whose library (iio,
 * which needs libpng/jpeg/tiff) is absent.  The reference objects are compiled with
 * one section per function and hidden visibility, and the library is linked with
 * section garbage collection, so only the conversions, the PGM pair and what they
 * call are kept; nothing stands in for iio.  lab_to_rgb is exported for 3 channels only:
 * its 4-channel branch reads the RGB planes through a pointer into a local image
 * that is released at the end of that branch (io_utility.cpp:281-286).
 */
#include <cstring>
#include <stdexcept>

#include "image.h"
#include "io_utility.h"

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API int ref_rgb_to_lab(const float *rgb, int W, int H, float *lab)
{
    Image<float> in((uint)W, (uint)H, 3u);
    std::memcpy(in.raw(), rgb, sizeof(float) * (size_t)W * H * 3);
    Image<float> out = IOUtility::rgb_to_lab(in);
    std::memcpy(lab, out.raw(), sizeof(float) * (size_t)W * H * 3);
    return 0;
}

REF_API int ref_lab_to_rgb(const float *lab, int W, int H, float *rgb)
{
    Image<float> in((uint)W, (uint)H, 3u);
    std::memcpy(in.raw(), lab, sizeof(float) * (size_t)W * H * 3);
    Image<float> out = IOUtility::lab_to_rgb(in);
    std::memcpy(rgb, out.raw(), sizeof(float) * (size_t)W * H * 3);
    return 0;
}

/* IOUtility::read_pgm_image / write_pgm_image (io_utility.cpp:26-117), which use stdio
 * only.  ref_read_pgm returns -1 where the reference returns an empty image; the size
 * comes back through W and H, the samples into out when it is not NULL. */
REF_API int ref_read_pgm(const char *path, int *W, int *H, float *out, long capacity)
{
    Image<float> img = IOUtility::read_pgm_image(path);
    if (img.is_empty()) return -1;
    *W = (int)img.get_size_x();
    *H = (int)img.get_size_y();
    if (!out) return 0;
    if (capacity < (long)*W * *H) return -2;
    std::memcpy(out, img.raw(), sizeof(float) * (size_t)*W * *H);
    return 0;
}

REF_API int ref_write_pgm(const char *path, const float *img, int W, int H, int C)
{
    Image<float> in((uint)W, (uint)H, (uint)C);
    std::memcpy(in.raw(), img, sizeof(float) * (size_t)W * H * C);
    IOUtility::write_pgm_image(path, in);
    return 0;
}
