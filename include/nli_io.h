/* nli_io.h — image files for the inpainting flow: PNG reading and writing with the
 * sample semantics of the reference's image I/O (io_utility.cpp:119-220, which calls
 * 3rdparty/iio/iio.c).  Host code only, exported by libnli_plugins.so, zlib-based (no
 * libpng), and the reference's own stdio PGM reader and writer (io_utility.cpp:26-117).
 *
 * Images are row-major, interleaved (HWC) float32, as everywhere in nli.h.
 * Every function returns NLI_OK (0) or a negative code and never throws.  The message
 * of the calling thread's last failure is nli_io_last_error().
 *
 * Reading follows iio's PNG reader (iio.c:1071-1130: libpng with PNG_TRANSFORM_EXPAND
 * and PNG_TRANSFORM_PACKING):
 *   - palette images become RGB, or RGBA when a tRNS chunk is present;
 *   - grey samples of 1, 2 or 4 bits become 8-bit values (0..255, bits replicated);
 *   - a tRNS chunk on a grey or RGB image adds an alpha channel (0 where the pixel
 *     equals the tRNS colour, else the maximum);
 *   - 16-bit samples stay 16-bit (0..65535); 8-bit samples are 0..255;
 *   - Adam7-interlaced files are read too.
 * Writing follows iio_save_image_float_vec for a ".png" name (iio.c:3367-3515): 8-bit
 * samples, each the float v taken to (uint8)clamp(0.5 + (double)v, 0, 255) (NaN: 0),
 * grey / grey+alpha / RGB / RGBA by channel count, not interlaced.
 */
#ifndef NLI_IO_H
#define NLI_IO_H

#include "nli.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NLI_E_IO (-5) /* file missing, unreadable, or not a PNG this reader decodes */

/* Size of a PNG file as the reader returns it: width, height, channels after the
 * expansion above (1..4) and bit depth (8 or 16).  Any pointer may be NULL. */
int nli_image_info(const char *path, int *W, int *H, int *channels, int *bit_depth);

/* iio_read_image_float_vec (iio.c, as io_utility.cpp:146 calls it): W*H*channels floats
 * into out; capacity is out's length in floats. */
int nli_read_image(const char *path, float *out, long capacity);

/* IOUtility::read_rgb_image (io_utility.cpp:143-168): three channels per pixel.  One
 * channel is replicated; with two (grey + alpha) the channels are (grey, alpha, grey);
 * a fourth channel is dropped.  capacity >= W*H*3. */
int nli_read_rgb_image(const char *path, float *rgb, long capacity);

/* IOUtility::read_mono_image (io_utility.cpp:119-141 over iio_read_image_float,
 * iio.c:3241-3259): one channel.  8-bit RGB and RGBA become
 * (uint8)(.299 R + .587 G + .114 B) in double (iio.c:873-930); grey + alpha and 16-bit
 * colour are refused, as iio refuses them ("non-scalarizable image", "uncolorize type
 * not supported").  capacity >= W*H. */
int nli_read_mono_image(const char *path, float *out, long capacity);

/* iio_save_image_float_vec with a ".png" or ".PNG" name: C = 1..4 channels written as
 * 8-bit samples (rounding above).  Other names are refused (the reference picks TIFF,
 * PGM and others by suffix; only PNG is implemented). */
int nli_write_image(const char *path, const float *img, int W, int H, int C);

/* IOUtility::write_rgb_image (io_utility.cpp:171-190): channels 0..2 of an image with
 * C >= 3 channels as an RGB PNG.  C < 3 is refused: the reference reads channels 1 and
 * 2 of such an image from the neighbouring pixels and, at the last pixels, past the end
 * of the buffer. */
int nli_write_rgb_image(const char *path, const float *img, int W, int H, int C);

/* IOUtility::write_mono_image (io_utility.cpp:192-220): C = 1 as is; C = 3 as the
 * luminance (float)(0.2126 R + 0.7152 G + 0.0722 B) computed in double.  Other channel
 * counts are refused: the reference leaves those samples unset. */
int nli_write_mono_image(const char *path, const float *img, int W, int H, int C);

/* IOUtility::read_pgm_image (io_utility.cpp:26-74; stdio, no iio): "P2" (ASCII) or "P5"
 * (one byte per sample, whatever the maximum value says).  Whitespace and '#' comments
 * are skipped between the header fields and also before the samples, so binary samples
 * at the start that are whitespace bytes are skipped, as there; a binary file that ends
 * early gives -1 for each missing sample (fgetc's EOF).  nli_pgm_info gives the size;
 * capacity >= W*H.  The reference returns an empty image for a file it cannot open or
 * whose magic is not P2/P5 and exits on a bad ASCII sample: here both are NLI_E_IO. */
int nli_pgm_info(const char *path, int *W, int *H);
int nli_read_pgm_image(const char *path, float *out, long capacity);

/* IOUtility::write_pgm_image (io_utility.cpp:80-117): "P5", maximum 255, one byte per
 * pixel: (char)v for C = 1, (char)(0.2126 R + 0.7152 G + 0.0722 B) in double for C = 3,
 * the header alone for other C.  (char) of a value outside char is undefined in C++;
 * the bytes are those x86-64 builds of the reference write: truncation to a 32-bit
 * integer (0x80000000 for NaN and out-of-range values), its low byte. */
int nli_write_pgm_image(const char *path, const float *img, int W, int H, int C);

/* Message of the calling thread's last failed nli_io call ("" if none). */
const char *nli_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
