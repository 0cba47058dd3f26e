// k_pyramid.hip -- image pyramid down/up-sampling (gfx950).
//
// Reference: sampling.cpp:114-202 (downsample_with_mask), :304-385 (normalised
// convolution), :414-461 (9-tap separable convolution, taps accumulated i = 8..0,
// symmetric boundaries :550-562), :464-506 (bilinear on the centred coarse grid,
// double-precision blend on the axis-aligned cases), :205-246/:387-405 (nearest
// upsampling, Neumann boundaries).  Built with -ffp-contract=off, every float
// operation happens in the reference's order, so images and masks are
// bit-exact.  Two kernels per level: a full-resolution x pass over C+1 planes
// (mask-weighted channels and the known-pixel weight), then a y pass fused with
// the bilinear sampling at the coarse grid points.
#include "nli_internal.h"

#include <cmath>

namespace nli {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxC = 4;

struct Filt {
    float f[9];
};

__device__ inline int sym_a(int x, int size)
{
    while ((x < 0) || (x >= size)) {
        if (x < 0) x = -x - 1;
        if (x >= size) x = 2 * size - x - 1;
    }
    return x;
}

__device__ inline int neumann(int x, int size) { return max(0, min(size - 1, x)); }

// mode 0: planes c < C = conv_x(m ? img_c : 0), plane C = conv_x(m), m = (mask == mask_val)
// mode 1: planes c < C = conv_x(img_c)
// mode 2: plane 0 = conv_x(mask ? 1 : 0)
__global__ void k_xpass(const float *__restrict__ img, const uint8_t *__restrict__ mask, int W, int H, int C,
                        int mask_val, int mode, Filt fx, float *__restrict__ buf)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t plane = (size_t)W * H;
    size_t row = (size_t)W * y;
    float csum[kMaxC] = {0.0f, 0.0f, 0.0f, 0.0f};
    float msum = 0.0f;
    for (int i = 8; i >= 0; i--) {
        int id = sym_a(x + 4 - i, W);
        size_t q = row + id;
        float fi = fx.f[i];
        if (mode == 0) {
            float m = ((mask_val != 0) == (mask[q] != 0)) ? 1.0f : 0.0f;
            msum += fi * m;
            for (int c = 0; c < kMaxC; c++)
                if (c < C) {
                    float v = (m != 0.0f) ? img[q * C + c] : 0.0f;
                    csum[c] += fi * v;
                }
        } else if (mode == 1) {
            for (int c = 0; c < kMaxC; c++)
                if (c < C) csum[c] += fi * img[q * C + c];
        } else {
            float m = mask[q] ? 1.0f : 0.0f;
            msum += fi * m;
        }
    }
    size_t o = row + x;
    if (mode == 0) {
        for (int c = 0; c < kMaxC; c++)
            if (c < C) buf[c * plane + o] = csum[c];
        buf[C * plane + o] = msum;
    } else if (mode == 1) {
        for (int c = 0; c < kMaxC; c++)
            if (c < C) buf[c * plane + o] = csum[c];
    } else {
        buf[o] = msum;
    }
}

// y pass of the separable convolution at one fine-grid point
__device__ inline float smooth_y(const float *__restrict__ p, int W, int H, int xx, int yy, const Filt &fy)
{
    float sum = 0.0f;
    for (int i = 8; i >= 0; i--) {
        int id = sym_a(yy + 4 - i, H);
        sum += fy.f[i] * p[(size_t)id * W + xx];
    }
    return sum;
}

__device__ float bilinear_smoothed(const float *__restrict__ p, int W, int H, float x, float y, const Filt &fy)
{
    int id_x = (int)floorf(x);
    int id_y = (int)floorf(y);
    int x0 = neumann(id_x, W), y0 = neumann(id_y, H);
    int x1 = neumann(id_x + 1, W), y1 = neumann(id_y + 1, H);
    const float p11 = smooth_y(p, W, H, x0, y0, fy);
    float a = x - (float)id_x;
    float b = y - (float)id_y;
    float result;
    if ((!a) || (!b)) {
        if ((!a) && (!b)) {
            result = p11;
        } else if (!a) {
            const float p12 = smooth_y(p, W, H, x0, y1, fy);
            float bp = b * p12;
            result = (float)((1.0 - (double)b) * (double)p11 + (double)bp);
        } else {
            const float p21 = smooth_y(p, W, H, x1, y0, fy);
            float ap = a * p21;
            result = (float)((1.0 - (double)a) * (double)p11 + (double)ap);
        }
    } else {
        float a_1 = (float)(1.0 - (double)a);
        float b_1 = (float)(1.0 - (double)b);
        const float p21 = smooth_y(p, W, H, x1, y0, fy);
        const float p12 = smooth_y(p, W, H, x0, y1, fy);
        const float p22 = smooth_y(p, W, H, x1, y1, fy);
        result = b_1 * (a_1 * p11 + a * p21) + b * (a_1 * p12 + a * p22);
    }
    return result;
}

__global__ void k_ysample(const float *__restrict__ buf, int W, int H, int C, int sW, int sH, float factor_x,
                          float factor_y, float min_x, float min_y, int mode, int mask_val, float threshold,
                          Filt fy, float *__restrict__ out_img, uint8_t *__restrict__ out_mask)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= sW) return;
    size_t plane = (size_t)W * H;
    float y = min_y + (float)(unsigned)i * factor_y;
    float x = min_x + (float)(unsigned)j * factor_x;
    size_t o = (size_t)i * sW + j;
    if (mode == 0) {
        float mv = bilinear_smoothed(buf + C * plane, W, H, x, y, fy);
        for (int c = 0; c < C; c++) {
            float iv = bilinear_smoothed(buf + c * plane, W, H, x, y, fy);
            out_img[o * C + c] = ((double)mv > 1e-6) ? iv / mv : 0.f;
        }
        out_mask[o] = mask_val ? (mv > threshold) : (mv < threshold);
    } else if (mode == 1) {
        for (int c = 0; c < C; c++) out_img[o * C + c] = bilinear_smoothed(buf + c * plane, W, H, x, y, fy);
    } else {
        float mv = bilinear_smoothed(buf, W, H, x, y, fy);
        out_mask[o] = mv > threshold;
    }
}

__global__ void k_upsample(const float *__restrict__ in, int W, int H, int C, int W2, int H2, float factor_x,
                           float factor_y, float min_x, float min_y, float *__restrict__ out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= W2) return;
    float y = ((float)(unsigned)i - min_y) / factor_y;
    float x = ((float)(unsigned)j - min_x) / factor_x;
    int ix = neumann((int)floor((double)x + 0.5), W);
    int iy = neumann((int)floor((double)y + 0.5), H);
    for (int c = 0; c < C; c++) out[((size_t)i * W2 + j) * C + c] = in[((size_t)iy * W + ix) * C + c];
}

struct Geometry {
    int sW, sH;
    float factor_x, factor_y, min_x, min_y;
    Filt fx, fy;
};

Geometry make_geometry(int W, int H, float factor)
{
    Geometry g;
    g.sW = sample_size(W, factor);
    g.sH = sample_size(H, factor);
    g.factor_x = (float)W / (float)g.sW;
    g.factor_y = (float)H / (float)g.sH;
    float sigma_x = (float)(0.62 * std::sqrt((double)g.factor_x * (double)g.factor_x - 1));
    float sigma_y = (float)(0.62 * std::sqrt((double)g.factor_y * (double)g.factor_y - 1));
    gaussian_weights(9, 1, sigma_x, 0.0f, g.fx.f);
    gaussian_weights(9, 1, sigma_y, 0.0f, g.fy.f);
    g.min_x = (float)((double)g.factor_x / 2.0 - 0.5);
    g.min_y = (float)((double)g.factor_y / 2.0 - 0.5);
    return g;
}

inline float clamp_threshold(float t) { return (float)std::fmax(0.0, std::fmin(1.0, (double)t)); }

}  // namespace

void downsample_with_mask(const float *img, const uint8_t *mask, int W, int H, int C, float factor,
                          float threshold, int mask_val, float *out_img, uint8_t *out_mask, float *scratch,
                          hipStream_t s)
{
    if (C > kMaxC) throw std::invalid_argument("downsample: at most 4 channels");
    Geometry g = make_geometry(W, H, factor);
    dim3 gx((W + kBlock - 1) / kBlock, H);
    k_xpass<<<gx, kBlock, 0, s>>>(img, mask, W, H, C, mask_val, 0, g.fx, scratch);
    NLI_LAUNCH_CHECK();
    dim3 gy((g.sW + kBlock - 1) / kBlock, g.sH);
    k_ysample<<<gy, kBlock, 0, s>>>(scratch, W, H, C, g.sW, g.sH, g.factor_x, g.factor_y, g.min_x, g.min_y, 0,
                                    mask_val, clamp_threshold(threshold), g.fy, out_img, out_mask);
    NLI_LAUNCH_CHECK();
}

void downsample(const float *img, int W, int H, int C, float factor, float *out, float *scratch, hipStream_t s)
{
    if (C > kMaxC) throw std::invalid_argument("downsample: at most 4 channels");
    Geometry g = make_geometry(W, H, factor);
    dim3 gx((W + kBlock - 1) / kBlock, H);
    k_xpass<<<gx, kBlock, 0, s>>>(img, nullptr, W, H, C, 0, 1, g.fx, scratch);
    NLI_LAUNCH_CHECK();
    dim3 gy((g.sW + kBlock - 1) / kBlock, g.sH);
    k_ysample<<<gy, kBlock, 0, s>>>(scratch, W, H, C, g.sW, g.sH, g.factor_x, g.factor_y, g.min_x, g.min_y, 1, 0,
                                    0.0f, g.fy, out, nullptr);
    NLI_LAUNCH_CHECK();
}

void downsample_mask(const uint8_t *mask, int W, int H, float factor, float threshold, uint8_t *out,
                     float *scratch, hipStream_t s)
{
    Geometry g = make_geometry(W, H, factor);
    dim3 gx((W + kBlock - 1) / kBlock, H);
    k_xpass<<<gx, kBlock, 0, s>>>(nullptr, mask, W, H, 1, 0, 2, g.fx, scratch);
    NLI_LAUNCH_CHECK();
    dim3 gy((g.sW + kBlock - 1) / kBlock, g.sH);
    k_ysample<<<gy, kBlock, 0, s>>>(scratch, W, H, 1, g.sW, g.sH, g.factor_x, g.factor_y, g.min_x, g.min_y, 2, 0,
                                    clamp_threshold(threshold), g.fy, nullptr, out);
    NLI_LAUNCH_CHECK();
}

void upsample(const float *in, int W, int H, int C, int W2, int H2, float *out, hipStream_t s)
{
    float factor_x = (float)(unsigned)W2 / (float)(unsigned)W;
    float factor_y = (float)(unsigned)H2 / (float)(unsigned)H;
    float min_x = (float)((double)factor_x / 2.0 - 0.5);
    float min_y = (float)((double)factor_y / 2.0 - 0.5);
    dim3 g((W2 + kBlock - 1) / kBlock, H2);
    k_upsample<<<g, kBlock, 0, s>>>(in, W, H, C, W2, H2, factor_x, factor_y, min_x, min_y, out);
    NLI_LAUNCH_CHECK();
}

}  // namespace nli
