// nli_host.cpp -- host-side numerics of libnli: intra-patch Gaussian weights,
// pyramid rate and sizes, and the coarsest-level Poisson fill.
//
// These run once per object / per process() on tiny inputs, so they stay on the
// host and call the host libm exactly where the reference does (double exp for the
// weights, float logf/expf for the rate), which keeps them bit-identical to it.
#include "nli_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nli {

// gaussian_weights.cpp:19-80 (one segment per cell): the first quadrant (centre row
// and column included) is evaluated in double and stored as float, mirrored to the
// other quadrants, then normalised by its float sum.  The 1-D form has no factor 2
// in the exponent (:37), the 2-D form has it (:50).
void gaussian_weights(int size_x, int size_y, float sigma_x, float sigma_y, float *out)
{
    const unsigned sx = (unsigned)size_x, sy = (unsigned)size_y;
    const double x0 = -((float)sx / 2), y0 = -((float)sy / 2);
    const double vx = (double)sigma_x * (double)sigma_x, vy = (double)sigma_y * (double)sigma_y;
    std::fill(out, out + (size_t)sx * sy, 0.0f);
    for (unsigned i = 0; i <= sx / 2; i++) {
        const double x = x0 + i + 0.0 + 0.5;
        if (sy == 1) {
            out[i] = (float)(0.0 + std::exp(-(x * x / vx)));
            continue;
        }
        for (unsigned j = 0; j <= sy / 2; j++) {
            const double y = y0 + j + 0.0 + 0.5;
            out[sx * j + i] = (float)(0.0 + std::exp(-(x * x / (2 * vx) + y * y / (2 * vy))));
        }
    }
    for (unsigned i = 0; i <= sx / 2; i++)
        for (unsigned j = 0; j <= sy / 2; j++) {
            const float v = out[sx * j + i];
            out[sx * j + (sx - 1 - i)] = v;
            if (sy > 1) {
                out[sx * (sy - 1 - j) + i] = v;
                out[sx * (sy - 1 - j) + (sx - 1 - i)] = v;
            }
        }
    float sum = 0.0f;
    for (unsigned k = 0; k < sx * sy; k++) sum += out[k];
    for (unsigned k = 0; k < sx * sy; k++) out[k] /= sum;
}

// image_inpainting.cpp:197-204 (float overloads of log and exp)
float subsampling_rate(float ratio, int scales)
{
    if (ratio < 0.0 || ratio > 1.0) return 0.0f;
    return expf(logf(ratio) / (float)(unsigned)(scales - 1));
}

// sampling.cpp:408-411
int sample_size(int size, float factor) { return (int)((float)(unsigned)size * factor + 0.5); }

namespace {

// Multiscale Gauss-Seidel Laplace extension of the NaN pixels of one plane
// (3rdparty/simpois/simpois.c:228-263 with tstep 0.25, 10 sweeps, every octave, no
// data term and no CG refinement, the fixed arguments of :289).
inline float at_clamped(const float *I, int w, int h, int i, int j)
{
    i = std::min(std::max(i, 0), w - 1);
    j = std::min(std::max(j, 0), h - 1);
    return I[i + (size_t)j * w];
}

void fill_level(float *u, const float *g, int w, int h, int octaves)
{
    std::vector<float> init((size_t)w * h, 0.0f);
    if (octaves > 1 && (w > 1 || h > 1)) {
        // 2x2 block means over the finite samples (NaN when none), solve, bilinear back
        const int ws = (int)std::ceil(w / 2.0), hs = (int)std::ceil(h / 2.0);
        std::vector<float> gs((size_t)ws * hs), us((size_t)ws * hs);
        for (int j = 0; j < hs; j++)
            for (int i = 0; i < ws; i++) {
                const float s[4] = {at_clamped(g, w, h, 2 * i, 2 * j), at_clamped(g, w, h, 2 * i + 1, 2 * j),
                                    at_clamped(g, w, h, 2 * i, 2 * j + 1), at_clamped(g, w, h, 2 * i + 1, 2 * j + 1)};
                float m = 0;
                int cnt = 0;
                for (float v : s)
                    if (std::isfinite(v)) {
                        m += v;
                        cnt += 1;
                    }
                gs[(size_t)ws * j + i] = cnt ? m / cnt : NAN;
            }
        fill_level(us.data(), gs.data(), ws, hs, octaves - 1);
        for (int j = 0; j < h; j++)
            for (int i = 0; i < w; i++) {
                const float px = (float)((i - 0.5) / 2), py = (float)((j - 0.5) / 2);
                const int ip = (int)std::floor((double)px), iq = (int)std::floor((double)py);
                const float a = at_clamped(us.data(), ws, hs, ip, iq), b = at_clamped(us.data(), ws, hs, ip + 1, iq);
                const float c = at_clamped(us.data(), ws, hs, ip, iq + 1);
                const float d = at_clamped(us.data(), ws, hs, ip + 1, iq + 1);
                const float fx = px - (float)ip, fy = py - (float)iq;
                float r = 0;
                r += a * (1 - fx) * (1 - fy);
                r += b * (fx) * (1 - fy);
                r += c * (1 - fx) * (fy);
                r += d * (fx) * (fy);
                init[(size_t)w * j + i] = r;
            }
    }
    std::vector<int> hole;
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++)
            if (std::isnan(g[(size_t)j * w + i])) hole.push_back(j * w + i);
    for (size_t k = 0; k < (size_t)w * h; k++) u[k] = std::isfinite(g[k]) ? g[k] : init[k];
    const float tstep = 0.25f;
    for (int sweep = 0; sweep < 10; sweep++)
        for (int ij : hole) {
            const int i = ij % w, j = ij / w;
            const float x = at_clamped(u, w, h, i, j);
            const float nb[4] = {at_clamped(u, w, h, i + 1, j), at_clamped(u, w, h, i, j + 1),
                                 at_clamped(u, w, h, i - 1, j), at_clamped(u, w, h, i, j - 1)};
            float l = 0;
            for (float v : nb)
                if (std::isfinite(v)) l += v - x;
            const float d = 0;
            u[ij] = u[ij] + tstep * (l - d);
        }
}

}  // namespace

void poisson_fill(float *img, const uint8_t *mask, int W, int H, int C)
{
    const size_t n = (size_t)W * H;
    std::vector<float> g(n), u(n);
    for (int c = 0; c < C; c++) {
        for (size_t k = 0; k < n; k++) g[k] = mask[k] ? NAN : img[k * C + c];
        fill_level(u.data(), g.data(), W, H, 99);
        for (size_t k = 0; k < n; k++) img[k * C + c] = u[k];
    }
}

}  // namespace nli
