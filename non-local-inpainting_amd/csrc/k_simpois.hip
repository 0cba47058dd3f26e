// k_simpois.hip -- coarsest-level Poisson fill on the device (gfx950).
//
// Reference: 3rdparty/simpois/simpois.c:228-291 as called by image_inpainting.cpp:125-153
// (poisson_solver_separable_simplest: per channel, octave recursion down to 1x1 with
// 2x2 block means over the known samples, bilinear zoom-in of the coarser solution as
// the initial guess, then 10 in-place Gauss-Seidel sweeps of the Neumann Laplacian with
// time step 0.25 over the unknown pixels in row-major order, :82-100).
//
// One workgroup per channel runs the whole recursion.  The in-place row-major
// Gauss-Seidel sweep is reproduced exactly by anti-diagonal wavefronts: pixel (i, j)
// reads (i-1, j) and (i, j-1), which the sequential sweep has already updated and which
// lie on the previous diagonal, and (i+1, j), (i, j+1), which it has not, on the next
// one.  So all unknowns of one diagonal are updated together, with the same float
// operations as the sequential sweep, and the result is bit-identical.
#include "nli_internal.h"

#include <cmath>
#include <vector>

namespace nli {
namespace {

constexpr int kThreads = 1024;
constexpr int kMaxLevels = 32;

struct Pyramid {
    int levels;
    int w[kMaxLevels], h[kMaxLevels];
    size_t off[kMaxLevels];   // level offset within one channel's g (and u) area
    size_t per_channel;       // floats of one channel's g area (u follows it)
};

__device__ inline float clamped(const float *I, int w, int h, int i, int j)
{
    i = min(max(i, 0), w - 1);
    j = min(max(j, 0), h - 1);
    return I[i + (size_t)j * w];
}

__global__ __launch_bounds__(kThreads) void k_simpois(float *__restrict__ img, const uint8_t *__restrict__ mask,
                                                      int C, Pyramid py, float *__restrict__ scratch)
{
    const int c = blockIdx.x;
    float *G = scratch + (size_t)c * 2 * py.per_channel;   // known samples, NaN = unknown
    float *U = G + py.per_channel;                          // solution per level
    const int W = py.w[0], H = py.h[0];
    const int tid = threadIdx.x;
    // level 0: the image channel with NaN in the hole (image_inpainting.cpp:133-140)
    for (size_t k = tid; k < (size_t)W * H; k += kThreads) G[k] = mask[k] ? NAN : img[k * C + c];
    __syncthreads();
    // 2x2 block means over the finite samples (simpois.c:152-175)
    for (int l = 1; l < py.levels; l++) {
        const int iw = py.w[l - 1], ih = py.h[l - 1], ow = py.w[l], oh = py.h[l];
        const float *in = G + py.off[l - 1];
        float *out = G + py.off[l];
        for (int k = tid; k < ow * oh; k += kThreads) {
            const int i = k % ow, j = k / ow;
            const float a[4] = {clamped(in, iw, ih, 2 * i, 2 * j), clamped(in, iw, ih, 2 * i + 1, 2 * j),
                                clamped(in, iw, ih, 2 * i, 2 * j + 1), clamped(in, iw, ih, 2 * i + 1, 2 * j + 1)};
            float m = 0;
            int cx = 0;
            for (int q = 0; q < 4; q++)
                if (isfinite(a[q])) {
                    m += a[q];
                    cx += 1;
                }
            out[k] = cx ? m / cx : NAN;
        }
        __syncthreads();
    }
    for (int l = py.levels - 1; l >= 0; l--) {
        const int w = py.w[l], h = py.h[l];
        const float *g = G + py.off[l];
        float *u = U + py.off[l];
        // initial guess: 0 at the coarsest level, else the bilinear zoom-in of the
        // coarser solution (simpois.c:188-226); known samples keep their value
        for (int k = tid; k < w * h; k += kThreads) {
            float init = 0.0f;
            if (l + 1 < py.levels) {
                const int i = k % w, j = k / w;
                const int ws = py.w[l + 1], hs = py.h[l + 1];
                const float *us = U + py.off[l + 1];
                const float px = (float)((i - 0.5) / 2), pq = (float)((j - 0.5) / 2);
                const int ip = (int)floor((double)px), iq = (int)floor((double)pq);
                const float a = clamped(us, ws, hs, ip, iq), b = clamped(us, ws, hs, ip + 1, iq);
                const float cc = clamped(us, ws, hs, ip, iq + 1), d = clamped(us, ws, hs, ip + 1, iq + 1);
                const float fx = px - (float)ip, fy = pq - (float)iq;
                float r = 0;
                r += a * (1 - fx) * (1 - fy);
                r += b * (fx) * (1 - fy);
                r += cc * (1 - fx) * (fy);
                r += d * (fx) * (fy);
                init = r;
            }
            u[k] = isfinite(g[k]) ? g[k] : init;
        }
        __syncthreads();
        // 10 Gauss-Seidel sweeps, tstep 0.25, no data term, by anti-diagonals
        for (int sweep = 0; sweep < 10; sweep++)
            for (int diag = 0; diag <= (w - 1) + (h - 1); diag++) {
                const int i_lo = max(0, diag - (h - 1)), i_hi = min(diag, w - 1);
                for (int i = i_lo + tid; i <= i_hi; i += kThreads) {
                    const int j = diag - i;
                    const size_t ij = (size_t)j * w + i;
                    if (!isnan(g[ij])) continue;
                    const float x = u[ij];
                    const float nb[4] = {clamped(u, w, h, i + 1, j), clamped(u, w, h, i, j + 1),
                                         clamped(u, w, h, i - 1, j), clamped(u, w, h, i, j - 1)};
                    float lap = 0;
                    for (int q = 0; q < 4; q++)
                        if (isfinite(nb[q])) lap += nb[q] - x;
                    const float dterm = 0;
                    u[ij] = u[ij] + 0.25f * (lap - dterm);
                }
                __syncthreads();
            }
    }
    for (size_t k = tid; k < (size_t)W * H; k += kThreads) img[k * C + c] = U[k];
}

Pyramid make_pyramid(int W, int H)
{
    Pyramid py{};
    int w = W, h = H, octaves = 99;
    size_t off = 0;
    for (;;) {
        if (py.levels == kMaxLevels) throw std::invalid_argument("poisson fill: image too large");
        py.w[py.levels] = w;
        py.h[py.levels] = h;
        py.off[py.levels] = off;
        off += (size_t)w * h;
        py.levels++;
        // simpois.c:234: recurse while octaves remain and the image is not 1x1
        if (!(octaves > 1 && (w > 1 || h > 1))) break;
        w = (int)std::ceil(w / 2.0);
        h = (int)std::ceil(h / 2.0);
        octaves--;
    }
    py.per_channel = off;
    return py;
}

}  // namespace

size_t poisson_fill_scratch(int W, int H, int C) { return 2 * make_pyramid(W, H).per_channel * (size_t)C; }

void poisson_fill_device(float *img, const uint8_t *mask, int W, int H, int C, float *scratch, hipStream_t s)
{
    if (W < 1 || H < 1 || C < 1) throw std::invalid_argument("poisson fill: empty image");
    const Pyramid py = make_pyramid(W, H);
    k_simpois<<<C, kThreads, 0, s>>>(img, mask, C, py, scratch);
    NLI_LAUNCH_CHECK();
}

}  // namespace nli
