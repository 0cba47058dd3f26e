/*
 * nli_oracle.cpp -- CPU oracle (clean-room restatement) of the reference's
 * inner inpainting loop.  TEST INFRASTRUCTURE ONLY: it is the checker of the
 * HIP path and the timed CPU baseline, never part of the product.
 *
 * Written as free functions over flat arrays in C-style C++ (g++), so that the
 * weighted median can use the same std::sort as the reference.  Every function
 * cites the reference file:line it restates.  Floating-point expressions keep
 * the reference's float/double types and evaluation order, so that on x86-64
 * (SSE, no FMA contraction: built with -ffp-contract=off) results match the
 * reference bit for bit.
 *
 * Not restated: glibc libm and rand()/srand() (called, not re-implemented),
 * libgomp scheduling, and the simpois environment "smart parameters"
 * (PMSFAC/PONLIT/CG_EPS, 3rdparty/simpois/smapa.h): the defaults are used.
 */
#include "nli_oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::vector<float> fvec;
typedef std::vector<uint8_t> bvec;
typedef std::vector<int32_t> ivec;

struct Pt { int x, y; };

/* mask.cpp:223-230 -- out-of-range test() is false (unsigned comparison) */
inline bool mtest(const uint8_t *m, int W, int H, int x, int y)
{
    if ((unsigned)x >= (unsigned)W || (unsigned)y >= (unsigned)H) return false;
    return m[(size_t)W * y + x] != 0;
}

/* mask.cpp:322-329, 459-486 -- masked points in row-major order */
std::vector<Pt> masked_points(const uint8_t *m, int W, int H)
{
    std::vector<Pt> pts;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            if (m[(size_t)W * y + x]) pts.push_back(Pt{x, y});
    return pts;
}

/* mask.cpp:459-486 bounding box; (-1,-1) for an empty mask */
void bounding_box(const uint8_t *m, int W, int H, Pt &tl, Pt &br)
{
    tl = Pt{W, H};
    br = Pt{-1, -1};
    bool any = false;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            if (m[(size_t)W * y + x]) {
                any = true;
                tl.x = std::min(tl.x, x); tl.y = std::min(tl.y, y);
                br.x = std::max(br.x, x); br.y = std::max(br.y, y);
            }
    if (!any) { tl = Pt{-1, -1}; br = Pt{-1, -1}; }
}

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* sampling.cpp:550-562 symmetric_boundary_condition_A */
inline int sym_a(int x, int size)
{
    while ((x < 0) || (x >= size)) {
        if (x < 0) x = -x - 1;
        if (x >= size) x = 2 * size - x - 1;
    }
    return x;
}

/* sampling.cpp:526-529 */
inline int neumann(int x, int size) { return std::max(0, std::min(size - 1, x)); }

/* sampling.cpp:414-461 -- 9-tap separable convolution, taps accumulated i = K-1 .. 0 */
void separate_convolution(const float *in, float *out, int W, int H,
                          const float *fx, const float *fy, int kx, int ky)
{
    fvec buf((size_t)W * H);
    int radius = (kx - 1) / 2;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float sum = 0.0f;
            for (int i = kx - 1; i >= 0; i--) {
                int id = sym_a(x + radius - i, W);
                sum += fx[i] * in[(size_t)y * W + id];
            }
            buf[(size_t)y * W + x] = sum;
        }
    radius = (ky - 1) / 2;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float sum = 0.0f;
            for (int i = ky - 1; i >= 0; i--) {
                int id = sym_a(y + radius - i, H);
                sum += fy[i] * buf[(size_t)id * W + x];
            }
            out[(size_t)y * W + x] = sum;
        }
}

/* sampling.cpp:464-506 -- note the double-precision blend on the axis-aligned cases */
float bilinear(const float *in, int W, int H, float x, float y)
{
    int id_x = (int)std::floor(x);
    int id_y = (int)std::floor(y);
    int x0 = neumann(id_x, W), y0 = neumann(id_y, H);
    int x1 = neumann(id_x + 1, W), y1 = neumann(id_y + 1, H);
    const float p11 = in[x0 + (size_t)W * y0];
    float a = x - (float)id_x;
    float b = y - (float)id_y;
    float result = 0.0f;
    if ((!a) || (!b)) {
        if ((!a) && (!b)) {
            result = p11;
        } else if (!a) {
            const float p12 = in[x0 + (size_t)W * y1];
            float bp = b * p12;
            result = (float)((1.0 - (double)b) * (double)p11 + (double)bp);
        } else {
            const float p21 = in[x1 + (size_t)W * y0];
            float ap = a * p21;
            result = (float)((1.0 - (double)a) * (double)p11 + (double)ap);
        }
    } else {
        float a_1 = (float)(1.0 - (double)a);
        float b_1 = (float)(1.0 - (double)b);
        const float p21 = in[x1 + (size_t)W * y0];
        const float p12 = in[x0 + (size_t)W * y1];
        const float p22 = in[x1 + (size_t)W * y1];
        result = b_1 * (a_1 * p11 + a * p21) + b * (a_1 * p12 + a * p22);
    }
    return result;
}

/* sampling.cpp:509-521 nearest (round = floor(x + 0.5) in double) */
inline float nearest(const float *in, int W, int H, float x, float y)
{
    int ix = (int)std::floor((double)x + 0.5);
    int iy = (int)std::floor((double)y + 0.5);
    ix = neumann(ix, W);
    iy = neumann(iy, H);
    return in[ix + (size_t)W * iy];
}

/* sampling.cpp:250-262 / 321-332 -- sigma and the two 9-tap 1-D kernels */
void pyramid_filters(float factor_x, float factor_y, float *fx, float *fy)
{
    float sigma_x = (float)(0.62 * std::sqrt((double)factor_x * (double)factor_x - 1));
    float sigma_y = (float)(0.62 * std::sqrt((double)factor_y * (double)factor_y - 1));
    nlo_gaussian_weights(9, 1, sigma_x, 0.0f, fx);
    nlo_gaussian_weights(9, 1, sigma_y, 0.0f, fy);
}

/* sampling.cpp:304-385 downsample_internal_with_mask, one channel */
void downsample_internal_with_mask(const float *img_in, const float *msk_in, float *img_out,
                                   float *msk_out, int W, int H, int sW, int sH)
{
    float factor_x = (float)W / (float)sW;
    float factor_y = (float)H / (float)sH;
    size_t n = (size_t)W * H;
    fvec img_x_msk(n), msk_f(n), img_f(n);
    for (size_t i = 0; i < n; i++) img_x_msk[i] = msk_in[i] ? img_in[i] : 0.0f;
    float fx[9], fy[9];
    pyramid_filters(factor_x, factor_y, fx, fy);
    separate_convolution(msk_in, msk_f.data(), W, H, fx, fy, 9, 9);
    separate_convolution(img_x_msk.data(), img_f.data(), W, H, fx, fy, 9, 9);
    float min_x = (float)((double)factor_x / 2.0 - 0.5);
    float min_y = (float)((double)factor_y / 2.0 - 0.5);
    for (int i = 0; i < sH; i++) {
        float y = min_y + (float)(unsigned)i * factor_y;
        for (int j = 0; j < sW; j++) {
            float x = min_x + (float)(unsigned)j * factor_x;
            float mv = bilinear(msk_f.data(), W, H, x, y);
            float iv = bilinear(img_f.data(), W, H, x, y);
            img_out[(size_t)i * sW + j] = ((double)mv > 1e-6) ? iv / mv : 0.f;
            msk_out[(size_t)i * sW + j] = mv;
        }
    }
}

/* sampling.cpp:248-302 downsample_internal, one channel */
void downsample_internal(const float *in, float *out, int W, int H, int sW, int sH)
{
    float factor_x = (float)W / (float)sW;
    float factor_y = (float)H / (float)sH;
    float fx[9], fy[9];
    pyramid_filters(factor_x, factor_y, fx, fy);
    fvec buf((size_t)W * H);
    separate_convolution(in, buf.data(), W, H, fx, fy, 9, 9);
    float min_x = (float)((double)factor_x / 2.0 - 0.5);
    float min_y = (float)((double)factor_y / 2.0 - 0.5);
    for (int i = 0; i < sH; i++) {
        float y = min_y + (float)(unsigned)i * factor_y;
        for (int j = 0; j < sW; j++) {
            float x = min_x + (float)(unsigned)j * factor_x;
            out[(size_t)i * sW + j] = bilinear(buf.data(), W, H, x, y);
        }
    }
}

inline float clamp_threshold(float t)
{
    return (float)std::fmax(0.0, std::fmin(1.0, (double)t));
}

/* distance_transform.cpp:106-115 */
inline float dt_f(int x, int xi, float gi)
{
    double dx = (double)(float)(x - xi);
    double g = (double)gi;
    return (float)std::sqrt(dx * dx + g * g);
}
inline int dt_sep(int i, int u, float gi, float gu)
{
    double du = (double)(float)u, di = (double)(float)i;
    double a = du * du - di * di + (double)gu * (double)gu - (double)gi * (double)gi;
    return (int)(a / (double)(2 * (u - i)));
}

/* ---------------- patch distances ---------------- */

struct DistCtx {
    nlo_dist_params p;
    fvec weights;
    const float *src, *tgt, *sg, *tg;
    int sW, tW, C;
};

float dist_eval(const DistCtx &d, int sx, int sy, int tx, int ty)
{
    return nlo_patch_distance(&d.p, d.weights.data(), d.src, d.sg, d.sW, d.tgt, d.tg, d.tW, d.C,
                              sx, sy, tx, ty);
}

inline bool needs_gradient(int kind) { return kind == NLO_DIST_L2COMBINED || kind == NLO_DIST_L2RGBD; }

/* ---------------- updaters ---------------- */

/* patch_non_local_medians.cpp:110-126 */
struct CompareByColor {
    bool operator()(const std::pair<float, float> &a, const std::pair<float, float> &b) const
    {
        return a.first < b.first;
    }
};

float weighted_median(std::vector<std::pair<float, float> > &m, float total_weight)
{
    std::sort(m.begin(), m.end(), CompareByColor());
    float median = 0.0f;
    float half = total_weight / 2;
    for (size_t i = 0; i < m.size(); i++) {
        total_weight -= m[i].second;
        if (total_weight < half) {
            median = m[i].first;
            break;
        }
    }
    return median;
}

/* patch_non_local_means.cpp:23-100 (in place, scan order) */
double nlmeans(int P, const float *weights, float *image, const uint8_t *domain,
               const int32_t *nnf, const float *conf, int W, int H, int C)
{
    int hx = P / 2, hy = P / 2;
    double total = 0.0;
    std::vector<float> tv(C);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            if (!domain[(size_t)W * y + x]) continue;
            int x_a = std::max(x - hx, 0), x_b = std::min(x + hx, W - 1);
            int y_a = std::max(y - hy, 0), y_b = std::min(y + hy, H - 1);
            for (int c = 0; c < C; c++) tv[c] = 0.0f;
            float tw = 0.0f;
            for (int i = x_a; i <= x_b; i++)
                for (int j = y_a; j <= y_b; j++) {
                    int cx = nnf[2 * ((size_t)W * j + i)], cy = nnf[2 * ((size_t)W * j + i) + 1];
                    if (cx == -1) continue;
                    cx += x - i;
                    cy += y - j;
                    if (cx < 0 || cy < 0 || cx >= W || cy >= H) continue;
                    float w = weights[(size_t)P * (j - y_a) + (i - x_a)];
                    if (conf && domain[(size_t)W * j + i]) w *= conf[(size_t)W * j + i];
                    tw += w;
                    for (int c = 0; c < C; c++) tv[c] += image[(size_t)C * ((size_t)W * cy + cx) + c] * w;
                }
            if (tw > 0.0) {
                for (int c = 0; c < C; c++) {
                    float v = tv[c] / tw;
                    float &px = image[(size_t)C * ((size_t)W * y + x) + c];
                    float prev = px;
                    float d = prev - v;
                    total += d * d;
                    px = v;
                }
            }
        }
    return total;
}

/* patch_non_local_medians.cpp:23-106 (in place, scan order) */
double nlmedians(int P, const float *weights, float *image, const uint8_t *domain,
                 const int32_t *nnf, const float *conf, int W, int H, int C)
{
    int hx = P / 2, hy = P / 2;
    double total = 0.0;
    std::vector<std::vector<std::pair<float, float> > > maps(C);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            if (!domain[(size_t)W * y + x]) continue;
            int x_a = std::max(x - hx, 0), x_b = std::min(x + hx, W - 1);
            int y_a = std::max(y - hy, 0), y_b = std::min(y + hy, H - 1);
            float tw = 0.0f;
            for (int c = 0; c < C; c++) maps[c].clear();
            for (int i = x_a; i <= x_b; i++)
                for (int j = y_a; j <= y_b; j++) {
                    int cx = nnf[2 * ((size_t)W * j + i)], cy = nnf[2 * ((size_t)W * j + i) + 1];
                    if (cx == -1) continue;
                    cx += x - i;
                    cy += y - j;
                    if (cx < 0 || cy < 0 || cx >= W || cy >= H) continue;
                    float w = weights[(size_t)P * (j - y_a) + (i - x_a)];
                    if (conf && domain[(size_t)W * j + i]) w *= conf[(size_t)W * j + i];
                    tw += w;
                    for (int c = 0; c < C; c++)
                        maps[c].push_back(std::make_pair(image[(size_t)C * ((size_t)W * cy + cx) + c], w));
                }
            if (maps[0].empty()) continue;
            for (int c = 0; c < C; c++) {
                float v = weighted_median(maps[c], tw);
                float &px = image[(size_t)C * ((size_t)W * y + x) + c];
                float prev = px;
                float d = prev - v;
                total += d * d;
                px = v;
            }
        }
    return total;
}

/* patch_non_local_poisson.cpp:36-121 */
double nlpoisson(int P, float sigma, float lambda, float cg_tol, int cg_limit, float *image,
                 const float *orig, const uint8_t *domain, const uint8_t *ext,
                 const int32_t *nnf, const float *conf, int W, int H, int C,
                 nlo_cg_stats *stats, int stats_offset)
{
    size_t n = (size_t)W * H;
    std::vector<double> a1(n), a2(n), f1(n * C), f2(n * C), pert(n * C, 0.0), u0(n * C);
    nlo_poisson_coefficients(P, sigma, lambda, image, ext, nnf, conf, W, H, C,
                             a1.data(), a2.data(), f1.data(), f2.data());
    /* split_image_into_channels(original) :358-381 */
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < C; c++) u0[(size_t)c * n + i] = orig[i * C + c];
    std::vector<Pt> pts = masked_points(domain, W, H);
    std::vector<double> channel(n);
    for (int c = 0; c < C; c++) {
        double *u0c = &u0[(size_t)c * n];
        double *f1c = &f1[(size_t)c * n];
        double *f2c = &f2[(size_t)c * n];
        double *pc = &pert[(size_t)c * n];
        std::memcpy(channel.data(), u0c, n * sizeof(double));
        nlo_anisotropic_laplacian(channel.data(), a2.data(), domain, W, H);
        for (size_t k = 0; k < pts.size(); k++) {
            size_t idx = (size_t)W * pts[k].y + pts[k].x;
            f1c[idx] += -f2c[idx] - a1[idx] * u0c[idx] + channel[idx];
            pc[idx] = (double)image[idx * C + c] - u0c[idx];
        }
        double nr0 = 0, nr = 0;
        int it = nlo_conjugate_gradient(pc, a1.data(), a2.data(), f1c, domain, W, H, cg_tol, cg_limit,
                                        &nr0, &nr);
        if (stats && stats_offset + c < 4) {
            stats->cg_iterations[stats_offset + c] = it;
            stats->cg_nr0[stats_offset + c] = nr0;
            stats->cg_nr[stats_offset + c] = nr;
            stats->channels_solved = stats_offset + c + 1;
        }
    }
    double total = 0.0;
    for (size_t k = 0; k < pts.size(); k++) {
        size_t idx = (size_t)W * pts[k].y + pts[k].x;
        for (int c = 0; c < C; c++) {
            float v = (float)(u0[(size_t)c * n + idx] + pert[(size_t)c * n + idx]);
            float prev = image[idx * C + c];
            float d = prev - v;
            total += d * d;
            image[idx * C + c] = v;
        }
    }
    return total;
}

/* ---------------- simpois (3rdparty/simpois/simpois.c) ---------------- */

inline float sp_get(const float *I, int w, int h, int i, int j)   /* :23-31 getpixel_1 */
{
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    if (i >= w) i = w - 1;
    if (j >= h) j = h - 1;
    return I[i + (size_t)j * w];
}

inline float sp_lap_neum(const float *I, int w, int h, int i, int j)   /* :61-79 */
{
    float x = sp_get(I, w, h, i, j);
    float a = sp_get(I, w, h, i + 1, j);
    float b = sp_get(I, w, h, i, j + 1);
    float c = sp_get(I, w, h, i - 1, j);
    float d = sp_get(I, w, h, i, j - 1);
    float r = 0;
    if (std::isfinite(a)) r += a - x;
    if (std::isfinite(b)) r += b - x;
    if (std::isfinite(c)) r += c - x;
    if (std::isfinite(d)) r += d - x;
    return r;
}

void sp_zoom_out(float *out, int ow, int oh, const float *in, int iw, int ih)   /* :146-169 */
{
    for (int j = 0; j < oh; j++)
        for (int i = 0; i < ow; i++) {
            float a[4], m = 0;
            a[0] = sp_get(in, iw, ih, 2 * i, 2 * j);
            a[1] = sp_get(in, iw, ih, 2 * i + 1, 2 * j);
            a[2] = sp_get(in, iw, ih, 2 * i, 2 * j + 1);
            a[3] = sp_get(in, iw, ih, 2 * i + 1, 2 * j + 1);
            int cx = 0;
            for (int k = 0; k < 4; k++)
                if (std::isfinite(a[k])) { m += a[k]; cx += 1; }
            out[(size_t)ow * j + i] = cx ? m / cx : NAN;
        }
}

inline float sp_bilinear(const float *x, int w, int h, float p, float q)   /* :172-196 */
{
    int ip = (int)std::floor((double)p);
    int iq = (int)std::floor((double)q);
    float a = sp_get(x, w, h, ip, iq);
    float b = sp_get(x, w, h, ip + 1, iq);
    float c = sp_get(x, w, h, ip, iq + 1);
    float d = sp_get(x, w, h, ip + 1, iq + 1);
    float fx = p - (float)ip, fy = q - (float)iq;
    float r = 0;
    r += a * (1 - fx) * (1 - fy);
    r += b * (fx) * (1 - fy);
    r += c * (1 - fx) * (fy);
    r += d * (fx) * (fy);
    return r;
}

void sp_zoom_in(float *out, int ow, int oh, const float *in, int iw, int ih)   /* :200-213 */
{
    for (int j = 0; j < oh; j++)
        for (int i = 0; i < ow; i++) {
            float x = (float)((i - 0.5) / 2);
            float y = (float)((j - 0.5) / 2);
            out[(size_t)ow * j + i] = sp_bilinear(in, iw, ih, x, y);
        }
}

/* :228-263 poisson_rec with f = NULL, cgit = 0 (the simplest entry point, :281-291) */
void sp_rec(float *u, const float *g, int w, int h, float tstep, int niter, int scale)
{
    fvec init((size_t)w * h);
    if (scale > 1 && (w > 1 || h > 1)) {
        int ws = (int)std::ceil(w / 2.0);
        int hs = (int)std::ceil(h / 2.0);
        fvec gs((size_t)ws * hs), us((size_t)ws * hs);
        sp_zoom_out(gs.data(), ws, hs, g, w, h);
        sp_rec(us.data(), gs.data(), ws, hs, tstep, niter, scale - 1);
        sp_zoom_in(init.data(), w, h, us.data(), ws, hs);
    } else {
        for (int i = 0; i < w * h; i++) init[i] = 0;
    }
    /* :113-143 poisson_extension_with_init; omega = NaN pixels in row-major order */
    std::vector<int> omega;
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++)
            if (std::isnan(g[(size_t)j * w + i])) omega.push_back(j * w + i);
    for (int i = 0; i < w * h; i++) u[i] = std::isfinite(g[i]) ? g[i] : init[i];
    for (int it = 0; it < niter; it++)
        for (size_t p = 0; p < omega.size(); p++) {   /* :82-100 gauss_seidel_iteration */
            int ij = omega[p];
            int i = ij % w, j = ij / w;
            float l = sp_lap_neum(u, w, h, i, j);
            float d = 0;
            u[ij] = u[ij] + tstep * (l - d);
        }
}

}  // namespace

/* ======================================================================== */
extern "C" {

/* gaussian_weights.cpp:19-80 (SEGMENTS = 1, gaussian_weights.h:23) */
void nlo_gaussian_weights(int size_x, int size_y, float sigma_x, float sigma_y, float *values)
{
    unsigned sx = (unsigned)size_x, sy = (unsigned)size_y;
    for (unsigned i = 0; i < sx * sy; i++) values[i] = 0.0f;
    double x_a = -((float)sx / 2);
    double y_a = -((float)sy / 2);
    const double half_step = 0.5;
    double ssx = (double)sigma_x * (double)sigma_x;
    double ssy = (double)sigma_y * (double)sigma_y;
    if (sy == 1) {
        for (unsigned i = 0; i <= sx / 2; i++) {
            double x = x_a + i + 0.0 + half_step;
            double value = 0.0;
            value += std::exp(-(x * x / ssx));
            values[i] = (float)value;
        }
    } else {
        for (unsigned i = 0; i <= sx / 2; i++)
            for (unsigned j = 0; j <= sy / 2; j++) {
                double x = x_a + i + 0.0 + half_step;
                double y = y_a + j + 0.0 + half_step;
                double value = 0.0;
                value += std::exp(-(x * x / (2 * ssx) + y * y / (2 * ssy)));
                values[sx * j + i] = (float)value;
            }
    }
    for (unsigned i = 0; i <= sx / 2; i++)
        for (unsigned j = 0; j <= sy / 2; j++) {
            float v = values[sx * j + i];
            values[sx * j + (sx - 1 - i)] = v;
            if (sy > 1) {
                values[sx * (sy - 1 - j) + i] = v;
                values[sx * (sy - 1 - j) + (sx - 1 - i)] = v;
            }
        }
    float sum = 0.0f;
    for (unsigned i = 0; i < sx * sy; i++) sum += values[i];
    for (unsigned i = 0; i < sx * sy; i++) values[i] /= sum;
}

/* ---- colour conversion (io_utility.cpp:268-318, per-pixel steps :380-450) ----
 * Each step keeps the reference's types: float arithmetic, except where a double
 * constant promotes an expression (the 0.04045 test and the /12.92 branch of the
 * gamma expansion, the a/500.0 term of the inverse), and powf for pow(float, float). */
static void lab_of_rgb(const float *in, float *out)
{
    float lin[3];
    for (int c = 0; c < 3; c++) {
        const float v = in[c] / 255.0f;                                 /* :382-384 */
        const float e = ((double)v > 0.04045) ? powf((v + 0.055f) / 1.055f, 2.4f)
                                               : (float)((double)v / 12.92);   /* :386-388 */
        lin[c] = e * 100.0f;                                            /* :390-392 */
    }
    /* :394-396, sRGB primaries to XYZ (D65); sums left to right */
    const float X = lin[0] * 0.412453f + lin[1] * 0.357580f + lin[2] * 0.180423f;
    const float Y = lin[0] * 0.212671f + lin[1] * 0.715160f + lin[2] * 0.072169f;
    const float Z = lin[0] * 0.019334f + lin[1] * 0.119193f + lin[2] * 0.950227f;
    /* :403-409, white point normalisation and the cube-root / linear toe */
    float f[3] = {X / 95.047f, Y / 100.000f, Z / 108.883f};
    for (int c = 0; c < 3; c++)
        f[c] = (f[c] > 0.008856f) ? powf(f[c], 1.0f / 3.0f) : (7.787f * f[c]) + (16.0f / 116.0f);
    out[0] = (116.0f * f[1]) - 16.0f;                                   /* :411-413 */
    out[1] = 500.0f * (f[0] - f[1]);
    out[2] = 200.0f * (f[1] - f[2]);
}

static void rgb_of_lab(const float *in, float *out)
{
    float f[3];
    f[1] = (in[0] + 16.0f) / 116.0f;                                    /* :419-421 */
    f[0] = (float)((double)in[1] / 500.0 + (double)f[1]);
    f[2] = f[1] - in[2] / 200.0f;
    const float white[3] = {95.047f, 100.000f, 108.883f};
    float xyz[3];
    for (int c = 0; c < 3; c++) {                                       /* :423-429 */
        const float cube = powf(f[c], 3.0f);
        const float t = (cube > 0.008856f) ? cube : (f[c] - 16.0f / 116.0f) / 7.787f;
        xyz[c] = t * white[c];
    }
    const float x = xyz[0] / 100.0f, y = xyz[1] / 100.0f, z = xyz[2] / 100.0f;   /* :435-437 */
    float lin[3];                                                       /* :439-441 */
    lin[0] = x * 3.240479f + y * -1.537150f + z * -0.498535f;
    lin[1] = x * -0.969256f + y * 1.875992f + z * 0.041556f;
    lin[2] = x * 0.055648f + y * -0.204043f + z * 1.057311f;
    for (int c = 0; c < 3; c++) {                                       /* :443-449 */
        const float g = (lin[c] > 0.0031308f) ? 1.055f * powf(lin[c], 1.0f / 2.4f) - 0.055f : 12.92f * lin[c];
        out[c] = g * 255.0f;
    }
}

void nlo_rgb_to_lab(const float *rgb, int W, int H, int C, float *lab)
{
    const size_t n = (size_t)W * H;
    if (C != 3) {                                                       /* :270-272 */
        std::memcpy(lab, rgb, sizeof(float) * n * C);
        return;
    }
    for (size_t i = 0; i < n; i++) lab_of_rgb(rgb + 3 * i, lab + 3 * i);
}

void nlo_lab_to_rgb(const float *lab, int W, int H, int C, float *rgb)
{
    const size_t n = (size_t)W * H;
    if (C != 3 && C != 4) {                                             /* :294-296 */
        std::memcpy(rgb, lab, sizeof(float) * n * C);
        return;
    }
    for (size_t i = 0; i < n; i++) rgb_of_lab(lab + (size_t)C * i, rgb + 3 * i);
}

/* gradient.cpp:20-76, forward differences in x and y, zero on the last column/row */
void nlo_gradient(const float *img, int W, int H, int C, float *grad)
{
    for (int x = 0; x < W; x++)
        for (int y = 0; y < H; y++) {
            size_t index = (size_t)C * ((size_t)W * y + x);
            for (int ch = 0; ch < C; ch++) {
                grad[(index + ch) * 2] = (x == W - 1) ? 0.0f : img[index + C + ch] - img[index + ch];
                grad[(index + ch) * 2 + 1] =
                    (y == H - 1) ? 0.0f : img[index + (size_t)C * W + ch] - img[index + ch];
            }
        }
}

/* distance_transform.cpp:21-98 (Meijster) */
void nlo_distance_transform(const uint8_t *mask, int W, int H, float *out)
{
    fvec g((size_t)W * H);
    const float INF = (float)(unsigned)(W + H);
#define G(x, y) g[(size_t)(y) * W + (x)]
    for (int x = 0; x < W; x++) {
        G(x, 0) = mask[x] ? INF : 0.0f;
        for (int y = 1; y < H; y++)
            G(x, y) = mask[(size_t)W * y + x] ? (float)((double)G(x, y - 1) + 1.0) : 0.0f;
        for (int y = H - 2; y >= 0; y--)
            if (G(x, y + 1) < G(x, y)) G(x, y) = (float)((double)G(x, y + 1) + 1.0);
    }
    std::vector<int> s(W + 1), t(W + 1);
    for (int y = 0; y < H; y++) {
        int q = 0;
        s[0] = 0;
        t[0] = 0;
        for (int u = 1; u < W; u++) {
            while (q >= 0 && dt_f(t[q], s[q], G(s[q], y)) > dt_f(t[q], u, G(u, y))) q--;
            if (q < 0) {
                q = 0;
                s[0] = u;
            } else {
                unsigned w = (unsigned)(1 + dt_sep(s[q], u, G(s[q], y), G(u, y)));
                if (w < (unsigned)W) {
                    q++;
                    s[q] = u;
                    t[q] = (int)w;
                }
            }
        }
        for (int u = W - 1; u >= 0; u--) {
            out[(size_t)W * y + u] = dt_f(u, s[q], G(s[q], y));
            if (u == t[q]) q--;
        }
    }
#undef G
}

/* image_inpainting.cpp:611-630 -- uses the float overload of exp (expf) */
void nlo_confidence(const uint8_t *mask, int W, int H, float decay, float asym, float *out)
{
    size_t n = (size_t)W * H;
    for (size_t i = 0; i < n; i++) out[i] = 1.0f;
    if (decay > 0) {
        fvec d(n);
        nlo_distance_transform(mask, W, H, d.data());
        for (size_t i = 0; i < n; i++)
            if (mask[i]) out[i] = (1 - asym) * expf(-d[i] / decay) + asym;
    } else {
        for (size_t i = 0; i < n; i++)
            if (mask[i]) out[i] = asym;
    }
}

/* image_inpainting.cpp:197-204 -- float overloads of log/exp */
float nlo_subsampling_rate(float ratio, int scales)
{
    if (ratio < 0.0 || ratio > 1.0) return 0.0f;
    return expf(logf(ratio) / (float)(unsigned)(scales - 1));
}

/* sampling.cpp:408-411 */
int nlo_sample_size(int size, float factor)
{
    return (int)((float)(unsigned)size * factor + 0.5);
}

/* sampling.cpp:114-202 */
int nlo_downsample_with_mask(const float *img, const uint8_t *mask, int W, int H, int C,
                             float factor, float threshold, int mask_val,
                             float *out_img, uint8_t *out_mask)
{
    if (factor <= 0 || factor >= 1.0) return -1;
    size_t n = (size_t)W * H;
    fvec fm(n);
    for (size_t i = 0; i < n; i++) fm[i] = ((mask_val != 0) == (mask[i] != 0)) ? 1.0f : 0.0f;
    int sW = nlo_sample_size(W, factor), sH = nlo_sample_size(H, factor);
    size_t sn = (size_t)sW * sH;
    fvec sfm(sn), cin(n), cout(sn);
    for (int c = 0; c < C; c++) {
        for (size_t i = 0; i < n; i++) cin[i] = img[i * C + c];
        downsample_internal_with_mask(cin.data(), fm.data(), cout.data(), sfm.data(), W, H, sW, sH);
        for (size_t i = 0; i < sn; i++) out_img[i * C + c] = cout[i];
    }
    float t = clamp_threshold(threshold);
    for (size_t i = 0; i < sn; i++) out_mask[i] = mask_val ? (sfm[i] > t) : (sfm[i] < t);
    return 0;
}

/* sampling.cpp:19-70 */
int nlo_downsample(const float *img, int W, int H, int C, float factor, float *out)
{
    if (factor <= 0 || factor >= 1.0) return -1;
    int sW = nlo_sample_size(W, factor), sH = nlo_sample_size(H, factor);
    size_t n = (size_t)W * H, sn = (size_t)sW * sH;
    fvec cin(n), cout(sn);
    for (int c = 0; c < C; c++) {
        for (size_t i = 0; i < n; i++) cin[i] = img[i * C + c];
        downsample_internal(cin.data(), cout.data(), W, H, sW, sH);
        for (size_t i = 0; i < sn; i++) out[i * C + c] = cout[i];
    }
    return 0;
}

/* sampling.cpp:73-111 */
int nlo_downsample_mask(const uint8_t *mask, int W, int H, float factor, float threshold, uint8_t *out)
{
    if (factor <= 0 || factor >= 1.0) return -1;
    size_t n = (size_t)W * H;
    fvec fm(n, 0.0f);
    for (size_t i = 0; i < n; i++) if (mask[i]) fm[i] = 1.0f;
    int sW = nlo_sample_size(W, factor), sH = nlo_sample_size(H, factor);
    fvec sfm((size_t)sW * sH);
    downsample_internal(fm.data(), sfm.data(), W, H, sW, sH);
    float t = clamp_threshold(threshold);
    for (size_t i = 0; i < (size_t)sW * sH; i++) out[i] = sfm[i] > t;
    return 0;
}

/* sampling.cpp:205-246, 387-405 -- nearest-neighbour upsampling */
void nlo_upsample(const float *in, int W, int H, int C, int W2, int H2, float *out)
{
    float factor_x = (float)(unsigned)W2 / (float)(unsigned)W;
    float factor_y = (float)(unsigned)H2 / (float)(unsigned)H;
    float min_x = (float)((double)factor_x / 2.0 - 0.5);
    float min_y = (float)((double)factor_y / 2.0 - 0.5);
    for (int i = 0; i < H2; i++) {
        float y = ((float)(unsigned)i - min_y) / factor_y;
        for (int j = 0; j < W2; j++) {
            float x = ((float)(unsigned)j - min_x) / factor_x;
            int ix = neumann((int)std::floor((double)x + 0.5), W);
            int iy = neumann((int)std::floor((double)y + 0.5), H);
            for (int c = 0; c < C; c++)
                out[((size_t)i * W2 + j) * C + c] = in[((size_t)iy * W + ix) * C + c];
        }
    }
}

/* image_inpainting.cpp:566-600 */
void nlo_extended_domain(const uint8_t *hole, int W, int H, int P, uint8_t *out)
{
    int rx = P / 2, ry = P / 2;
    std::memcpy(out, hole, (size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            if (!hole[(size_t)W * y + x]) continue;
            bool surrounded = mtest(hole, W, H, x + 1, y) && mtest(hole, W, H, x - 1, y) &&
                              mtest(hole, W, H, x, y + 1) && mtest(hole, W, H, x, y - 1);
            if (surrounded) continue;
            int x_a = std::max(x - rx, 0), x_b = std::min(x + rx, W - 1);
            int y_a = std::max(y - ry, 0), y_b = std::min(y + ry, H - 1);
            for (int i = x_a; i <= x_b; i++)
                for (int j = y_a; j <= y_b; j++)
                    if (!hole[(size_t)W * j + i]) out[(size_t)W * j + i] = 1;
        }
}

/* image_inpainting.cpp:641-657 */
void nlo_add_margin(uint8_t *a, uint8_t *b, int W, int H, int margin)
{
    for (int i = 0; i < margin; i++) {
        for (int x = 0; x < W; x++) {
            int rows[2] = {i, H - i - 1};
            for (int k = 0; k < 2; k++)
                if ((unsigned)rows[k] < (unsigned)H) {
                    a[(size_t)W * rows[k] + x] = 0;
                    b[(size_t)W * rows[k] + x] = 0;
                }
        }
        for (int y = 0; y < H; y++) {
            int cols[2] = {i, W - i - 1};
            for (int k = 0; k < 2; k++)
                if ((unsigned)cols[k] < (unsigned)W) {
                    a[(size_t)W * y + cols[k]] = 0;
                    b[(size_t)W * y + cols[k]] = 0;
                }
        }
    }
}

/* image_inpainting.cpp:359-384 (erode = 1: inpaint_internal) and
 * :496-502 / :543-548 (erode = 0: propagate_weights).  Returns 1 when the eroded
 * source is empty before the margin (the reference throws, :382-384). */
int nlo_scale_masks(const uint8_t *hole, int W, int H, int P, int erode,
                    uint8_t *ext, uint8_t *target, uint8_t *source)
{
    size_t n = (size_t)W * H;
    nlo_extended_domain(hole, W, H, P, ext);
    std::memcpy(target, ext, n);
    for (size_t i = 0; i < n; i++) source[i] = !ext[i];
    int empty = 1;
    if (erode) {
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++)
                if (mtest(source, W, H, i, j)) {
                    if (!mtest(source, W, H, i + 1, j)) source[(size_t)W * j + i] = 0;
                    else if (!mtest(source, W, H, i, j + 1)) source[(size_t)W * j + i] = 0;
                    else empty = 0;
                }
    } else {
        empty = 0;
    }
    nlo_add_margin(source, target, W, H, P / 2);
    return empty;
}

/* image_inpainting.cpp:511-530, 660-663 -- NNF scaling with round-half-away */
void nlo_scale_nnf(const int32_t *lnnf, const uint8_t *lhole, int lW, int lH, int uW, int uH,
                   int32_t *unnf)
{
    float scale_x = (float)(unsigned)uW / (float)(unsigned)lW;
    float scale_y = (float)(unsigned)uH / (float)(unsigned)lH;
    auto rnd = [](float v) -> int { return (v > 0.0) ? (int)std::floor((double)v + 0.5) : (int)std::ceil((double)v - 0.5); };
    for (size_t i = 0; i < (size_t)uW * uH * 2; i++) unnf[i] = -1;
    for (int y = 0; y < lH; y++)
        for (int x = 0; x < lW; x++) {
            if (!lhole[(size_t)lW * y + x]) continue;
            int nx = lnnf[2 * ((size_t)lW * y + x)], ny = lnnf[2 * ((size_t)lW * y + x) + 1];
            if (nx < 0) continue;
            nx = rnd((float)nx * scale_x);
            ny = rnd((float)ny * scale_y);
            int ux = rnd((float)x * scale_x);
            int uy = rnd((float)y * scale_y);
            if (ux < 0 || uy < 0 || ux >= uW || uy >= uH) continue;   /* never taken: rounding stays in range */
            unnf[2 * ((size_t)uW * uy + ux)] = nx;
            unnf[2 * ((size_t)uW * uy + ux) + 1] = ny;
        }
}

/* l2_norm_patch_distance.cpp:22-104, l1_norm_patch_distance.cpp:22-105,
 * l2_combined_patch_distance.cpp:39-148, l2_rgbd_patch_distance.cpp:41-117.
 * Returns NaN for RGBD with C != 4 (the reference throws, :111-112). */
float nlo_patch_distance(const nlo_dist_params *d, const float *w,
                         const float *src, const float *sg, int sW,
                         const float *tgt, const float *tg, int tW, int C,
                         int sx, int sy, int tx, int ty)
{
    const int P = d->patch;
    const int r = P / 2;
    const float lambda = d->lambda;
    const float *pw = w;
    if (d->kind == NLO_DIST_L2 || d->kind == NLO_DIST_L1) {
        const bool l1 = d->kind == NLO_DIST_L1;
        float distance = 0.0f;
        for (int dy = -r; dy < P - r; dy++)
            for (int dx = -r; dx < P - r; dx++, ++pw) {
                float weight = *pw;
                int s_id = C * (sW * (sy + dy) + (sx + dx));
                int t_id = C * (tW * (ty + dy) + (tx + dx));
                float norm;
                if (C == 3) {
                    float a = src[s_id] - tgt[t_id], b = src[s_id + 1] - tgt[t_id + 1],
                          c = src[s_id + 2] - tgt[t_id + 2];
                    norm = l1 ? std::fabs(a) + std::fabs(b) + std::fabs(c) : a * a + b * b + c * c;
                } else if (C == 1) {
                    float a = src[s_id] - tgt[t_id];
                    norm = l1 ? std::fabs(a) : a * a;
                } else {
                    norm = 0.0f;
                    for (int ch = 0; ch < C; ch++) {
                        float a = src[s_id + ch] - tgt[t_id + ch];
                        norm += l1 ? std::fabs(a) : a * a;
                    }
                }
                distance += weight * norm;
            }
        return distance / (float)(unsigned)(P * P);
    }
    if (d->kind == NLO_DIST_L2COMBINED) {
        double distance = 0.0, gdistance = 0.0;
        for (int dy = -r; dy < P - r; dy++)
            for (int dx = -r; dx < P - r; dx++, ++pw) {
                float weight = *pw;
                int s_id = C * (sW * (sy + dy) + (sx + dx));
                int t_id = C * (tW * (ty + dy) + (tx + dx));
                if (C == 3 || C == 1) {
                    float norm = 0.0f, gnorm = 0.0f;
                    for (int ch = 0; ch < C; ch++) {
                        float a = src[s_id + ch] - tgt[t_id + ch];
                        norm = (ch == 0) ? a * a : norm + a * a;
                    }
                    for (int k = 0; k < 2 * C; k++) {
                        float a = sg[s_id * 2 + k] - tg[t_id * 2 + k];
                        gnorm = (k == 0) ? a * a : gnorm + a * a;
                    }
                    distance += weight * norm;
                    gdistance += weight * gnorm;
                } else {
                    double norm = 0.0, gnorm = 0.0;
                    for (int ch = 0; ch < C; ch++) {
                        float a = src[s_id + ch] - tgt[t_id + ch];
                        norm += a * a;
                        float gx = sg[(s_id + ch) * 2] - tg[(t_id + ch) * 2];
                        float gy = sg[(s_id + ch) * 2 + 1] - tg[(t_id + ch) * 2 + 1];
                        gnorm += gx * gx + gy * gy;
                    }
                    distance += weight * norm;
                    gdistance += weight * gnorm;
                }
            }
        return (float)((lambda * distance + (1 - lambda) * gdistance) / (unsigned)(P * P));
    }
    if (d->kind == NLO_DIST_L2RGBD) {
        if (C != 4) return NAN;
        const float lambda_d = d->lambda_d;
        double dist = 0.0, dist_d = 0.0, gdist = 0.0, gdist_d = 0.0;
        for (int dy = -r; dy < P - r; dy++)
            for (int dx = -r; dx < P - r; dx++, ++pw) {
                float weight = *pw;
                double norm = 0.0, norm_d = 0.0, gnorm = 0.0, gnorm_d = 0.0;
                int s_id = C * (sW * (sy + dy) + (sx + dx));
                int t_id = C * (tW * (ty + dy) + (tx + dx));
                for (int ch = 0; ch < C; ch++) {
                    float a = src[s_id + ch] - tgt[t_id + ch];
                    float gx = sg[(s_id + ch) * 2] - tg[(t_id + ch) * 2];
                    float gy = sg[(s_id + ch) * 2 + 1] - tg[(t_id + ch) * 2 + 1];
                    if (ch == 3) {
                        norm_d += a * a;
                        gnorm_d += gx * gx + gy * gy;
                    } else {
                        norm += a * a;
                        gnorm += gx * gx + gy * gy;
                    }
                }
                dist += weight * norm;
                dist_d += weight * norm_d;
                gdist += weight * gnorm;
                gdist_d += weight * gnorm_d;
            }
        return (float)((lambda * dist + (1 - lambda) * gdist + lambda_d * dist_d +
                        (1 - lambda_d) * gdist_d) / (unsigned)(P * P));
    }
    return NAN;
}

/* Distance of every target point to its NNF entry (FLT_MAX where the entry is -1). */
int nlo_patch_distance_field(const nlo_dist_params *d, const float *src, int sW, int sH,
                             const float *tgt, int tW, int tH, int C,
                             const int32_t *nnf, const uint8_t *tmask, float *out)
{
    fvec w((size_t)d->patch * d->patch), sg, tg;
    nlo_gaussian_weights(d->patch, d->patch, d->sigma, d->sigma, w.data());
    if (needs_gradient(d->kind)) {
        sg.resize((size_t)sW * sH * C * 2);
        tg.resize((size_t)tW * tH * C * 2);
        nlo_gradient(src, sW, sH, C, sg.data());
        nlo_gradient(tgt, tW, tH, C, tg.data());
    }
    for (int y = 0; y < tH; y++)
        for (int x = 0; x < tW; x++) {
            size_t i = (size_t)tW * y + x;
            out[i] = FLT_MAX;
            if (!tmask[i] || nnf[2 * i] < 0) continue;
            out[i] = nlo_patch_distance(d, w.data(), src, sg.empty() ? nullptr : sg.data(), sW, tgt,
                                        tg.empty() ? nullptr : tg.data(), tW, C, nnf[2 * i],
                                        nnf[2 * i + 1], x, y);
        }
    (void)sH;
    return 0;
}

/* patch_match.cpp:311-512 (serial build): glibc rand() seeded with srand(seed).
 * energy_per_iter (optional, length pm->iterations) is the METRICS total distance
 * (patch_match.cpp:501-508). Returns -1 on a size mismatch (the reference returns an
 * empty field, :317-322). */
static int pm_serial(const nlo_dist_params *dp, const nlo_pm_params *pm,
                     const float *src, const uint8_t *smask, int sW, int sH,
                     const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                     const int32_t *init_nnf, bool reseed, unsigned seed,
                     int32_t *out_nnf, float *out_dist, double *energy_per_iter)
{
    if (!dp || !pm) return -1;
    if (reseed) srand(seed);
    DistCtx d;
    d.p = *dp;
    d.weights.resize((size_t)dp->patch * dp->patch);
    nlo_gaussian_weights(dp->patch, dp->patch, dp->sigma, dp->sigma, d.weights.data());
    fvec sg, tg;
    if (needs_gradient(dp->kind)) {
        sg.resize((size_t)sW * sH * C * 2);
        tg.resize((size_t)tW * tH * C * 2);
        nlo_gradient(src, sW, sH, C, sg.data());
        nlo_gradient(tgt, tW, tH, C, tg.data());
    }
    d.src = src; d.tgt = tgt; d.sg = sg.empty() ? nullptr : sg.data(); d.tg = tg.empty() ? nullptr : tg.data();
    d.sW = sW; d.tW = tW; d.C = C;

    size_t tn = (size_t)tW * tH;
    fvec dist(tn, std::numeric_limits<float>::max());
    ivec nb(2 * tn, -1);
    std::vector<Pt> tp = masked_points(tmask, tW, tH);
    auto NX = [&](int x, int y) -> int32_t & { return nb[2 * ((size_t)tW * y + x)]; };
    auto NY = [&](int x, int y) -> int32_t & { return nb[2 * ((size_t)tW * y + x) + 1]; };
    auto DI = [&](int x, int y) -> float & { return dist[(size_t)tW * y + x]; };

    if (init_nnf) {
        for (size_t i = 0; i < tp.size(); i++) {
            Pt p = tp[i];
            int nx = init_nnf[2 * ((size_t)tW * p.y + p.x)], ny = init_nnf[2 * ((size_t)tW * p.y + p.x) + 1];
            int tries = 0;
            while (!mtest(smask, sW, sH, nx, ny) && tries < pm->shots) {
                nx = (int)((unsigned)rand() % (unsigned)sW);
                ny = (int)((unsigned)rand() % (unsigned)sH);
                tries++;
            }
            if (mtest(smask, sW, sH, nx, ny)) {
                NX(p.x, p.y) = nx; NY(p.x, p.y) = ny;
                DI(p.x, p.y) = dist_eval(d, nx, ny, p.x, p.y);
            }
        }
    } else {
        for (size_t i = 0; i < tp.size(); i++) {
            Pt p = tp[i];
            int nx = -1, ny = -1;
            int tries = 0;
            while ((nx < 0 || !smask[(size_t)sW * ny + nx]) && tries < pm->shots) {
                nx = (int)((unsigned)rand() % (unsigned)sW);
                ny = (int)((unsigned)rand() % (unsigned)sH);
                tries++;
            }
            if (mtest(smask, sW, sH, nx, ny)) {
                NX(p.x, p.y) = nx; NY(p.x, p.y) = ny;
                DI(p.x, p.y) = dist_eval(d, nx, ny, p.x, p.y);
            }
        }
    }

    const int max_window = (pm->window != -1) ? pm->window : std::max(sW, sH);
    for (int iter = 0; iter < pm->iterations; iter++) {
        double total = 0.0;
        int index, index_end, shift;
        if (iter % 2 == 0) { index = 0; index_end = (int)tp.size(); shift = -1; }
        else { index = (int)tp.size() - 1; index_end = -1; shift = 1; }
        for (; index != index_end; index -= shift) {
            int x = tp[index].x, y = tp[index].y;
            float distance = DI(x, y);
            float original = distance;
            int bx = -1, by = -1;
            if (mtest(tmask, tW, tH, x + shift, y)) {
                int cx = NX(x + shift, y) - shift, cy = NY(x + shift, y);
                if (mtest(smask, sW, sH, cx, cy)) {
                    float cd = dist_eval(d, cx, cy, x, y);
                    if (cd < distance) { distance = cd; bx = cx; by = cy; }
                }
            }
            if (mtest(tmask, tW, tH, x, y + shift)) {
                int cx = NX(x, y + shift), cy = NY(x, y + shift) - shift;
                if (mtest(smask, sW, sH, cx, cy)) {
                    float cd = dist_eval(d, cx, cy, x, y);
                    if (cd < distance) { distance = cd; bx = cx; by = cy; }
                }
            }
            if (bx < 0) { bx = NX(x, y); by = NY(x, y); }
            int scx = bx, scy = by;
            for (int ws = max_window; ws >= 1; ws /= 2) {
                int x_min = std::max(scx - ws, 0), y_min = std::max(scy - ws, 0);
                int x_max = std::min(scx + ws + 1, sW), y_max = std::min(scy + ws + 1, sH);
                for (int k = 0; k < pm->shots; k++) {
                    int cx = x_min + rand() % (x_max - x_min);
                    int cy = y_min + rand() % (y_max - y_min);
                    if (mtest(smask, sW, sH, cx, cy)) {
                        float cd = dist_eval(d, cx, cy, x, y);
                        if (cd < distance) { distance = cd; bx = cx; by = cy; }
                        break;
                    }
                }
            }
            if (original > distance) {
                DI(x, y) = distance;
                NX(x, y) = bx; NY(x, y) = by;
            }
            total += std::min(original, distance);
        }
        if (energy_per_iter) energy_per_iter[iter] = total;
    }
    std::memcpy(out_nnf, nb.data(), nb.size() * sizeof(int32_t));
    if (out_dist) std::memcpy(out_dist, dist.data(), dist.size() * sizeof(float));
    return 0;
}

/* patch_match.cpp:57-302 (OpenMP build): chunks of the row-major target list per
 * thread, odd/even double buffers, chunk-front copy after each sweep; per-thread
 * rand_r() seeds (seed + thread id) replace the time(NULL) base seed (:145).
 * threads = 0 reproduces the reference's thread-count formula (:149). */
static int pm_omp(const nlo_dist_params *dp, const nlo_pm_params *pm,
                  const float *src, const uint8_t *smask, int sW, int sH,
                  const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                  const int32_t *init_nnf, bool reseed, unsigned seed, int threads, int32_t *out_nnf)
{
#ifndef _OPENMP
    (void)threads;
    return pm_serial(dp, pm, src, smask, sW, sH, tgt, tmask, tW, tH, C, init_nnf, reseed, seed,
                     out_nnf, nullptr, nullptr);
#else
    if (reseed) srand(seed);
    DistCtx d;
    d.p = *dp;
    d.weights.resize((size_t)dp->patch * dp->patch);
    nlo_gaussian_weights(dp->patch, dp->patch, dp->sigma, dp->sigma, d.weights.data());
    fvec sg, tg;
    if (needs_gradient(dp->kind)) {
        sg.resize((size_t)sW * sH * C * 2);
        tg.resize((size_t)tW * tH * C * 2);
        nlo_gradient(src, sW, sH, C, sg.data());
        nlo_gradient(tgt, tW, tH, C, tg.data());
    }
    d.src = src; d.tgt = tgt; d.sg = sg.empty() ? nullptr : sg.data(); d.tg = tg.empty() ? nullptr : tg.data();
    d.sW = sW; d.tW = tW; d.C = C;
    size_t tn = (size_t)tW * tH;
    fvec dist_odd(tn, std::numeric_limits<float>::max()), dist_even(tn, std::numeric_limits<float>::max());
    ivec nb_odd(2 * tn, -1), nb_even(2 * tn, -1);
    std::vector<Pt> tp = masked_points(tmask, tW, tH);
    for (size_t i = 0; i < tp.size(); i++) {
        Pt p = tp[i];
        size_t pi = (size_t)tW * p.y + p.x;
        int nx, ny, tries = 0;
        if (init_nnf) {
            nx = init_nnf[2 * pi]; ny = init_nnf[2 * pi + 1];
            while (!mtest(smask, sW, sH, nx, ny) && tries < pm->shots) {
                nx = (int)((unsigned)rand() % (unsigned)sW);
                ny = (int)((unsigned)rand() % (unsigned)sH);
                tries++;
            }
        } else {
            nx = -1; ny = -1;
            while ((nx < 0 || !smask[(size_t)sW * ny + nx]) && tries < pm->shots) {
                nx = (int)((unsigned)rand() % (unsigned)sW);
                ny = (int)((unsigned)rand() % (unsigned)sH);
                tries++;
            }
        }
        if (mtest(smask, sW, sH, nx, ny)) {
            nb_odd[2 * pi] = nb_even[2 * pi] = nx;
            nb_odd[2 * pi + 1] = nb_even[2 * pi + 1] = ny;
            float dd = dist_eval(d, nx, ny, p.x, p.y);
            dist_odd[pi] = dist_even[pi] = dd;
        }
    }
    Pt tl, br;
    bounding_box(tmask, tW, tH, tl, br);
    int domain_width = br.x - tl.x + 1;
    int nt = threads;
    if (nt <= 0) {
        nt = std::min(omp_get_max_threads(), (int)tp.size() / (int)(2 * domain_width));
        if (nt <= 0) nt = omp_get_max_threads();   /* libgomp: num_threads(0) means the default */
    }
    const int max_window = (pm->window != -1) ? pm->window : std::max(sW, sH);
    #pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        int nth = omp_get_num_threads();
        unsigned my_seed = seed + (unsigned)tid;
        int chunk = (int)tp.size() / nth;
        ivec *my_nb = (tid % 2 != 0) ? &nb_odd : &nb_even;
        ivec *ot_nb = (tid % 2 != 0) ? &nb_even : &nb_odd;
        fvec *my_di = (tid % 2 != 0) ? &dist_odd : &dist_even;
        fvec *ot_di = (tid % 2 != 0) ? &dist_even : &dist_odd;
        for (int iter = 0; iter < pm->iterations; iter++) {
            int ib, ie, shift;
            if (iter % 2 == 0) {
                ib = chunk * tid;
                ie = (tid < nth - 1) ? chunk * (tid + 1) : (int)tp.size();
                shift = -1;
            } else {
                ib = (tid < nth - 1) ? chunk * (tid + 1) - 1 : (int)tp.size() - 1;
                ie = chunk * tid - 1;
                shift = 1;
            }
            for (int index = ib; index != ie; index -= shift) {
                int x = tp[index].x, y = tp[index].y;
                size_t pi = (size_t)tW * y + x;
                float distance = (*my_di)[pi];
                float original = distance;
                int bx = -1, by = -1;
                if (mtest(tmask, tW, tH, x + shift, y)) {
                    size_t qi = (size_t)tW * y + (x + shift);
                    int cx = (*my_nb)[2 * qi] - shift, cy = (*my_nb)[2 * qi + 1];
                    if (mtest(smask, sW, sH, cx, cy)) {
                        float cd = dist_eval(d, cx, cy, x, y);
                        if (cd < distance) { distance = cd; bx = cx; by = cy; }
                    }
                }
                if (mtest(tmask, tW, tH, x, y + shift)) {
                    size_t qi = (size_t)tW * (y + shift) + x;
                    int cx = (*my_nb)[2 * qi], cy = (*my_nb)[2 * qi + 1] - shift;
                    if (mtest(smask, sW, sH, cx, cy)) {
                        float cd = dist_eval(d, cx, cy, x, y);
                        if (cd < distance) { distance = cd; bx = cx; by = cy; }
                    }
                }
                if (bx < 0) { bx = (*my_nb)[2 * pi]; by = (*my_nb)[2 * pi + 1]; }
                int scx = bx, scy = by;
                for (int ws = max_window; ws >= 1; ws /= 2) {
                    int x_min = std::max(scx - ws, 0), y_min = std::max(scy - ws, 0);
                    int x_max = std::min(scx + ws + 1, sW), y_max = std::min(scy + ws + 1, sH);
                    for (int k = 0; k < pm->shots; k++) {
                        int cx = x_min + rand_r(&my_seed) % (x_max - x_min);
                        int cy = y_min + rand_r(&my_seed) % (y_max - y_min);
                        if (mtest(smask, sW, sH, cx, cy)) {
                            float cd = dist_eval(d, cx, cy, x, y);
                            if (cd < distance) { distance = cd; bx = cx; by = cy; }
                            break;
                        }
                    }
                }
                if (original > distance) {
                    (*my_di)[pi] = distance;
                    (*my_nb)[2 * pi] = bx; (*my_nb)[2 * pi + 1] = by;
                }
            }
            #pragma omp barrier
            if (iter < pm->iterations - 1) {
                int count = 0;
                for (int index = ie + shift; (index != ib + shift) && (count < domain_width); index += shift, count++) {
                    size_t pi = (size_t)tW * tp[index].y + tp[index].x;
                    (*ot_di)[pi] = (*my_di)[pi];
                    (*ot_nb)[2 * pi] = (*my_nb)[2 * pi];
                    (*ot_nb)[2 * pi + 1] = (*my_nb)[2 * pi + 1];
                }
            } else {
                for (int index = ie + shift; index != ib + shift; index += shift) {
                    size_t pi = (size_t)tW * tp[index].y + tp[index].x;
                    (*ot_nb)[2 * pi] = (*my_nb)[2 * pi];
                    (*ot_nb)[2 * pi + 1] = (*my_nb)[2 * pi + 1];
                }
            }
            #pragma omp barrier
        }
    }
    std::memcpy(out_nnf, nb_odd.data(), nb_odd.size() * sizeof(int32_t));
    return 0;
#endif
}

int nlo_patchmatch(const nlo_dist_params *dp, const nlo_pm_params *pm,
                   const float *src, const uint8_t *smask, int sW, int sH,
                   const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                   const int32_t *init_nnf, unsigned seed,
                   int32_t *out_nnf, float *out_dist, double *energy_per_iter)
{
    return pm_serial(dp, pm, src, smask, sW, sH, tgt, tmask, tW, tH, C, init_nnf, true, seed,
                     out_nnf, out_dist, energy_per_iter);
}

int nlo_patchmatch_omp(const nlo_dist_params *dp, const nlo_pm_params *pm,
                       const float *src, const uint8_t *smask, int sW, int sH,
                       const float *tgt, const uint8_t *tmask, int tW, int tH, int C,
                       const int32_t *init_nnf, unsigned seed, int threads, int32_t *out_nnf)
{
    return pm_omp(dp, pm, src, smask, sW, sH, tgt, tmask, tW, tH, C, init_nnf, true, seed, threads,
                  out_nnf);
}

/* patch_non_local_poisson.cpp:139-232 (scatter over extended-domain centres, no
 * (-1,-1) check on the NNF entry, :187) and :235-285 (backward divergence). */
void nlo_poisson_coefficients(int P, float sigma, float lambda, const float *image,
                              const uint8_t *ext, const int32_t *nnf, const float *conf,
                              int W, int H, int C, double *a1, double *a2, double *f1, double *f2)
{
    size_t n = (size_t)W * H;
    fvec w((size_t)P * P);
    nlo_gaussian_weights(P, P, sigma, sigma, w.data());
    fvec grad(n * C * 2);
    nlo_gradient(image, W, H, C, grad.data());
    std::vector<double> F1(n * C, 0.0), F2(n * C, 0.0);
    for (size_t i = 0; i < n; i++) { a1[i] = 0; a2[i] = 0; }
    for (size_t i = 0; i < n * C; i++) { f1[i] = 0; f2[i] = 0; }
    int hx = P / 2, hy = P / 2;
    for (int py = 0; py < H; py++)
        for (int px = 0; px < W; px++) {
            size_t pidx = (size_t)W * py + px;
            if (!ext[pidx]) continue;
            int x_first = std::max(px - hx, 0), x_last = std::min(px + hx, W - 1);
            int y_first = std::max(py - hy, 0), y_last = std::min(py + hy, H - 1);
            int nbx = nnf[2 * pidx], nby = nnf[2 * pidx + 1];
            float confidence = conf ? conf[pidx] : 1.0f;
            for (int x = x_first; x <= x_last; x++)
                for (int y = y_first; y <= y_last; y++) {
                    int n_x = nbx + x - px, n_y = nby + y - py;
                    if (n_x < 0 || n_y < 0 || n_x >= W || n_y >= H) continue;
                    size_t index = (size_t)W * y + x;
                    double weight = (double)(confidence * w[(size_t)P * (y - py + hy) + (x - px + hx)]);
                    a1[index] += weight * lambda;
                    a2[index] += weight * (1 - lambda);
                    size_t nidx = (size_t)W * n_y + n_x;
                    for (int ch = 0; ch < C; ch++) {
                        f1[(size_t)ch * n + index] += weight * image[nidx * C + ch] * lambda;
                        F1[(size_t)ch * n + index] += weight * grad[(nidx * C + ch) * 2] * (1 - lambda);
                        F2[(size_t)ch * n + index] += weight * grad[(nidx * C + ch) * 2 + 1] * (1 - lambda);
                    }
                }
        }
    if (lambda != 1) {
        for (int ch = 0; ch < C; ch++) {
            const double *fx = &F1[(size_t)ch * n], *fy = &F2[(size_t)ch * n];
            double *dv = &f2[(size_t)ch * n];
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    size_t idx = (size_t)W * y + x;
                    if (!ext[idx]) continue;
                    if (x == 0) dv[idx] = fx[idx];
                    else if (x == W - 1) dv[idx] = -fx[idx - 1];
                    else dv[idx] = fx[idx] - fx[idx - 1];
                    if (y == 0) dv[idx] += fy[idx];
                    else if (y == H - 1) dv[idx] += -fy[idx - W];
                    else dv[idx] += fy[idx] - fy[idx - W];
                }
        }
    }
}

/* patch_non_local_poisson.cpp:291-356: right/bottom use a(centre), left/top use
 * a(left)/a(top); zero outside the image; evaluated at mask points, written back
 * after all are computed. */
void nlo_anisotropic_laplacian(double *u, const double *a, const uint8_t *mask, int W, int H)
{
    std::vector<Pt> pts = masked_points(mask, W, H);
    std::vector<double> buf(pts.size());
    for (size_t i = 0; i < pts.size(); i++) {
        int x = pts[i].x, y = pts[i].y;
        size_t idx = (size_t)W * y + x;
        double a_t, a_b, a_r, a_l, u_t, u_b, u_r, u_l;
        if (y != 0) { a_t = a[idx - W]; u_t = u[idx - W]; } else { a_t = 0; u_t = 0; }
        if (x != 0) { a_l = a[idx - 1]; u_l = u[idx - 1]; } else { a_l = 0; u_l = 0; }
        if (y != H - 1) { a_b = a[idx]; u_b = u[idx + W]; } else { a_b = 0; u_b = 0; }
        if (x != W - 1) { a_r = a[idx]; u_r = u[idx + 1]; } else { a_r = 0; u_r = 0; }
        buf[i] = u_t * a_t + u_l * a_l + u_b * a_b + u_r * a_r - u[idx] * (a_t + a_l + a_b + a_r);
    }
    for (size_t i = 0; i < pts.size(); i++) u[(size_t)W * pts[i].y + pts[i].x] = buf[i];
}

/* patch_non_local_poisson.cpp:387-477: 2 matvecs per iteration (r recomputed as
 * b - A x), stop on nr < tol*nr0 (squared norms) or iteration >= limit, so at most
 * limit+1 iterations. Returns the number of iterations run. */
int nlo_conjugate_gradient(double *x, const double *a1, const double *a2, const double *b,
                           const uint8_t *mask, int W, int H, float tol, int limit,
                           double *nr0_out, double *nr_out)
{
    size_t n = (size_t)W * H;
    std::vector<Pt> pts = masked_points(mask, W, H);
    std::vector<double> al(n), d(n, 0.0), r(pts.size(), 0.0);
    std::memcpy(al.data(), x, n * sizeof(double));
    nlo_anisotropic_laplacian(al.data(), a2, mask, W, H);
    double nr0 = 0.0;
    for (size_t i = 0; i < pts.size(); i++) {
        size_t idx = (size_t)W * pts[i].y + pts[i].x;
        r[i] = b[idx] - a1[idx] * x[idx] + al[idx];
        d[idx] = r[i];
        nr0 += r[i] * r[i];
    }
    int iteration = 0;
    double nr = nr0;
    bool stop = !(nr0 > 0);
    while (!stop) {
        std::memcpy(al.data(), d.data(), n * sizeof(double));
        nlo_anisotropic_laplacian(al.data(), a2, mask, W, H);
        double rdotq = 0.0;
        for (size_t i = 0; i < pts.size(); i++) {
            size_t idx = (size_t)W * pts[i].y + pts[i].x;
            rdotq += r[i] * (a1[idx] * d[idx] - al[idx]);
        }
        double alpha = nr / rdotq;
        for (size_t i = 0; i < pts.size(); i++) {
            size_t idx = (size_t)W * pts[i].y + pts[i].x;
            x[idx] += alpha * d[idx];
        }
        double nr_old = nr;
        std::memcpy(al.data(), x, n * sizeof(double));
        nlo_anisotropic_laplacian(al.data(), a2, mask, W, H);
        nr = 0.0;
        for (size_t i = 0; i < pts.size(); i++) {
            size_t idx = (size_t)W * pts[i].y + pts[i].x;
            r[i] = b[idx] - a1[idx] * x[idx] + al[idx];
            nr += r[i] * r[i];
        }
        double beta = nr / nr_old;
        for (size_t i = 0; i < pts.size(); i++) {
            size_t idx = (size_t)W * pts[i].y + pts[i].x;
            d[idx] = r[i] + beta * d[idx];
        }
        stop = ((nr < tol * nr0) || (iteration >= limit));
        iteration++;
    }
    if (nr0_out) *nr0_out = nr0;
    if (nr_out) *nr_out = nr;
    return iteration;
}

/* a_image_updating.h:34-39 dispatch over the four updaters;
 * patch_non_local_depth.cpp:36-55 for NLO_UPD_DEPTH (orig split from image, :49).
 * Returns NaN when the channel count does not fit the updater. */
double nlo_update(const nlo_upd_params *u, float *image, const float *orig,
                  const uint8_t *domain, const uint8_t *ext, const int32_t *nnf,
                  const float *conf, int W, int H, int C, nlo_cg_stats *stats)
{
    if (stats) std::memset(stats, 0, sizeof(*stats));
    fvec w((size_t)u->patch * u->patch);
    nlo_gaussian_weights(u->patch, u->patch, u->sigma, u->sigma, w.data());
    switch (u->kind) {
    case NLO_UPD_MEANS:
        return nlmeans(u->patch, w.data(), image, domain, nnf, conf, W, H, C);
    case NLO_UPD_MEDIANS:
        return nlmedians(u->patch, w.data(), image, domain, nnf, conf, W, H, C);
    case NLO_UPD_POISSON:
        return nlpoisson(u->patch, u->sigma, u->lambda, u->cg_tol, u->cg_limit, image, orig, domain,
                         ext, nnf, conf, W, H, C, stats, 0);
    case NLO_UPD_DEPTH: {
        if (C != 4) return NAN;
        size_t n = (size_t)W * H;
        fvec rgb(n * 3), dep(n), rgb_o(n * 3), dep_o(n);
        for (size_t i = 0; i < n; i++) {
            for (int c = 0; c < 3; c++) rgb[i * 3 + c] = rgb_o[i * 3 + c] = image[i * 4 + c];
            dep[i] = dep_o[i] = image[i * 4 + 3];
        }
        double total = 0;
        total += nlpoisson(u->patch, u->sigma, u->lambda, 0.000001f, 1000, rgb.data(), rgb_o.data(),
                           domain, ext, nnf, conf, W, H, 3, stats, 0);
        total += nlpoisson(u->patch, u->sigma, u->lambda_d, 0.000001f, 1000, dep.data(), dep_o.data(),
                           domain, ext, nnf, conf, W, H, 1, stats, 3);
        for (size_t i = 0; i < n; i++) {
            for (int c = 0; c < 3; c++) image[i * 4 + c] = rgb[i * 3 + c];
            image[i * 4 + 3] = dep[i];
        }
        (void)orig;
        return total;
    }
    }
    return NAN;
}

/* 3rdparty/simpois/simpois.c:281-291: per channel (planar), poisson_rec(.25, 10, 99, 0) */
void nlo_simpois(float *out, const float *in, int w, int h, int pd)
{
    for (int l = 0; l < pd; l++)
        sp_rec(out + (size_t)w * h * l, in + (size_t)w * h * l, w, h, .25f, 10, 99);
}

/* image_inpainting.cpp:68-190 (process), :349-417 (inpaint_internal),
 * :492-557 (propagate_weights), main.cpp:245-266 (method pairing).
 * Returns 0, or -2 for an empty source mask (the reference throws, :382-384),
 * -1 for bad arguments. */
int nlo_process(const nlo_process_params *p, const float *in, const uint8_t *mask,
                int W, int H, int C, float *out, nlo_process_stats *st)
{
    if (!p || p->scales < 1 || p->scales > 32) return -1;
    double t_start = now_s(), t_pyr = 0, t_pm = 0, t_upd = 0;
    nlo_process_stats local;
    if (!st) st = &local;
    std::memset(st, 0, sizeof(*st));

    nlo_dist_params dp;
    nlo_upd_params up;
    std::memset(&dp, 0, sizeof(dp));
    std::memset(&up, 0, sizeof(up));
    dp.patch = up.patch = p->patch;
    dp.sigma = up.sigma = p->sigma;
    switch (p->method) {
    case NLO_UPD_MEANS: dp.kind = NLO_DIST_L2; up.kind = NLO_UPD_MEANS; break;
    case NLO_UPD_MEDIANS: dp.kind = NLO_DIST_L1; up.kind = NLO_UPD_MEDIANS; break;
    case NLO_UPD_POISSON:
        dp.kind = NLO_DIST_L2COMBINED; dp.lambda = p->lambda;
        up.kind = NLO_UPD_POISSON; up.lambda = p->lambda; up.cg_tol = p->cg_tol; up.cg_limit = p->cg_limit;
        break;
    case NLO_UPD_DEPTH:
        if (C != 4) return -1;
        dp.kind = NLO_DIST_L2RGBD; dp.lambda = p->lambda_rgb; dp.lambda_d = p->lambda_d;
        up.kind = NLO_UPD_DEPTH; up.lambda = p->lambda_rgb; up.lambda_d = p->lambda_d;
        break;
    default: return -1;
    }
    nlo_pm_params pm = {p->pm_iters, p->pm_shots, p->pm_window};
    float rate = nlo_subsampling_rate(p->coarse, p->scales);
    const int S = p->scales;
    const int P = p->patch;
    unsigned pm_calls = 0;
    srand(p->seed);   /* main.cpp:273 seeds once; every PatchMatch call continues the stream */

    auto run_pm = [&](const float *img, const uint8_t *sm, const uint8_t *tm, int w, int h,
                      const int32_t *init, int32_t *outn) {
        double t0 = now_s();
        unsigned sd = p->seed + 7919u * (pm_calls++);
        if (p->pm_threads > 0)
            pm_omp(&dp, &pm, img, sm, w, h, img, tm, w, h, C, init, false, sd, p->pm_threads, outn);
        else
            pm_serial(&dp, &pm, img, sm, w, h, img, tm, w, h, C, init, false, 0, outn, nullptr, nullptr);
        t_pm += now_s() - t0;
    };
    auto run_upd = [&](float *img, const float *orig, const uint8_t *hole, const uint8_t *ext,
                       const int32_t *nnf, const float *conf, int w, int h) {
        double t0 = now_s();
        double r = nlo_update(&up, img, orig, hole, ext, nnf, conf, w, h, C, nullptr);
        t_upd += now_s() - t0;
        return r;
    };

    std::vector<fvec> pyr(S);
    std::vector<bvec> mpyr(S);
    std::vector<int> ws(S), hs(S);
    ws[0] = W; hs[0] = H;
    pyr[0].assign(in, in + (size_t)W * H * C);
    mpyr[0].assign(mask, mask + (size_t)W * H);
    double t0 = now_s();
    for (int i = 1; i < S; i++) {
        int w = nlo_sample_size(ws[i - 1], rate), h = nlo_sample_size(hs[i - 1], rate);
        ws[i] = w; hs[i] = h;
        pyr[i].resize((size_t)w * h * C);
        mpyr[i].resize((size_t)w * h);
        if (p->init != NLO_INIT_NONE) {
            nlo_downsample_with_mask(pyr[i - 1].data(), mpyr[i - 1].data(), ws[i - 1], hs[i - 1], C, rate,
                                     0.4f, 0, pyr[i].data(), mpyr[i].data());
        } else {
            nlo_downsample(pyr[i - 1].data(), ws[i - 1], hs[i - 1], C, rate, pyr[i].data());
            nlo_downsample_mask(mpyr[i - 1].data(), ws[i - 1], hs[i - 1], rate, 0.4f, mpyr[i].data());
        }
    }
    t_pyr += now_s() - t0;
    {
        /* coarsest-level initialisation, :104-154 */
        fvec &img = pyr[S - 1];
        bvec &m = mpyr[S - 1];
        int w = ws[S - 1], h = hs[S - 1];
        if (p->init == NLO_INIT_AVG || p->init == NLO_INIT_BLACK) {
            std::vector<float> color(C, 0.0f);
            if (p->init == NLO_INIT_AVG)
                for (int c = 0; c < C; c++) {
                    float sum = 0.0f;
                    int count = 0;
                    for (int y = 0; y < h; y++)
                        for (int x = 0; x < w; x++)
                            if (!m[(size_t)w * y + x]) { sum += img[((size_t)w * y + x) * C + c]; count++; }
                    if (count > 0) color[c] = sum / (float)count;
                }
            for (size_t i = 0; i < (size_t)w * h; i++)
                if (m[i]) for (int c = 0; c < C; c++) img[i * C + c] = color[c];
        } else if (p->init == NLO_INIT_POISSON) {
            size_t n = (size_t)w * h;
            fvec pin(n * C), pout(n * C);
            for (int c = 0; c < C; c++)
                for (size_t i = 0; i < n; i++) pin[i + c * n] = m[i] ? NAN : img[i * C + c];
            nlo_simpois(pout.data(), pin.data(), w, h, C);
            for (int c = 0; c < C; c++)
                for (size_t i = 0; i < n; i++) img[i * C + c] = pout[i + c * n];
        }
    }

    int cap = p->iters;
    if (p->max_iters_per_scale > 0 && p->max_iters_per_scale < cap) cap = p->max_iters_per_scale;

    /* inpaint_internal, :349-417 */
    auto inpaint_internal = [&](int level, const float *conf, ivec &nnf) -> int {
        int w = ws[level], h = hs[level];
        size_t n = (size_t)w * h;
        bvec ext(n), tgt(n), srcm(n);
        int empty = nlo_scale_masks(mpyr[level].data(), w, h, P, 1, ext.data(), tgt.data(), srcm.data());
        if (empty) return -2;
        fvec original = pyr[level];
        double total = std::numeric_limits<double>::max();
        int i = 0;
        const double t_level = now_s();
        ivec next(2 * n);
        for (i = 0; i < cap && total > p->tol; i++) {
            run_pm(pyr[level].data(), srcm.data(), tgt.data(), w, h, nnf.empty() ? nullptr : nnf.data(),
                   next.data());
            nnf = next;
            total = run_upd(pyr[level].data(), original.data(), mpyr[level].data(), ext.data(), nnf.data(),
                            conf, w, h);
        }
        st->iterations[level] = i;
        st->seconds_iterations[level] = now_s() - t_level;
        return 0;
    };

    fvec conf((size_t)ws[S - 1] * hs[S - 1]);
    nlo_confidence(mpyr[S - 1].data(), ws[S - 1], hs[S - 1], p->conf_t, p->conf_a, conf.data());
    ivec nnf;
    if (inpaint_internal(S - 1, conf.data(), nnf) != 0) return -2;
    for (int i = S - 2; i >= 0; i--) {
        int uw = ws[i], uh = hs[i], lw = ws[i + 1], lh = hs[i + 1];
        size_t un = (size_t)uw * uh, ln = (size_t)lw * lh;
        conf.assign(un, 0.0f);
        const double t_prop = now_s();
        nlo_confidence(mpyr[i].data(), uw, uh, p->conf_t, p->conf_a, conf.data());
        /* propagate_weights, :492-557 */
        bvec lext(ln), ltgt(ln), lsrc(ln);
        nlo_scale_masks(mpyr[i + 1].data(), lw, lh, P, 0, lext.data(), ltgt.data(), lsrc.data());
        ivec lnnf(2 * ln);
        run_pm(pyr[i + 1].data(), lsrc.data(), ltgt.data(), lw, lh, nullptr, lnnf.data());
        ivec snnf(2 * un);
        nlo_scale_nnf(lnnf.data(), mpyr[i + 1].data(), lw, lh, uw, uh, snnf.data());
        fvec upimg(un * C);
        nlo_upsample(pyr[i + 1].data(), lw, lh, C, uw, uh, upimg.data());
        for (size_t k = 0; k < un; k++)
            if (mpyr[i][k]) for (int c = 0; c < C; c++) pyr[i][k * C + c] = upimg[k * C + c];
        bvec uext(un), utgt(un), usrc(un);
        nlo_scale_masks(mpyr[i].data(), uw, uh, P, 0, uext.data(), utgt.data(), usrc.data());
        ivec rnnf(2 * un);
        run_pm(upimg.data(), usrc.data(), utgt.data(), uw, uh, snnf.data(), rnnf.data());
        {
            fvec orig_copy = pyr[i];   /* image and orig share one buffer in the reference (:554) */
            run_upd(pyr[i].data(), orig_copy.data(), mpyr[i].data(), uext.data(), rnnf.data(), conf.data(), uw, uh);
        }
        nnf = rnnf;
        st->seconds_propagate[i] = now_s() - t_prop;
        if (inpaint_internal(i, conf.data(), nnf) != 0) return -2;
    }
    std::memcpy(out, pyr[0].data(), (size_t)W * H * C * sizeof(float));
    for (int i = 0; i < S; i++) { st->width[i] = ws[i]; st->height[i] = hs[i]; }
    st->seconds_pyramid = t_pyr;
    st->seconds_patchmatch = t_pm;
    st->seconds_update = t_upd;
    st->seconds_other = now_s() - t_start - t_pyr - t_pm - t_upd;
    return 0;
}

}  /* extern "C" */
