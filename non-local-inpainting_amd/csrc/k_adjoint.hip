// k_adjoint.hip -- the patch energy of NL-Poisson, its gradient with respect to the
// pixels, and the patch extraction / aggregation operator pair with its hand-written
// adjoint (gfx950).
//
// For a fixed NNF phi and confidence c, the NL-Poisson update minimises over the hole
// pixels the patch energy
//
//   E(u) = sum_{p in ext} sum_{h valid} w(p,h) sum_ch [ lambda (u(p+h) - u(s))^2
//          + (1-lambda) ( (dx u(p+h) - dx u(s))^2 + (dy u(p+h) - dy u(s))^2 ) ],
//   s = phi(p) + h,  w(p,h) = (double)(c(p) * g(h))   (float product, as the reference),
//   lambda and 1 - lambda taken in float as the reference's coefficients take them,
//
// with forward differences dx u(x,y) = u(x+1,y) - u(x,y) (0 on the last column, likewise
// dy), and (p,h) valid when p+h and phi(p)+h are inside the image.  These are exactly
// the terms the reference accumulates into a1, a2, f1, F (patch_non_local_poisson.cpp:
// 139-232, no (-1,-1) check on phi, :187), so dE/du = 2 (a1 u - f1 - L_a2 u + div F)
// on every pixel: twice the residual of the reference's linear system (:36-121,
// :235-356) on the hole pixels.
//
// Operators (u per pixel and channel, patches per (centre e, offset h, channel)):
//   extract_t  (Et u)[e,h,ch] = u(p_e + h)        extract_s  (Es u)[e,h,ch] = u(phi(p_e) + h)
//   aggregate_t = Et^T: each pixel q gathers w[e,h] from the centres p_e = q - h (the
//                 gather form of the scatter, deterministic order)
//   aggregate_s = Es^T: scatter-add to phi(p_e) + h (FP64 atomics)
// The gradient kernels apply the same pair without materialising the patches: the
// per-(p,h) residuals are recomputed from the NNF inside the aggregation.
#include "nli_internal.h"

#include <cfloat>
#include <vector>

namespace nli {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxP = 15;
constexpr int kMaxC = 4;

unsigned grid_n(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct EArgs {
    const float *img;
    const uint8_t *ext;
    const int2 *nnf;
    const float *conf;
    const float *w;   // P x P Gaussian weights
    int P, W, H, C;
    double lam, lam1;   // lambda and (1 - lambda) as the reference forms them in float
    int tx, ty, tc;     // perturbed pixel/channel (tx < 0: none)
    double delta;
};

__device__ inline double val(const EArgs &a, int x, int y, int c)
{
    double v = (double)a.img[((size_t)a.W * y + x) * a.C + c];
    if (x == a.tx && y == a.ty && c == a.tc) v += a.delta;
    return v;
}
__device__ inline double dxv(const EArgs &a, int x, int y, int c)
{
    return x + 1 < a.W ? val(a, x + 1, y, c) - val(a, x, y, c) : 0.0;
}
__device__ inline double dyv(const EArgs &a, int x, int y, int c)
{
    return y + 1 < a.H ? val(a, x, y + 1, c) - val(a, x, y, c) : 0.0;
}
// does a term reading u at (x,y), (x+1,y), (x,y+1) depend on the perturbed pixel?
__device__ inline bool touches(const EArgs &a, int x, int y)
{
    return (a.tx == x && a.ty == y) || (a.tx == x + 1 && a.ty == y && x + 1 < a.W) ||
           (a.tx == x && a.ty == y + 1 && y + 1 < a.H);
}

__device__ inline void block_sum_store(double v, double *partials)
{
    __shared__ double sh[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int i = 1; i < kBlock / 64; i++) t += sh[i];
        partials[blockIdx.x] = t;
    }
}

// energy (or, with a perturbed pixel, only the terms that read it): one thread per
// extended-domain centre, block partials summed on the host in block order
__global__ __launch_bounds__(kBlock) void k_energy(EArgs a, const int *__restrict__ pts, int n,
                                                   double *__restrict__ partials)
{
    __shared__ float wl[kMaxP * kMaxP];
    for (int i = threadIdx.x; i < a.P * a.P; i += blockDim.x) wl[i] = a.w[i];
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (k < n) {
        const int p = pts[k];
        const int px = p % a.W, py = p / a.W;
        const int2 f = a.nnf[p];
        const float cw = a.conf ? a.conf[p] : 1.0f;
        const int r = a.P / 2;
        for (int oy = -r; oy <= r; oy++)
            for (int ox = -r; ox <= r; ox++) {
                const int tx = px + ox, ty = py + oy, sx = f.x + ox, sy = f.y + oy;
                if ((unsigned)tx >= (unsigned)a.W || (unsigned)ty >= (unsigned)a.H) continue;
                if ((unsigned)sx >= (unsigned)a.W || (unsigned)sy >= (unsigned)a.H) continue;
                if (a.tx >= 0 && !touches(a, tx, ty) && !touches(a, sx, sy)) continue;
                const double wt = (double)(cw * wl[a.P * (oy + r) + (ox + r)]);
                double t = 0.0;
                for (int c = 0; c < a.C; c++) {
                    const double d0 = val(a, tx, ty, c) - val(a, sx, sy, c);
                    const double gx = dxv(a, tx, ty, c) - dxv(a, sx, sy, c);
                    const double gy = dyv(a, tx, ty, c) - dyv(a, sx, sy, c);
                    t += a.lam * d0 * d0 + a.lam1 * (gx * gx + gy * gy);
                }
                acc += wt * t;
            }
    }
    block_sum_store(acc, partials);
}

// gradient, target side (gather over the centres p = q - h): S = sum 2 w lam (u(q) - u(s)),
// Rx/Ry = sum 2 w (1-lam) (d u(q) - d u(s)); planes [c * W*H + q]
__global__ __launch_bounds__(kBlock) void k_grad_gather(EArgs a, double *__restrict__ S, double *__restrict__ Rx,
                                                        double *__restrict__ Ry)
{
    __shared__ float wl[kMaxP * kMaxP];
    for (int i = threadIdx.x; i < a.P * a.P; i += blockDim.x) wl[i] = a.w[i];
    __syncthreads();
    const int qx = blockIdx.x * blockDim.x + threadIdx.x, qy = blockIdx.y;
    if (qx >= a.W) return;
    const size_t plane = (size_t)a.W * a.H, q = (size_t)a.W * qy + qx;
    const int r = a.P / 2;
    double s0[kMaxC] = {0, 0, 0, 0}, sx_[kMaxC] = {0, 0, 0, 0}, sy_[kMaxC] = {0, 0, 0, 0};
    double uq[kMaxC], gq[kMaxC], hq[kMaxC];
    for (int c = 0; c < a.C; c++) {
        uq[c] = val(a, qx, qy, c);
        gq[c] = dxv(a, qx, qy, c);
        hq[c] = dyv(a, qx, qy, c);
    }
    for (int oy = -r; oy <= r; oy++)
        for (int ox = -r; ox <= r; ox++) {
            const int px = qx - ox, py = qy - oy;
            if ((unsigned)px >= (unsigned)a.W || (unsigned)py >= (unsigned)a.H) continue;
            const size_t p = (size_t)a.W * py + px;
            if (!a.ext[p]) continue;
            const int2 f = a.nnf[p];
            const int sx = f.x + ox, sy = f.y + oy;
            if ((unsigned)sx >= (unsigned)a.W || (unsigned)sy >= (unsigned)a.H) continue;
            const float cw = a.conf ? a.conf[p] : 1.0f;
            const double wt = (double)(cw * wl[a.P * (oy + r) + (ox + r)]);
            for (int c = 0; c < a.C; c++) {
                s0[c] += 2.0 * wt * a.lam * (uq[c] - val(a, sx, sy, c));
                sx_[c] += 2.0 * wt * a.lam1 * (gq[c] - dxv(a, sx, sy, c));
                sy_[c] += 2.0 * wt * a.lam1 * (hq[c] - dyv(a, sx, sy, c));
            }
        }
    for (int c = 0; c < a.C; c++) {
        S[c * plane + q] = s0[c];
        Rx[c * plane + q] = sx_[c];
        Ry[c * plane + q] = sy_[c];
    }
}

// gradient, source side (scatter to s = phi(p) + h): one thread per centre
__global__ __launch_bounds__(kBlock) void k_grad_scatter(EArgs a, const int *__restrict__ pts, int n,
                                                         double *__restrict__ S, double *__restrict__ Rx,
                                                         double *__restrict__ Ry)
{
    __shared__ float wl[kMaxP * kMaxP];
    for (int i = threadIdx.x; i < a.P * a.P; i += blockDim.x) wl[i] = a.w[i];
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t plane = (size_t)a.W * a.H;
    const int p = pts[k];
    const int px = p % a.W, py = p / a.W;
    const int2 f = a.nnf[p];
    const float cw = a.conf ? a.conf[p] : 1.0f;
    const int r = a.P / 2;
    for (int oy = -r; oy <= r; oy++)
        for (int ox = -r; ox <= r; ox++) {
            const int tx = px + ox, ty = py + oy, sx = f.x + ox, sy = f.y + oy;
            if ((unsigned)tx >= (unsigned)a.W || (unsigned)ty >= (unsigned)a.H) continue;
            if ((unsigned)sx >= (unsigned)a.W || (unsigned)sy >= (unsigned)a.H) continue;
            const double wt = (double)(cw * wl[a.P * (oy + r) + (ox + r)]);
            const size_t s = (size_t)a.W * sy + sx;
            for (int c = 0; c < a.C; c++) {
                atomicAdd(&S[c * plane + s], -2.0 * wt * a.lam * (val(a, tx, ty, c) - val(a, sx, sy, c)));
                atomicAdd(&Rx[c * plane + s],
                          -2.0 * wt * a.lam1 * (dxv(a, tx, ty, c) - dxv(a, sx, sy, c)));
                atomicAdd(&Ry[c * plane + s],
                          -2.0 * wt * a.lam1 * (dyv(a, tx, ty, c) - dyv(a, sx, sy, c)));
            }
        }
}

// dE/du(q) = S(q) + Rx(q - ex) - Rx(q) + Ry(q - ey) - Ry(q): the adjoint of the forward
// differences (dx u(q') reads u(q'+ex) with +1 and u(q') with -1, and is 0 on the last
// column); output HWC
__global__ void k_grad_combine(const double *__restrict__ S, const double *__restrict__ Rx,
                               const double *__restrict__ Ry, int W, int H, int C, double *__restrict__ grad)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const size_t plane = (size_t)W * H, q = (size_t)W * y + x;
    for (int c = 0; c < C; c++) {
        const double *rx = Rx + c * plane, *ry = Ry + c * plane;
        double g = S[c * plane + q];
        if (x >= 1) g += rx[q - 1];
        if (x + 1 < W) g -= rx[q];
        if (y >= 1) g += ry[q - W];
        if (y + 1 < H) g -= ry[q];
        grad[q * C + c] = g;
    }
}

// ---- the operator pair on materialised patches (u and patches in FP64) ----
// patches[(e * P*P + h) * C + c]; invalid (p,h) hold 0
__global__ void k_extract(const double *__restrict__ u, const int *__restrict__ pts, int n, const int2 *__restrict__ nnf,
                          int P, int W, int H, int C, int side, double *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t PP = (size_t)P * P;
    if (i >= (size_t)n * PP) return;
    const int e = (int)(i / PP), hh = (int)(i % PP);
    const int r = P / 2, ox = hh % P - r, oy = hh / P - r;
    const int p = pts[e];
    const int px = p % W, py = p / W;
    const int2 f = nnf[p];
    const int tx = px + ox, ty = py + oy, sx = f.x + ox, sy = f.y + oy;
    const bool ok = (unsigned)tx < (unsigned)W && (unsigned)ty < (unsigned)H && (unsigned)sx < (unsigned)W &&
                    (unsigned)sy < (unsigned)H;
    const size_t src = !ok ? 0 : (side == 0 ? (size_t)W * ty + tx : (size_t)W * sy + sx);
    for (int c = 0; c < C; c++) out[i * C + c] = ok ? u[src * C + c] : 0.0;
}

// Et^T as a gather: pixel q sums the patch entries of the centres q - h
__global__ void k_aggregate_t(const double *__restrict__ patches, const int *__restrict__ lidx,
                              const int2 *__restrict__ nnf, int P, int W, int H, int C, double *__restrict__ out)
{
    const int qx = blockIdx.x * blockDim.x + threadIdx.x, qy = blockIdx.y;
    if (qx >= W) return;
    const int r = P / 2;
    double acc[kMaxC] = {0, 0, 0, 0};
    for (int oy = -r; oy <= r; oy++)
        for (int ox = -r; ox <= r; ox++) {
            const int px = qx - ox, py = qy - oy;
            if ((unsigned)px >= (unsigned)W || (unsigned)py >= (unsigned)H) continue;
            const int e = lidx[(size_t)W * py + px];
            if (e < 0) continue;
            const int2 f = nnf[(size_t)W * py + px];
            const int sx = f.x + ox, sy = f.y + oy;
            if ((unsigned)sx >= (unsigned)W || (unsigned)sy >= (unsigned)H) continue;
            const size_t i = (size_t)e * P * P + (size_t)P * (oy + r) + (ox + r);
            for (int c = 0; c < C; c++) acc[c] += patches[i * C + c];
        }
    const size_t q = (size_t)W * qy + qx;
    for (int c = 0; c < C; c++) out[q * C + c] = acc[c];
}

// Es^T: scatter-add to phi(p) + h
__global__ void k_aggregate_s(const double *__restrict__ patches, const int *__restrict__ pts, int n,
                              const int2 *__restrict__ nnf, int P, int W, int H, int C, double *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t PP = (size_t)P * P;
    if (i >= (size_t)n * PP) return;
    const int e = (int)(i / PP), hh = (int)(i % PP);
    const int r = P / 2, ox = hh % P - r, oy = hh / P - r;
    const int p = pts[e];
    const int px = p % W, py = p / W;
    const int2 f = nnf[p];
    const int tx = px + ox, ty = py + oy, sx = f.x + ox, sy = f.y + oy;
    if ((unsigned)tx >= (unsigned)W || (unsigned)ty >= (unsigned)H || (unsigned)sx >= (unsigned)W ||
        (unsigned)sy >= (unsigned)H)
        return;
    const size_t s = (size_t)W * sy + sx;
    for (int c = 0; c < C; c++) atomicAdd(&out[s * C + c], patches[i * C + c]);
}

__global__ void k_lidx_fill(const int *pts, int n, int *lidx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) lidx[pts[k]] = k;
}

void check_desc(const EnergyDesc &d, int C)
{
    if (d.patch < 1 || d.patch > kMaxP || d.patch % 2 == 0)
        throw std::invalid_argument("patch side must be odd and at most 15");
    if (C < 1 || C > kMaxC) throw std::invalid_argument("1 to 4 channels");
}

EArgs make_args(const EnergyDesc &d, const float *img, const uint8_t *ext, const int32_t *nnf, const float *conf,
                int W, int H, int C, AdjWork &w, hipStream_t s)
{
    if (w.w_patch != d.patch || w.w_sigma != d.sigma) {
        std::vector<float> wh((size_t)d.patch * d.patch);
        gaussian_weights(d.patch, d.patch, d.sigma, d.sigma, wh.data());
        w.weights.ensure(wh.size());
        NLI_HIP(hipMemcpyAsync(w.weights.p, wh.data(), wh.size() * sizeof(float), hipMemcpyHostToDevice, s));
        NLI_HIP(hipStreamSynchronize(s));
        w.w_patch = d.patch;
        w.w_sigma = d.sigma;
    }
    EArgs a{img, ext, reinterpret_cast<const int2 *>(nnf), conf, w.weights.p, d.patch, W, H, C,
            (double)d.lambda, (double)(1.0f - d.lambda), -1, -1, -1, 0.0};
    return a;
}

int compact_ext(const uint8_t *ext, int W, int H, AdjWork &w, hipStream_t s)
{
    const size_t n = (size_t)W * H;
    w.pts.ensure(n);
    w.counts.ensure(n / 1024 + 2);
    w.total.ensure(2);
    return compact_indices(ext, (int)n, w.pts.p, w.counts.p, w.total.p, s);
}

}  // namespace

double patch_energy(const EnergyDesc &d, const float *img, const uint8_t *ext, const int32_t *nnf, const float *conf,
                    int W, int H, int C, int tx, int ty, int tc, double delta, AdjWork &w, hipStream_t s)
{
    check_desc(d, C);
    EArgs a = make_args(d, img, ext, nnf, conf, W, H, C, w, s);
    if (tx >= 0) {
        if (tx >= W || ty < 0 || ty >= H || tc < 0 || tc >= C) throw std::invalid_argument("perturbed pixel outside");
        a.tx = tx;
        a.ty = ty;
        a.tc = tc;
        a.delta = delta;
    }
    const int n = compact_ext(ext, W, H, w, s);
    if (n == 0) return 0.0;
    const unsigned nb = grid_n((size_t)n);
    w.partials.ensure(nb);
    k_energy<<<nb, kBlock, 0, s>>>(a, w.pts.p, n, w.partials.p);
    NLI_LAUNCH_CHECK();
    std::vector<double> part(nb);
    NLI_HIP(hipMemcpyAsync(part.data(), w.partials.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    NLI_HIP(hipStreamSynchronize(s));
    double e = 0.0;
    for (double v : part) e += v;
    return e;
}

void energy_gradient(const EnergyDesc &d, const float *img, const uint8_t *ext, const int32_t *nnf, const float *conf,
                     int W, int H, int C, double *grad, AdjWork &w, hipStream_t s)
{
    check_desc(d, C);
    EArgs a = make_args(d, img, ext, nnf, conf, W, H, C, w, s);
    const size_t plane = (size_t)W * H;
    double *S = w.S.ensure(plane * C), *Rx = w.Rx.ensure(plane * C), *Ry = w.Ry.ensure(plane * C);
    const dim3 g2((unsigned)((W + kBlock - 1) / kBlock), (unsigned)H);
    k_grad_gather<<<g2, kBlock, 0, s>>>(a, S, Rx, Ry);
    NLI_LAUNCH_CHECK();
    const int n = compact_ext(ext, W, H, w, s);
    if (n > 0) {
        k_grad_scatter<<<grid_n((size_t)n), kBlock, 0, s>>>(a, w.pts.p, n, S, Rx, Ry);
        NLI_LAUNCH_CHECK();
    }
    k_grad_combine<<<g2, kBlock, 0, s>>>(S, Rx, Ry, W, H, C, grad);
    NLI_LAUNCH_CHECK();
}

int patch_extract(int P, const double *u, const uint8_t *ext, const int32_t *nnf, int W, int H, int C, int side,
                  double *out, size_t out_capacity, AdjWork &w, hipStream_t s)
{
    check_desc(EnergyDesc{P, 1.0f, 0.0f}, C);
    const int n = compact_ext(ext, W, H, w, s);
    const size_t total = (size_t)n * P * P;
    if (total * C > out_capacity) throw std::invalid_argument("patch buffer too small for the extended domain");
    if (total)
        k_extract<<<grid_n(total), kBlock, 0, s>>>(u, w.pts.p, n, reinterpret_cast<const int2 *>(nnf), P, W, H, C,
                                                    side, out);
    NLI_LAUNCH_CHECK();
    return n;
}

void adjoint_aggregate(int P, const double *patches, const uint8_t *ext, const int32_t *nnf, int W, int H, int C,
                       int side, double *out, AdjWork &w, hipStream_t s)
{
    check_desc(EnergyDesc{P, 1.0f, 0.0f}, C);
    const int n = compact_ext(ext, W, H, w, s);
    const size_t plane = (size_t)W * H;
    const int2 *f = reinterpret_cast<const int2 *>(nnf);
    if (side == 0) {
        int *lidx = w.lidx.ensure(plane);
        fill_i32(lidx, -1, plane, s);
        if (n) k_lidx_fill<<<grid_n((size_t)n), kBlock, 0, s>>>(w.pts.p, n, lidx);
        NLI_LAUNCH_CHECK();
        const dim3 g2((unsigned)((W + kBlock - 1) / kBlock), (unsigned)H);
        k_aggregate_t<<<g2, kBlock, 0, s>>>(patches, lidx, f, P, W, H, C, out);
    } else {
        fill_f64(out, 0.0, plane * C, s);
        const size_t total = (size_t)n * P * P;
        if (total) k_aggregate_s<<<grid_n(total), kBlock, 0, s>>>(patches, w.pts.p, n, f, P, W, H, C, out);
    }
    NLI_LAUNCH_CHECK();
}

}  // namespace nli
