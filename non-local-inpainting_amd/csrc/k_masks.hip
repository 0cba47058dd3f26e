// k_masks.hip -- per-scale mask kernels, order-preserving compaction, the exact
// Euclidean distance transform and the confidence mask (gfx950).
//
// Reference: image_inpainting.cpp:566-600 (extended domain), :359-384 and
// :641-657 (target/source masks, erosion, margin), :511-530 (NNF scaling),
// :536-540 (copy of the upsampled image into the hole), :611-630 (confidence),
// distance_transform.cpp:21-115 (Meijster EDT).  All integer/mask results are
// bit-exact; the EDT reproduces the reference's double-precision f()/sep().
#include "nli_internal.h"

#include <cfloat>
#include <cmath>

namespace nli {
namespace {

constexpr int kBlock = 256;

__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int32_t *p, int32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float *p, float v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// one or two NNF fields set to (-1, -1) and their distances to FLT_MAX (n pixels each)
__global__ void k_fill_fields(int2 *n0, float *d0, int2 *n1, float *d1, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    n0[i] = make_int2(-1, -1);
    d0[i] = FLT_MAX;
    if (n1) {
        n1[i] = make_int2(-1, -1);
        d1[i] = FLT_MAX;
    }
}
__global__ void k_fill_f64(double *p, double v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---- order-preserving compaction: 1024 flags per block, 4 per thread ----
constexpr int kItems = 4;
constexpr int kChunk = kBlock * kItems;

__device__ inline int block_exclusive_scan(int v, int *sh)
{
    // Hillis-Steele inclusive scan of 256 values in LDS, returned exclusive.
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int add = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    int incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ void k_count(const uint8_t *flag, int n, int *counts)
{
    __shared__ int sh[kBlock];
    int base = blockIdx.x * kChunk + threadIdx.x * kItems;
    int c = 0;
    for (int k = 0; k < kItems; k++)
        if (base + k < n && flag[base + k]) c++;
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = sh[0];
}

// single block: exclusive scan of nb counts in place, total to *total
__global__ void k_scan_counts(int *counts, int nb, int *total)
{
    __shared__ int sh[kBlock];
    int carry = 0;
    for (int base = 0; base < nb; base += kBlock) {
        int i = base + threadIdx.x;
        int v = (i < nb) ? counts[i] : 0;
        int ex = block_exclusive_scan(v, sh);
        if (i < nb) counts[i] = carry + ex;
        // last thread holds ex + v of the chunk
        __shared__ int chunk_total;
        if (threadIdx.x == kBlock - 1) chunk_total = ex + v;
        __syncthreads();
        carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_write(const uint8_t *flag, int n, const int *offsets, int *out)
{
    __shared__ int sh[kBlock];
    int base = blockIdx.x * kChunk + threadIdx.x * kItems;
    int c = 0;
    for (int k = 0; k < kItems; k++)
        if (base + k < n && flag[base + k]) c++;
    int pos = offsets[blockIdx.x] + block_exclusive_scan(c, sh);
    for (int k = 0; k < kItems; k++)
        if (base + k < n && flag[base + k]) out[pos++] = base + k;
}

// ---- extended domain (separable dilation of the boundary hole pixels) ----
__device__ inline bool mt(const uint8_t *m, int W, int H, int x, int y)
{
    return (unsigned)x < (unsigned)W && (unsigned)y < (unsigned)H && m[(size_t)W * y + x];
}

// boundary[q] = hole(q) and not all four neighbours in the hole (out of range = not hole)
__global__ void k_boundary_hdilate(const uint8_t *hole, int W, int H, int r, uint8_t *hb)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    uint8_t any = 0;
    int xa = max(x - r, 0), xb = min(x + r, W - 1);
    for (int i = xa; i <= xb && !any; i++) {
        if (!hole[(size_t)W * y + i]) continue;
        bool surrounded = mt(hole, W, H, i + 1, y) && mt(hole, W, H, i - 1, y) &&
                          mt(hole, W, H, i, y + 1) && mt(hole, W, H, i, y - 1);
        if (!surrounded) any = 1;
    }
    hb[(size_t)W * y + x] = any;
}

__global__ void k_vdilate_or(const uint8_t *hole, const uint8_t *hb, int W, int H, int r, uint8_t *ext)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t q = (size_t)W * y + x;
    uint8_t v = hole[q];
    int ya = max(y - r, 0), yb = min(y + r, H - 1);
    for (int j = ya; j <= yb && !v; j++) v = hb[(size_t)W * j + x];
    ext[q] = v ? 1 : 0;
}

__global__ void k_scale_masks(const uint8_t *ext, int W, int H, int margin, int erode, uint8_t *target,
                              uint8_t *source, int *nonempty)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t q = (size_t)W * y + x;
    bool in_margin = x < margin || x >= W - margin || y < margin || y >= H - margin;
    bool src = !ext[q];
    if (erode && src) {
        bool right = (x + 1 < W) && !ext[q + 1];
        bool down = (y + 1 < H) && !ext[q + W];
        src = right && down;
        if (src) atomicOr(nonempty, 1);
    }
    target[q] = (ext[q] && !in_margin) ? 1 : 0;
    source[q] = (src && !in_margin) ? 1 : 0;
}

__device__ inline int round_half_away(float v)
{
    return (v > 0.0) ? (int)floor((double)v + 0.5) : (int)ceil((double)v - 0.5);
}

__global__ void k_scale_nnf(const int32_t *lnnf, const int *pts, int n, int lW, int lH, int uW, int uH,
                            float sx, float sy, int32_t *unnf)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int p = pts[i];
    int x = p % lW, y = p / lW;
    int nx = lnnf[2 * (size_t)p], ny = lnnf[2 * (size_t)p + 1];
    if (nx < 0) return;
    nx = round_half_away((float)nx * sx);
    ny = round_half_away((float)ny * sy);
    int ux = round_half_away((float)x * sx);
    int uy = round_half_away((float)y * sy);
    if (ux < 0 || uy < 0 || ux >= uW || uy >= uH) return;
    unnf[2 * ((size_t)uW * uy + ux)] = nx;
    unnf[2 * ((size_t)uW * uy + ux) + 1] = ny;
}

__global__ void k_copy_into_hole(const float *src, float *dst, const int *pts, int n, int C)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t p = (size_t)pts[i];
    for (int c = 0; c < C; c++) dst[p * C + c] = src[p * C + c];
}

// ---- exact EDT (distance_transform.cpp:21-115) ----
__global__ void k_dt_columns(const uint8_t *mask, int W, int H, float *g)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W) return;
    const float INF = (float)(unsigned)(W + H);
    // the loads of 8 rows are issued before the 8 dependent steps that use them
    constexpr int kR = 8;
    float prev = mask[x] ? INF : 0.0f;
    g[x] = prev;
    for (int y0 = 1; y0 < H; y0 += kR) {
        uint8_t m[kR];
#pragma unroll
        for (int j = 0; j < kR; j++) m[j] = y0 + j < H ? mask[(size_t)W * (y0 + j) + x] : 0;
#pragma unroll
        for (int j = 0; j < kR; j++)
            if (y0 + j < H) {
                const float v = m[j] ? (float)((double)prev + 1.0) : 0.0f;
                g[(size_t)W * (y0 + j) + x] = v;
                prev = v;
            }
    }
    float below = g[(size_t)W * (H - 1) + x];
    for (int y0 = H - 2; y0 >= 0; y0 -= kR) {
        float c[kR];
#pragma unroll
        for (int j = 0; j < kR; j++) c[j] = y0 - j >= 0 ? g[(size_t)W * (y0 - j) + x] : 0.0f;
#pragma unroll
        for (int j = 0; j < kR; j++)
            if (y0 - j >= 0) {
                float cur = c[j];
                if (below < cur) {
                    cur = (float)((double)below + 1.0);
                    g[(size_t)W * (y0 - j) + x] = cur;
                }
                below = cur;
            }
    }
}

__device__ inline float dt_f(int x, int xi, float gi)
{
    double dx = (double)(float)(x - xi);
    double gg = (double)gi;
    return (float)sqrt(dx * dx + gg * gg);
}

__device__ inline int dt_sep(int i, int u, float gi, float gu)
{
    double du = (double)(float)u, di = (double)(float)i;
    double a = du * du - di * di + (double)gu * (double)gu - (double)gi * (double)gi;
    return (int)(a / (double)(2 * (u - i)));
}

// One thread per row.  The envelope stacks s, t (and g at the stacked columns) live in
// LDS when a row's 3W words fit (LDS_ROWS threads per block), else in the global scratch
// st: the scan's dependent stack accesses then cost LDS latency, not memory latency.
template <bool LDS_STACK>
__global__ void k_dt_rows(const float *g, int W, int H, float *out, int *st)
{
    extern __shared__ int dt_lds[];
    int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H) return;
    const float *gr = g + (size_t)W * y;
    int *s = LDS_STACK ? dt_lds + (size_t)3 * W * threadIdx.x : st + (size_t)2 * W * y;
    int *t = s + W;
    float *gs = LDS_STACK ? reinterpret_cast<float *>(t + W) : nullptr;   // g at s[q]
    auto g_at = [&](int q) { return LDS_STACK ? gs[q] : gr[s[q]]; };
    int q = 0;
    s[0] = 0;
    t[0] = 0;
    if (LDS_STACK) gs[0] = gr[0];
    for (int u = 1; u < W; u++) {
        const float gu = gr[u];
        while (q >= 0 && dt_f(t[q], s[q], g_at(q)) > dt_f(t[q], u, gu)) q--;
        if (q < 0) {
            q = 0;
            s[0] = u;
            if (LDS_STACK) gs[0] = gu;
        } else {
            unsigned w = (unsigned)(1 + dt_sep(s[q], u, g_at(q), gu));
            if (w < (unsigned)W) {
                q++;
                s[q] = u;
                t[q] = (int)w;
                if (LDS_STACK) gs[q] = gu;
            }
        }
    }
    for (int u = W - 1; u >= 0; u--) {
        out[(size_t)W * y + u] = dt_f(u, s[q], g_at(q));
        if (u == t[q]) q--;
    }
}

// (1 - a) * exp(-d / tau) + a inside the mask, 1 elsewhere.  expf is evaluated as the
// float rounding of the double exponential (within one ulp of the host expf).
__global__ void k_confidence(const uint8_t *mask, const float *dt, size_t n, float decay, float asym,
                             float *out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = 1.0f;
    if (mask[i]) {
        if (decay > 0) {
            float t = -dt[i] / decay;
            float e = (float)exp((double)t);
            v = (1 - asym) * e + asym;
        } else {
            v = asym;
        }
    }
    out[i] = v;
}

}  // namespace

void fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s)
{
    if (!n) return;
    k_fill_u8<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
    NLI_LAUNCH_CHECK();
}
void fill_i32(int32_t *p, int32_t v, size_t n, hipStream_t s)
{
    if (!n) return;
    k_fill_i32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
    NLI_LAUNCH_CHECK();
}
void fill_fields(int32_t *n0, float *d0, int32_t *n1, float *d1, size_t n, hipStream_t s)
{
    if (!n) return;
    k_fill_fields<<<grid_for(n), kBlock, 0, s>>>((int2 *)n0, d0, (int2 *)n1, d1, n);
    NLI_LAUNCH_CHECK();
}
void fill_f32(float *p, float v, size_t n, hipStream_t s)
{
    if (!n) return;
    k_fill_f32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
    NLI_LAUNCH_CHECK();
}
void fill_f64(double *p, double v, size_t n, hipStream_t s)
{
    if (!n) return;
    k_fill_f64<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
    NLI_LAUNCH_CHECK();
}

int compact_indices(const uint8_t *flag, int n, int *out, int *counts, int *d_total, hipStream_t s)
{
    if (n <= 0) return 0;
    int nb = (n + kChunk - 1) / kChunk;
    k_count<<<nb, kBlock, 0, s>>>(flag, n, counts);
    NLI_LAUNCH_CHECK();
    k_scan_counts<<<1, kBlock, 0, s>>>(counts, nb, d_total);
    NLI_LAUNCH_CHECK();
    k_write<<<nb, kBlock, 0, s>>>(flag, n, counts, out);
    NLI_LAUNCH_CHECK();
    int total = 0;
    NLI_HIP(hipMemcpyAsync(&total, d_total, sizeof(int), hipMemcpyDeviceToHost, s));
    NLI_HIP(hipStreamSynchronize(s));
    return total;
}

void extended_domain(const uint8_t *hole, int W, int H, int P, uint8_t *ext, uint8_t *scratch, hipStream_t s)
{
    int r = P / 2;
    dim3 grid((W + kBlock - 1) / kBlock, H);
    k_boundary_hdilate<<<grid, kBlock, 0, s>>>(hole, W, H, r, scratch);
    NLI_LAUNCH_CHECK();
    k_vdilate_or<<<grid, kBlock, 0, s>>>(hole, scratch, W, H, r, ext);
    NLI_LAUNCH_CHECK();
}

void scale_masks(const uint8_t *ext, int W, int H, int margin, int erode, uint8_t *target, uint8_t *source,
                 int *d_nonempty, hipStream_t s)
{
    NLI_HIP(hipMemsetAsync(d_nonempty, 0, sizeof(int), s));
    dim3 grid((W + kBlock - 1) / kBlock, H);
    k_scale_masks<<<grid, kBlock, 0, s>>>(ext, W, H, margin, erode, target, source, d_nonempty);
    NLI_LAUNCH_CHECK();
}

void scale_nnf(const int32_t *lnnf, const int *pts, int n, int lW, int lH, int uW, int uH, int32_t *unnf,
               hipStream_t s)
{
    fill_i32(unnf, -1, (size_t)uW * uH * 2, s);
    if (!n) return;
    float sx = (float)(unsigned)uW / (float)(unsigned)lW;
    float sy = (float)(unsigned)uH / (float)(unsigned)lH;
    k_scale_nnf<<<grid_for(n), kBlock, 0, s>>>(lnnf, pts, n, lW, lH, uW, uH, sx, sy, unnf);
    NLI_LAUNCH_CHECK();
}

void copy_into_hole(const float *src, float *dst, const int *pts, int n, int C, hipStream_t s)
{
    if (!n) return;
    k_copy_into_hole<<<grid_for(n), kBlock, 0, s>>>(src, dst, pts, n, C);
    NLI_LAUNCH_CHECK();
}

void distance_transform(const uint8_t *mask, int W, int H, float *out, float *g, int *st, hipStream_t s)
{
    k_dt_columns<<<grid_for(W), kBlock, 0, s>>>(mask, W, H, g);
    NLI_LAUNCH_CHECK();
    const size_t row_bytes = (size_t)3 * W * sizeof(int);
    const int per_block = (int)std::min<size_t>(64, (64 << 10) / row_bytes);
    if (per_block >= 1) {
        k_dt_rows<true><<<(H + per_block - 1) / per_block, per_block, per_block * row_bytes, s>>>(g, W, H, out, st);
    } else {
        k_dt_rows<false><<<(H + 63) / 64, 64, 0, s>>>(g, W, H, out, st);
    }
    NLI_LAUNCH_CHECK();
}

void confidence(const uint8_t *mask, const float *dt, int W, int H, float decay, float asym, float *out,
                hipStream_t s)
{
    size_t n = (size_t)W * H;
    k_confidence<<<grid_for(n), kBlock, 0, s>>>(mask, dt, n, decay, asym, out);
    NLI_LAUNCH_CHECK();
}

}  // namespace nli
