// k_patchmatch.hip -- gradient, patch distances and PatchMatch NNF search (gfx950).
//
// Reference: gradient.cpp:20-76; l2_norm_patch_distance.cpp:22-104,
// l1_norm_patch_distance.cpp:22-105, l2_combined_patch_distance.cpp:39-148,
// l2_rgbd_patch_distance.cpp:41-117 (distances); patch_match.cpp:57-302 / 304-512
// (PatchMatch: initialisation, propagation, random search).
//
// Distances keep the reference's float/double types and summation order
// (-ffp-contract=off), so a distance at a fixed NNF is bit-exact.
//
// Search schedule.  The reference's propagation is a sequential scan (each target
// reads its left/top neighbour updated in the same sweep), which has no parallel
// equivalent.  Each sweep here is one launch in which every target, in parallel,
// reads the previous sweep's field (Jacobi) and tries the shifted entries of its four
// neighbours at distance `step` (jump flooding: step = s0, s0/2, ..., 1, repeated),
// then runs the reference's random search unchanged: windows max(W,H)/2^k around the
// post-propagation best, up to `shots` uniform draws per window, the first draw inside
// the source evaluated.  Candidates are accepted only when strictly better, as in the
// reference.  A candidate's distance evaluation stops early once its partial sum proves
// it cannot be accepted; the accept/reject decision is then identical to the full
// evaluation (all terms are non-negative and rounding is monotone).
//
// Work layout.  16x4 target tiles, one 256-thread block each (4 lanes per target in a
// sweep, 1 in the initialisation; only tiles holding a target are launched).  The block stages the target
// patch data of its tile plus the patch halo in LDS as planar float planes (colour
// channels, then the 2C gradient planes); source patches are gathered from global
// memory (L2 / MALL resident), row by row.
#include "nli_internal.h"

#include <atomic>
#include <map>
#include <mutex>

#include <cfloat>
#include <climits>
#include <cmath>
#include <string>
#include <vector>

namespace nli {
namespace {

constexpr int kBlock = 256;
constexpr int TX = 16;     // target tile: 16 x 4 targets
constexpr int TY = 4;
constexpr int kLanes = 4;  // sweep lanes per target (kBlock = TX * TY * kLanes)
constexpr int kXCDs = 8;
static_assert(kBlock == TX * TY * kLanes, "one sweep block covers one tile");

// ---- counter-based RNG: a 32-bit integer hash (xor-shift / multiply rounds with the
// constants of Wellons' "lowbias32", a published avalanche-tested hash) over a Weyl
// counter; the counter starts at a hash of (seed, stream, index), so every draw is a
// pure function of its key and its position in the sequence.  Two 32-bit multiplies per
// draw: the random-search generator draws for every target and window of a sweep.
__device__ inline uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}
struct Rng {
    uint32_t s;
    __device__ uint32_t next()
    {
        s += 0x9E3779B9u;
        return hash32(s);
    }
    // uniform in [0, n) by multiply-high
    __device__ int below(int n) { return (int)(((uint64_t)next() * (uint32_t)n) >> 32); }
};
__device__ inline Rng make_rng(uint64_t seed, uint32_t stream, uint32_t idx)
{
    const uint32_t key = hash32((uint32_t)seed ^ hash32((uint32_t)(seed >> 32) + stream));
    return Rng{hash32(idx ^ key)};
}

struct PMArgs {
    const float *src, *sgrad;
    const uint8_t *smask;
    int sW, sH;
    const float *tgt, *tgrad;
    const uint8_t *tmask;
    int tW, tH;
    int P;
    const float *w;
    float lambda, lambda_d;
    const int2 *nnf_in;
    const float *dist_in;
    int2 *nnf_out;
    float *dist_out;
    const int *tiles;
    int ntiles_x, ntiles;
    int maxw, shots, step;
    uint64_t seed;
    uint32_t stream;
    const int2 *init;   // initial field (k_pm_init), may be null
    int row_scan;       // pool sweep: carry improvements along each tile row (see k_pm_sweep_pool)
    // batched launches (one launch for several images): sweep k reads field k & 1 of the
    // pair and writes the other, at jump-flood step s0 >> (k mod (levels + 1)); an image
    // whose *active is 0 is skipped (its stopping rule has ended its iterations)
    int2 *nnf_pp[2];
    float *dist_pp[2];
    int s0, levels;
    const int *active;
};

// The image of a block in a batched launch: start[i] is the first block of image i
// (start[n] = the total), so the image is the last i with start[i] <= b.
__device__ inline int batch_image(const int *__restrict__ start, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Centres whose patch leaves the image are never targets or sources here (the
// reference would read outside the image for them; its masks always exclude them).
__device__ inline bool is_source(const PMArgs &a, int x, int y)
{
    int r = a.P / 2;
    return x >= r && y >= r && x < a.sW - r && y < a.sH - r && a.smask[(size_t)a.sW * y + x];
}
__device__ inline bool is_target(const PMArgs &a, int x, int y)
{
    int r = a.P / 2;
    return x >= r && y >= r && x < a.tW - r && y < a.tH - r && a.tmask[(size_t)a.tW * y + x];
}

// L2RGBD with lambda_rgb = 1 and lambda_d = 0 (the RGB-D workload's weights), a
// pool-sweep kernel variant only: the colour-gradient and depth-value terms then enter
// the distance multiplied by exactly 0, so the distance is the colour-value sum plus the
// depth-gradient sum, accumulated as in the general kernel (l2_rgbd_patch_distance.cpp:
// 114-116 with those weights) -- the same bits with about half the arithmetic.
constexpr int DIST_L2RGBD_CD = 4;
static_assert(DIST_L2RGBD_CD != DIST_L2 && DIST_L2RGBD_CD != DIST_L1 && DIST_L2RGBD_CD != DIST_L2COMBINED &&
                  DIST_L2RGBD_CD != DIST_L2RGBD,
              "kernel variant id outside the distance kinds");

template <int KIND>
struct Feat {
    static constexpr bool grad = (KIND == DIST_L2COMBINED || KIND == DIST_L2RGBD || KIND == DIST_L2RGBD_CD);
};

// target patch accessors: f < C colour channel f, f >= C gradient component f - C
struct TgtLds {
    const float *lds;
    int plane, LW, ox, oy;
    __device__ float v(int f, int dx, int dy) const { return lds[f * plane + (oy + dy) * LW + ox + dx]; }
};
template <int C>
struct TgtGlobal {
    const float *img, *grad;
    int W, ox, oy;
    __device__ float v(int f, int dx, int dy) const
    {
        size_t p = (size_t)(oy + dy) * W + (ox + dx);
        return f < C ? img[p * C + f] : grad[p * 2 * C + (f - C)];
    }
};

// One row (dy) of a gradient distance: cur = image row sy - r + dy, nxt = the row below
// (the same row again at the bottom image border, so its vertical differences are the
// exact +0 of gradient.cpp:49-52); column P of both is column x0 + P, or column
// x0 + P - 1 at the right border (+0 horizontal differences, :34-37).
template <int KIND, int C, int P, class T>
__device__ inline void fly_row(const float (&cur)[P + 1][C], const float (&nxt)[P + 1][C],
                               const float *__restrict__ w, const T &t, int dy, double &dist, double &gdist,
                               double &dist_d, double &gdist_d)
{
#pragma unroll
    for (int dx = 0; dx < P; dx++) {
        const float wt = w[dy * P + dx];
        float g[2 * C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            g[2 * c] = cur[dx + 1][c] - cur[dx][c];
            g[2 * c + 1] = nxt[dx][c] - cur[dx][c];
        }
        if constexpr (KIND == DIST_L2COMBINED) {
            float norm = 0.0f, gnorm = 0.0f;
#pragma unroll
            for (int c = 0; c < C; c++) {
                float d = cur[dx][c] - t.v(c, dx, dy);
                norm = (c == 0) ? d * d : norm + d * d;
            }
#pragma unroll
            for (int k = 0; k < 2 * C; k++) {
                float d = g[k] - t.v(C + k, dx, dy);
                gnorm = (k == 0) ? d * d : gnorm + d * d;
            }
            dist += (double)(wt * norm);
            gdist += (double)(wt * gnorm);
        } else if constexpr (KIND == DIST_L2RGBD_CD) {   // C == 4: colour values, depth gradient
            double norm = 0.0, gnorm_d = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float d = cur[dx][c] - t.v(c, dx, dy);
                norm += (double)(d * d);
            }
            {
                float gx = g[6] - t.v(C + 6, dx, dy);
                float gy = g[7] - t.v(C + 7, dx, dy);
                gnorm_d += (double)(gx * gx + gy * gy);
            }
            dist += (double)wt * norm;
            gdist_d += (double)wt * gnorm_d;
        } else {   // DIST_L2RGBD, C == 4
            double norm = 0.0, norm_d = 0.0, gnorm = 0.0, gnorm_d = 0.0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                float d = cur[dx][c] - t.v(c, dx, dy);
                float gx = g[2 * c] - t.v(C + 2 * c, dx, dy);
                float gy = g[2 * c + 1] - t.v(C + 2 * c + 1, dx, dy);
                if (c == 3) {
                    norm_d += (double)(d * d);
                    gnorm_d += (double)(gx * gx + gy * gy);
                } else {
                    norm += (double)(d * d);
                    gnorm += (double)(gx * gx + gy * gy);
                }
            }
            dist += (double)wt * norm;
            dist_d += (double)wt * norm_d;
            gdist += (double)wt * gnorm;
            gdist_d += (double)wt * gnorm_d;
        }
    }
}

template <int C, int P>
__device__ inline void load_row(float (&row)[P + 1][C], const float *__restrict__ src, int last)
{
#pragma unroll
    for (int dx = 0; dx < P; dx++)
#pragma unroll
        for (int c = 0; c < C; c++) row[dx][c] = src[dx * C + c];
#pragma unroll
    for (int c = 0; c < C; c++) row[P][c] = src[last * C + c];
}

// Gradient distances (L2Combined with C = 1 or 3, L2RGBD) for a compile-time patch
// side: the source gradient is formed on the fly from the image rows the colour term
// loads anyway (forward differences, the exact float subtractions of
// gradient.cpp:32-73), so a source patch costs (P+1)^2 C floats of gathers instead of
// 3 P^2 C.  Arithmetic and summation order are those of patch_distance below, so the
// results are identical.
template <int KIND, int C, int PS, class T>
__device__ float patch_distance_fly(const PMArgs &a, const float *__restrict__ w, int sx, int sy, const T &t,
                                    float bd)
{
    constexpr int P = PS, r = PS / 2;
    constexpr unsigned PP = (unsigned)(P * P);
    const double bound = (double)bd * (double)PP * (1.0 + 1e-6);
    const double lam = (double)a.lambda, lam1 = (double)(1 - a.lambda);
    const double lamd = (double)a.lambda_d, lamd1 = (double)(1 - a.lambda_d);
    double dist = 0.0, gdist = 0.0, dist_d = 0.0, gdist_d = 0.0;
    const int x0 = sx - r;
    const int last = (sx + r + 1 < a.sW) ? P : P - 1;   // right border: repeat the last column
    const size_t stride = (size_t)a.sW * C;
    const float *row = a.src + ((size_t)(sy - r) * a.sW + x0) * C;
    auto partial = [&]() {
        double v = lam * dist + lam1 * gdist;
        if constexpr (KIND == DIST_L2RGBD) v = v + lamd * dist_d + lamd1 * gdist_d;
        return v;
    };
    float A[P + 1][C], B[P + 1][C];
    load_row<C, P>(A, row, last);
    for (int dy = 0; dy < P; dy += 2) {
        // rows dy (in A) and dy + 1 (into B)
        const float *nrow = (sy - r + dy + 1 < a.sH) ? row + stride : row;
        load_row<C, P>(B, nrow, last);
        fly_row<KIND, C, P>(A, B, w, t, dy, dist, gdist, dist_d, gdist_d);
        if (partial() > bound) return INFINITY;
        if (dy + 1 >= P) break;
        row = nrow;
        const float *nrow2 = (sy - r + dy + 2 < a.sH) ? row + stride : row;
        load_row<C, P>(A, nrow2, last);
        fly_row<KIND, C, P>(B, A, w, t, dy + 1, dist, gdist, dist_d, gdist_d);
        if (partial() > bound) return INFINITY;
        row = nrow2;
    }
    return (float)(partial() / (double)PP);
}

// Patch distance between the source patch centred at (sx, sy) and the target patch of
// `t`.  Returns +inf as soon as the partial sum proves the result is >= bd.
template <int KIND, int C, int PS, class T>
__device__ float patch_distance(const PMArgs &a, const float *__restrict__ w, int sx, int sy, const T &t, float bd)
{
    if constexpr (PS > 0 && (KIND == DIST_L2RGBD || (KIND == DIST_L2COMBINED && (C == 1 || C == 3))))
        return patch_distance_fly<KIND, C, PS>(a, w, sx, sy, t, bd);
    const int P = PS ? PS : a.P, r = P / 2;
    const unsigned PP = (unsigned)(P * P);
    if constexpr (KIND == DIST_L2 || KIND == DIST_L1) {
        const float bound = bd * (float)PP * (1.0f + 0x1p-20f);
        float acc = 0.0f;
        for (int dy = 0; dy < P; dy++) {
            const float *srow = a.src + ((size_t)(sy - r + dy) * a.sW + (sx - r)) * C;
#pragma unroll
            for (int dx = 0; dx < (PS ? PS : P); dx++) {
                float norm = 0.0f;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    float d = srow[dx * C + c] - t.v(c, dx, dy);
                    float term = (KIND == DIST_L1) ? fabsf(d) : d * d;
                    norm = (c == 0) ? term : norm + term;
                }
                acc += w[dy * P + dx] * norm;
            }
            if (acc > bound) return INFINITY;
        }
        return acc / (float)PP;
    } else {
        const double bound = (double)bd * (double)PP * (1.0 + 1e-6);
        const double lam = (double)a.lambda, lam1 = (double)(1 - a.lambda);
        const double lamd = (double)a.lambda_d, lamd1 = (double)(1 - a.lambda_d);
        double dist = 0.0, gdist = 0.0, dist_d = 0.0, gdist_d = 0.0;
        for (int dy = 0; dy < P; dy++) {
            const size_t rowp = (size_t)(sy - r + dy) * a.sW + (sx - r);
            const float *srow = a.src + rowp * C;
            const float *grow = a.sgrad + rowp * 2 * C;
#pragma unroll
            for (int dx = 0; dx < (PS ? PS : P); dx++) {
                const float wt = w[dy * P + dx];
                if constexpr (KIND == DIST_L2COMBINED && (C == 1 || C == 3)) {
                    float norm = 0.0f, gnorm = 0.0f;
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        float d = srow[dx * C + c] - t.v(c, dx, dy);
                        norm = (c == 0) ? d * d : norm + d * d;
                    }
#pragma unroll
                    for (int k = 0; k < 2 * C; k++) {
                        float d = grow[dx * 2 * C + k] - t.v(C + k, dx, dy);
                        gnorm = (k == 0) ? d * d : gnorm + d * d;
                    }
                    dist += (double)(wt * norm);
                    gdist += (double)(wt * gnorm);
                } else if constexpr (KIND == DIST_L2COMBINED) {
                    double norm = 0.0, gnorm = 0.0;
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        float d = srow[dx * C + c] - t.v(c, dx, dy);
                        norm += (double)(d * d);
                        float gx = grow[dx * 2 * C + 2 * c] - t.v(C + 2 * c, dx, dy);
                        float gy = grow[dx * 2 * C + 2 * c + 1] - t.v(C + 2 * c + 1, dx, dy);
                        gnorm += (double)(gx * gx + gy * gy);
                    }
                    dist += (double)wt * norm;
                    gdist += (double)wt * gnorm;
                } else {   // DIST_L2RGBD, C == 4
                    double norm = 0.0, norm_d = 0.0, gnorm = 0.0, gnorm_d = 0.0;
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        float d = srow[dx * C + c] - t.v(c, dx, dy);
                        float gx = grow[dx * 2 * C + 2 * c] - t.v(C + 2 * c, dx, dy);
                        float gy = grow[dx * 2 * C + 2 * c + 1] - t.v(C + 2 * c + 1, dx, dy);
                        if (c == 3) {
                            norm_d += (double)(d * d);
                            gnorm_d += (double)(gx * gx + gy * gy);
                        } else {
                            norm += (double)(d * d);
                            gnorm += (double)(gx * gx + gy * gy);
                        }
                    }
                    dist += (double)wt * norm;
                    dist_d += (double)wt * norm_d;
                    gdist += (double)wt * gnorm;
                    gdist_d += (double)wt * gnorm_d;
                }
            }
            double partial = lam * dist + lam1 * gdist;
            if constexpr (KIND == DIST_L2RGBD) partial = partial + lamd * dist_d + lamd1 * gdist_d;
            if (partial > bound) return INFINITY;
        }
        double total = lam * dist + lam1 * gdist;
        if constexpr (KIND == DIST_L2RGBD) total = total + lamd * dist_d + lamd1 * gdist_d;
        return (float)(total / (double)PP);
    }
}

__global__ void k_gradient(const float *__restrict__ img, int W, int H, int C, float *__restrict__ grad)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    size_t index = (size_t)C * ((size_t)W * y + x);
    for (int ch = 0; ch < C; ch++) {
        float v = img[index + ch];
        grad[(index + ch) * 2] = (x == W - 1) ? 0.0f : img[index + C + ch] - v;
        grad[(index + ch) * 2 + 1] = (y == H - 1) ? 0.0f : img[index + (size_t)C * W + ch] - v;
    }
}

// flags[t] = tile t holds a target; *count += targets (the mask, without the border rule)
__global__ void k_tile_flags(const uint8_t *tmask, int W, int H, int ntx, int nty, uint8_t *flags, int *count)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntx * nty) return;
    int x0 = (t % ntx) * TX, y0 = (t / ntx) * TY;
    int n = 0;
    for (int y = y0; y < min(y0 + TY, H); y++)
        for (int x = x0; x < min(x0 + TX, W); x++) n += tmask[(size_t)W * y + x] ? 1 : 0;
    flags[t] = n > 0;
    if (n) atomicAdd(count, n);
}

// Initialisation (patch_match.cpp:94-141): the given entry, or uniform draws until one
// lands in the source (at most `shots`), then its distance.
// One thread per target of a 16x4 tile; the tile's target features (with the patch
// halo) are staged in LDS first, as in the sweeps.
template <int KIND, int C, int PS>
__device__ __attribute__((always_inline)) inline void pm_init_body(const PMArgs &a, int tile_slot)
{
    extern __shared__ float init_lds[];
    constexpr int F = Feat<KIND>::grad ? 3 * C : C;
    const int P = PS ? PS : a.P;
    const int hr = P / 2;
    const int LW = TX + 2 * hr, LH = TY + 2 * hr;
    const int plane = LW * LH;
    float *wl = init_lds + F * plane;
    int tile = a.tiles[tile_slot];
    const int x0 = (tile % a.ntiles_x) * TX, y0 = (tile / a.ntiles_x) * TY;
    for (int i = threadIdx.x; i < P * P; i += blockDim.x) wl[i] = a.w[i];
    for (int i = threadIdx.x; i < plane; i += blockDim.x) {
        const int gx = x0 - hr + i % LW, gy = y0 - hr + i / LW;
        const bool in = (unsigned)gx < (unsigned)a.tW && (unsigned)gy < (unsigned)a.tH;
        const size_t q = in ? (size_t)gy * a.tW + gx : 0;
#pragma unroll
        for (int c = 0; c < C; c++) init_lds[c * plane + i] = in ? a.tgt[q * C + c] : 0.0f;
        if constexpr (F > C) {
#pragma unroll
            for (int k = 0; k < 2 * C; k++) init_lds[(C + k) * plane + i] = in ? a.tgrad[q * 2 * C + k] : 0.0f;
        }
    }
    __syncthreads();
    const int lx = (int)threadIdx.x % TX, ly = (int)threadIdx.x / TX;
    int x = x0 + lx;
    int y = y0 + ly;
    if (x >= a.tW || y >= a.tH || !is_target(a, x, y)) return;
    size_t p = (size_t)a.tW * y + x;
    Rng rng = make_rng(a.seed, a.stream, (uint32_t)p);
    int cx, cy, tries = 0;
    if (a.init) {
        int2 c = a.init[p];
        cx = c.x;
        cy = c.y;
        while (!is_source(a, cx, cy) && tries < a.shots) {
            cx = rng.below(a.sW);
            cy = rng.below(a.sH);
            tries++;
        }
    } else {
        cx = -1;
        cy = -1;
        while (!is_source(a, cx, cy) && tries < a.shots) {
            cx = rng.below(a.sW);
            cy = rng.below(a.sH);
            tries++;
        }
    }
    if (is_source(a, cx, cy)) {
        const TgtLds t{init_lds, plane, LW, lx, ly};
        a.nnf_out[p] = make_int2(cx, cy);
        a.dist_out[p] = patch_distance<KIND, C, PS>(a, wl, cx, cy, t, FLT_MAX);
    }
}

template <int KIND, int C, int PS>
__global__ __launch_bounds__(kBlock) void k_pm_init(PMArgs a)
{
    pm_init_body<KIND, C, PS>(a, blockIdx.x);
}

// batched: one block per target tile of every image; writes field 0 of each image's pair
template <int KIND, int C, int PS>
__global__ __launch_bounds__(kBlock) void k_pm_init_batch(const PMArgs *__restrict__ args,
                                                          const int *__restrict__ start, int nimg)
{
    const int b = (int)blockIdx.x;
    if (b >= start[nimg]) return;
    const int img = batch_image(start, nimg, b);
    PMArgs a = args[img];
    if (a.active && *a.active == 0) return;
    a.nnf_out = a.nnf_pp[0];
    a.dist_out = a.dist_pp[0];
    pm_init_body<KIND, C, PS>(a, b - start[img]);
}

// (distance, candidate, order) minimum over the 4 lanes of a target group: smallest
// distance, ties to the earliest candidate -- the outcome of evaluating the candidates
// one after the other with a strict "<".
__device__ inline void group_min(float &d, int2 &c, int &order)
{
#pragma unroll
    for (int m = 1; m <= 2; m <<= 1) {
        float od = __shfl_xor(d, m, kLanes);
        int ox = __shfl_xor(c.x, m, kLanes);
        int oy = __shfl_xor(c.y, m, kLanes);
        int oo = __shfl_xor(order, m, kLanes);
        if (od < d || (od == d && oo < order)) {
            d = od;
            c = make_int2(ox, oy);
            order = oo;
        }
    }
}

// One sweep: jump-flood propagation at distance `step` + random search, 4 lanes per
// target.  Phase 1: lane k evaluates the shifted entry of neighbour k (left, right,
// up, down at distance `step`).  Phase 2: the search windows, centred on the phase-1
// best, are dealt round-robin to the 4 lanes.  A lane stops a distance evaluation
// early against its own running best, which is never below the sequential running
// best at that candidate, so every accept/reject decision is the sequential one.
template <int KIND, int C, int PS>
__global__ __launch_bounds__(kBlock) void k_pm_sweep(PMArgs a)
{
    extern __shared__ float lds[];
    constexpr int F = Feat<KIND>::grad ? 3 * C : C;
    const int P = PS ? PS : a.P;
    const int r = P / 2;
    const int LW = TX + 2 * r, LH = TY + 2 * r;
    const int plane = LW * LH;
    float *wl = lds + F * plane;
    // blocks are dealt to the 8 XCDs round-robin: give each XCD a contiguous band of
    // tiles so that neighbouring targets' candidates share its L2
    const int per = (int)(gridDim.x / kXCDs);
    const int slot = (int)(blockIdx.x % kXCDs) * per + (int)(blockIdx.x / kXCDs);
    if (slot >= a.ntiles) return;
    const int tile = a.tiles[slot];
    const int x0 = (tile % a.ntiles_x) * TX, y0 = (tile / a.ntiles_x) * TY;
    for (int i = threadIdx.x; i < P * P; i += blockDim.x) wl[i] = a.w[i];
    for (int i = threadIdx.x; i < plane; i += blockDim.x) {
        int gx = x0 - r + i % LW, gy = y0 - r + i / LW;
        bool in = (unsigned)gx < (unsigned)a.tW && (unsigned)gy < (unsigned)a.tH;
        size_t q = in ? (size_t)gy * a.tW + gx : 0;
#pragma unroll
        for (int c = 0; c < C; c++) lds[c * plane + i] = in ? a.tgt[q * C + c] : 0.0f;
        if constexpr (F > C) {
#pragma unroll
            for (int k = 0; k < 2 * C; k++) lds[(C + k) * plane + i] = in ? a.tgrad[q * 2 * C + k] : 0.0f;
        }
    }
    __syncthreads();
    const int lane = (int)threadIdx.x & (kLanes - 1);
    const int tl = (int)threadIdx.x / kLanes;
    const int lx = tl % TX, ly = tl / TX;
    const int x = x0 + lx, y = y0 + ly;
    // the 4 lanes of a group share the target, so groups leave together
    if (x >= a.tW || y >= a.tH || !is_target(a, x, y)) return;
    const size_t p = (size_t)a.tW * y + x;
    const TgtLds t{lds, plane, LW, lx, ly};
    int2 best = a.nnf_in[p];
    float bd = a.dist_in[p];

    // phase 1: propagation.  A candidate equal to the current best is skipped: its
    // distance is exactly bd, so the reference's strict comparison never takes it.
    {
        const int s = a.step;
        const int ox = lane == 0 ? -s : (lane == 1 ? s : 0);
        const int oy = lane == 2 ? -s : (lane == 3 ? s : 0);
        float d = INFINITY;
        int2 c = best;
        int order = INT_MAX;
        const int qx = x + ox, qy = y + oy;
        if (is_target(a, qx, qy)) {
            int2 e = a.nnf_in[(size_t)a.tW * qy + qx];
            e.x -= ox;
            e.y -= oy;
            if ((e.x != best.x || e.y != best.y) && is_source(a, e.x, e.y)) {
                d = patch_distance<KIND, C, PS>(a, wl, e.x, e.y, t, bd);
                c = e;
                order = lane;
            }
        }
        group_min(d, c, order);
        if (d < bd) {
            bd = d;
            best = c;
        }
    }

    // phase 2: random search around the phase-1 best, in rounds of one window per lane;
    // after each round the group's best becomes every lane's pruning bound
    {
        Rng rng = make_rng(a.seed, a.stream * kLanes + (uint32_t)lane, (uint32_t)p);
        const int scx = best.x, scy = best.y;
        int R = 0;
        for (int ws = a.maxw; ws >= 1; ws /= 2) R++;
        float gd = INFINITY;   // group best of phase 2 so far
        int2 gc = best;
        int gorder = INT_MAX;
        for (int base = 0; base < R; base += kLanes) {
            const int wi = base + lane;
            float d = INFINITY;
            int2 c = best;
            int order = INT_MAX;
            if (wi < R) {
                const int ws = a.maxw >> wi;
                int x_min = max(scx - ws, 0), y_min = max(scy - ws, 0);
                int x_max = min(scx + ws + 1, a.sW), y_max = min(scy + ws + 1, a.sH);
                for (int k = 0; k < a.shots; k++) {
                    int cx = x_min + rng.below(x_max - x_min);
                    int cy = y_min + rng.below(y_max - y_min);
                    if (is_source(a, cx, cy)) {
                        if (cx == best.x && cy == best.y) break;   // the phase-1 best: never strictly better
                        const float bound = fminf(bd, gd);
                        float dd = patch_distance<KIND, C, PS>(a, wl, cx, cy, t, bound);
                        if (dd < bound) {
                            d = dd;
                            c = make_int2(cx, cy);
                            order = wi;
                        }
                        break;
                    }
                }
            }
            group_min(d, c, order);
            if (d < gd || (d == gd && order < gorder)) {
                gd = d;
                gc = c;
                gorder = order;
            }
        }
        if (gd < bd) {
            bd = gd;
            best = gc;
        }
    }
    if (lane == 0) {
        a.nnf_out[p] = best;
        a.dist_out[p] = bd;
    }
}

// ---- sweep with per-wave candidate pools (compile-time patch) ----
// A block takes one or two target tiles; wave w takes row w of them (16 or 32 targets), or,
// in the split-row arrangement, one tile with two waves per row (8 targets a wave).
// The sweep's candidates -- the 4 jump-flood neighbours' shifted matches, then, around the
// resulting best, the reference's random search (windows max(W,H)/2^k, at most `shots`
// tries, the first in-source draw kept) -- are first written to LDS per wave and then
// evaluated as a pool: a lane whose candidate is pruned by early termination takes the
// next one, so lanes do not wait for the slowest candidate of a round.  The two phases
// share one pool (see sweep_pool_body).  A candidate equal
// to the current best is dropped: its distance is exactly the current one and the
// reference's strict comparison never takes it.  Each target's best (distance, order) is
// an LDS 64-bit atomic minimum; a candidate is pruned only when its partial sum exceeds
// the target's current best, which never drops below the final minimum, and ties go to
// the lower order -- the result equals the sequential strict-improvement acceptance over
// the same candidates (neighbours in order left, right, up, down; windows largest first).
constexpr int kMaxR = 16;                      // random-search windows (max(W, H) < 65536)
constexpr int kWaves = kBlock / 64;
static_assert(kWaves == TY, "wave w covers row w of each tile");
// Tile arrangements (the autotuner chooses among them; fields do not depend on the choice):
//   1: one tile per block of kBlock threads, wave w takes tile row w (TX targets);
//   2: two tiles per block, wave w takes row w of both (2 TX targets);
//   3: one tile per block of 2 kBlock threads, waves 2w and 2w + 1 take the two halves of
//      tile row w (TX / 2 targets each): twice the waves per tile, each with half the
//      work, for scales whose tiles leave most of the GPU idle.  The row scan needs whole
//      tile rows in one wave, so this arrangement runs only with the row scan off.
// In the kernels TPB is the tiles per block and WPR the waves per tile row.
constexpr int kArrangements = 3;

__device__ inline int lane_rank(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// Orders this wave's LDS writes before its later LDS reads (the pool tables are private
// to a wave, so no workgroup barrier is needed).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// candidates are packed as x | y << 16 in 32 unsigned bits (the pool sweep runs only for
// images below 65536 pixels a side)
__device__ inline uint32_t pack_xy(int x, int y) { return (uint32_t)x | ((uint32_t)y << 16); }
__device__ inline int2 unpack_xy(uint32_t v) { return make_int2((int)(v & 0xFFFFu), (int)(v >> 16)); }
__device__ inline uint64_t pool_key(float d, int order)
{
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)order;
}
__device__ inline float key_dist(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
constexpr uint64_t kNoKey = ((uint64_t)0x7F800000u << 32) | 0xFFFFFFFFu;   // (+inf, max)

// the target tiles are kept pixel-interleaved: F features padded to a multiple of 4
// floats per pixel, so one pixel's features are 16-byte LDS reads at immediate offsets
__host__ __device__ constexpr int pool_feat_stride(int F) { return (F + 3) & ~3; }
__host__ __device__ inline int pool_tasks_per_target(int R) { return 4 + R; }   // orders 0-3: propagation, 4 + wi: window wi
__host__ __device__ inline size_t pool_tables_offset(int F, int P, int tpb)
{
    const int r = P / 2;
    const size_t tiles = (size_t)tpb * pool_feat_stride(F) * (TX + 2 * r) * (TY + 2 * r);
    return ((tiles + (size_t)P * P) * sizeof(float) + 15) & ~(size_t)15;
}
// per wave: two keys (u64), best (int2), distance (float) per target; candidate (int)
// and list entry (u16) per task
__host__ __device__ inline size_t pool_wave_bytes(int R, int tpb, int wpr)
{
    const size_t targets = (size_t)tpb * TX / wpr;
    return targets * (2 * sizeof(uint64_t) + sizeof(int2) + sizeof(float)) +
           targets * pool_tasks_per_target(R) * (sizeof(int) + sizeof(uint16_t));
}
template <int FP>
struct TgtLdsI {
    const float *lds;
    int LW, ox, oy;
    __device__ float v(int f, int dx, int dy) const { return lds[((oy + dy) * LW + ox + dx) * FP + f]; }
};

struct WavePool {
    uint64_t *keys;   // [2 kWaveTargets] best (distance, order) of the running pool: tasks of
                      // order < 4 (propagation, row scan) in [t], random search in [kWaveTargets + t]
    int2 *best;       // [kWaveTargets] current best match
    float *bd;        // [kWaveTargets] its distance (-1: not a target)
    uint32_t *cand;   // [kWaveTargets * tasks]: pack_xy(cx, cy) for task order * kWaveTargets + target
    uint16_t *list;   // [kWaveTargets * tasks]: drawn tasks
};

// Evaluates the `nvalid` listed tasks of a wave (task j: target j % kWaveTargets, order
// j / kWaveTargets, candidate cand[j]); records each completed candidate by atomic minimum
// in keys[target] (order < 4) or keys[kWaveTargets + target] (order >= 4), and prunes it
// against that key and the target's current distance.  A task of order >= 4 (random
// search) is dropped as soon as keys[target] lies below the target's distance: then
// propagation has improved the target and its random search runs again around the new
// best (sweep_pool_body), so these results would be discarded.
template <int KIND, int C, int P, int FP, bool LAST_FIRST, int TPB, int WPR>
__device__ void run_pool(const PMArgs &a, const float *__restrict__ wl, const float *tiles, int LW, int tile_floats,
                         int wave, const WavePool &wp, int nvalid)
{
    constexpr int kSeg = TX / WPR;               // targets of one tile row in this wave
    constexpr int kWaveTargets = TPB * kSeg;
    const int wrow = wave / WPR, wcol = (wave % WPR) * kSeg;
    constexpr bool GRAD = Feat<KIND>::grad;
    constexpr int r = P / 2;
    constexpr unsigned PP = (unsigned)(P * P);
    const double lam = (double)a.lambda, lam1 = (double)(1 - a.lambda);
    const double lamd = (double)a.lambda_d, lamd1 = (double)(1 - a.lambda_d);
    const size_t stride = (size_t)a.sW * C;
    bool active = false;
    int t = 0, order = 0, kslot = 0, cy = 0, dy = 0, last = P;
    float bound = 0.0f, bd_t = 0.0f;
    const float *row = a.src;
    double dist = 0.0, gdist = 0.0, dist_d = 0.0, gdist_d = 0.0;   // gradient distances (FP64)
    float facc = 0.0f;                                             // L2 / L1 norms (FP32)
    float A[GRAD ? P + 1 : 1][C];
    auto partial = [&]() {
        double v = lam * dist + lam1 * gdist;
        if constexpr (KIND == DIST_L2RGBD) v = v + lamd * dist_d + lamd1 * gdist_d;
        if constexpr (KIND == DIST_L2RGBD_CD) v = dist + gdist_d;   // the weights 1, 0, 0, 1
        return v;
    };
    auto record = [&](float d) {
        if (d <= bound) atomicMin((unsigned long long *)&wp.keys[kslot], (unsigned long long)pool_key(d, order));
    };
    int next = 0;
    for (;;) {
        const bool need = !active;
        const uint64_t m = __ballot(need);
        if (need) {
            const int k = next + lane_rank(m);
            if (k < nvalid) {
                const int j = wp.list[LAST_FIRST ? nvalid - 1 - k : k];
                t = j % kWaveTargets;
                order = j / kWaveTargets;
                kslot = order < 4 ? t : kWaveTargets + t;
                const int2 cxy = unpack_xy(wp.cand[j]);
                const int cx = cxy.x;
                cy = cxy.y;
                bd_t = wp.bd[t];
                bound = fminf(bd_t, key_dist(wp.keys[kslot]));
                dy = 0;
                row = a.src + ((size_t)(cy - r) * a.sW + (cx - r)) * C;
                if constexpr (GRAD) {
                    dist = gdist = dist_d = gdist_d = 0.0;
                    last = (cx + r + 1 < a.sW) ? P : P - 1;
                    load_row<C, P>(A, row, last);
                } else {
                    facc = 0.0f;
                }
                active = true;
            }
        }
        next += __popcll(m);
        if (__ballot(active) == 0) break;
        if (active) {
            bound = fminf(bound, key_dist(wp.keys[kslot]));
            if (kslot != t && key_dist(wp.keys[t]) < bd_t) bound = -1.0f;   // superseded: the first check drops it
            const TgtLdsI<FP> tg{tiles + (t / kSeg) * tile_floats, LW, wcol + t % kSeg, wrow};
            if constexpr (GRAD) {
                // one patch row per pass: row dy is in A, row dy + 1 goes to B, then B -> A
                float B[P + 1][C];
                const float *nrow = (cy - r + dy + 1 < a.sH) ? row + stride : row;
                load_row<C, P>(B, nrow, last);
                fly_row<KIND, C, P>(A, B, wl, tg, dy, dist, gdist, dist_d, gdist_d);
                if (partial() > (double)bound * (double)PP * (1.0 + 1e-6)) {
                    active = false;
                } else if (dy + 1 >= P) {
                    record((float)(partial() / (double)PP));
                    active = false;
                } else {
#pragma unroll
                    for (int i = 0; i <= P; i++)
#pragma unroll
                        for (int c = 0; c < C; c++) A[i][c] = B[i][c];
                    row = nrow;
                    dy++;
                }
            } else {
                // colour-only distances: float accumulation in row-major order, as
                // l2_norm_patch_distance.cpp / l1_norm_patch_distance.cpp; two patch rows
                // per pass (a row's arithmetic costs about as much as the pool's
                // bookkeeping of a pass), the bound tested after each row
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (!active) break;
#pragma unroll
                    for (int dx = 0; dx < P; dx++) {
                        float norm = 0.0f;
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            float d = row[dx * C + c] - tg.v(c, dx, dy);
                            float term = (KIND == DIST_L1) ? fabsf(d) : d * d;
                            norm = (c == 0) ? term : norm + term;
                        }
                        facc += wl[dy * P + dx] * norm;
                    }
                    if (facc > bound * (float)PP * (1.0f + 0x1p-20f)) {
                        active = false;
                    } else if (dy + 1 >= P) {
                        record(facc / (float)PP);
                        active = false;
                    } else {
                        row += stride;
                        dy++;
                    }
                }
            }
        }
    }
}

template <int KIND, int C, int PS, int TPB, int WPR>
__device__ __attribute__((always_inline)) inline void sweep_pool_body(const PMArgs &a, int slot)
{
    static_assert(WPR == 1 || (WPR == 2 && TPB == 1), "arrangements: TPB tiles of one wave a row, or one tile of two");
    constexpr int kTilesPerBlock = TPB;
    constexpr int NT = kBlock * WPR;             // threads per block
    constexpr int kSeg = TX / WPR;               // targets of one tile row in a wave
    constexpr int kWaveTargets = TPB * kSeg;
    static_assert(PS > 0 && (KIND == DIST_L2RGBD || KIND == DIST_L2RGBD_CD || KIND == DIST_L2 || KIND == DIST_L1 ||
                             (KIND == DIST_L2COMBINED && (C == 1 || C == 3))),
                  "pool sweep: compile-time patch; gradient distances need the on-the-fly source gradient");
    extern __shared__ __align__(16) float plds[];
    constexpr bool GRAD = Feat<KIND>::grad;
    constexpr int F = GRAD ? 3 * C : C;
    constexpr int FP = pool_feat_stride(F);
    constexpr int P = PS, r = PS / 2;
    constexpr int LW = TX + 2 * r, LH = TY + 2 * r;
    constexpr int plane = LW * LH;
    constexpr int tile_floats = plane * FP;
    float *tiles = plds;
    float *wl = plds + kTilesPerBlock * tile_floats;
    int R = 0;
    for (int ws = a.maxw; ws >= 1; ws /= 2) R++;
    const int ntask = pool_tasks_per_target(R);
    const int wave = (int)threadIdx.x / 64;
    const int lw = (int)threadIdx.x & 63;
    const int wrow = wave / WPR, wcol = (wave % WPR) * kSeg;   // this wave's tile row and first column
    WavePool wp;
    {
        char *pb =
            reinterpret_cast<char *>(plds) + pool_tables_offset(F, P, TPB) + wave * pool_wave_bytes(R, TPB, WPR);
        wp.keys = reinterpret_cast<uint64_t *>(pb);
        wp.best = reinterpret_cast<int2 *>(pb + 2 * kWaveTargets * sizeof(uint64_t));
        wp.bd = reinterpret_cast<float *>(pb + kWaveTargets * (2 * sizeof(uint64_t) + sizeof(int2)));
        wp.cand = reinterpret_cast<uint32_t *>(pb + kWaveTargets * (2 * sizeof(uint64_t) + sizeof(int2) + sizeof(float)));
        wp.list = reinterpret_cast<uint16_t *>(wp.cand + kWaveTargets * ntask);
    }

    if (slot * kTilesPerBlock >= a.ntiles) return;   // whole block: uniform
    int tx0[kTilesPerBlock], ty0[kTilesPerBlock];
    for (int ti = 0; ti < kTilesPerBlock; ti++) {
        const int s = slot * kTilesPerBlock + ti;
        const int tile = s < a.ntiles ? a.tiles[s] : -1;
        tx0[ti] = tile >= 0 ? (tile % a.ntiles_x) * TX : -1;
        ty0[ti] = tile >= 0 ? (tile / a.ntiles_x) * TY : -1;
    }
    // targets of this wave: t = kSeg * tile + i, column wcol + i of tile row wrow
    // (returns whether target t's tile exists)
    auto target_coords = [&](int t, int &x, int &y) {
        int bx = tx0[0], by = ty0[0];   // selected, not indexed: the arrays stay in registers
#pragma unroll
        for (int ti = 1; ti < kTilesPerBlock; ti++)
            if (t / kSeg == ti) {
                bx = tx0[ti];
                by = ty0[ti];
            }
        x = bx + wcol + t % kSeg;
        y = by + wrow;
        return bx >= 0;
    };
    // Set-up loads, all issued before any is used (one memory round trip): lanes below
    // kWaveTargets load their target's mask, match and distance; every thread loads the
    // patch weight and its pixels of the tiles with their halo.  Addresses outside the
    // images become element 0, whose value is then discarded.  A target is valid exactly
    // when its distance in the wave's table is >= 0 (-1 otherwise), which the candidate
    // generators below rely on.
    auto tile_origin = [&](int ti, int &bx, int &by) {   // selected, not indexed
        bx = tx0[0];
        by = ty0[0];
#pragma unroll
        for (int u = 1; u < kTilesPerBlock; u++)
            if (ti == u) {
                bx = tx0[u];
                by = ty0[u];
            }
    };
    int tgx, tgy;
    const bool ttile = target_coords(lw % kWaveTargets, tgx, tgy);
    const bool tinb = lw < kWaveTargets && ttile && tgx >= r && tgy >= r && tgx < a.tW - r && tgy < a.tH - r;
    const size_t tp = tinb ? (size_t)a.tW * tgy + tgx : 0;
    const uint8_t tm = a.tmask[tp];
    const int2 tn = a.nnf_in[tp];
    const float td = a.dist_in[tp];
    static_assert(P * P <= NT, "one thread per patch weight");
    const float wv = a.w[(int)threadIdx.x < P * P ? (int)threadIdx.x : 0];
    constexpr int kPix = (kTilesPerBlock * plane + NT - 1) / NT;       // tile pixels per thread
    float pv[kPix][F];
    bool pin[kPix];
#pragma unroll
    for (int ip = 0; ip < kPix; ip++) {
        const int i = (int)threadIdx.x + ip * NT;
        const int ti = i / plane, k = i % plane;
        int bx, by;
        tile_origin(ti, bx, by);
        const int gx = bx - r + k % LW, gy = by - r + k / LW;
        pin[ip] = i < kTilesPerBlock * plane && bx >= 0 && (unsigned)gx < (unsigned)a.tW &&
                  (unsigned)gy < (unsigned)a.tH;
        const size_t q = pin[ip] ? (size_t)gy * a.tW + gx : 0;
#pragma unroll
        for (int c = 0; c < C; c++) pv[ip][c] = a.tgt[q * C + c];
        if constexpr (GRAD) {
#pragma unroll
            for (int k2 = 0; k2 < 2 * C; k2++) pv[ip][C + k2] = a.tgrad[q * 2 * C + k2];
        }
    }
    const bool my_valid = tinb && tm;
    const int2 my_best = my_valid ? tn : make_int2(-1, -1);
    const float my_bd = my_valid ? td : -1.0f;
    if ((int)threadIdx.x < P * P) wl[threadIdx.x] = wv;
#pragma unroll
    for (int ip = 0; ip < kPix; ip++) {
        const int i = (int)threadIdx.x + ip * NT;
        if (i < kTilesPerBlock * plane) {
            float *px = tiles + (size_t)(i / plane) * tile_floats + (size_t)(i % plane) * FP;
#pragma unroll
            for (int c = 0; c < F; c++) px[c] = pin[ip] ? pv[ip][c] : 0.0f;
#pragma unroll
            for (int k2 = F; k2 < FP; k2++) px[k2] = 0.0f;
        }
    }
    __syncthreads();

    if (lw < kWaveTargets) {
        wp.best[lw] = my_best;
        wp.bd[lw] = my_bd;
        wp.keys[lw] = kNoKey;
        wp.keys[kWaveTargets + lw] = kNoKey;
    }
    wave_sync();

    // Random-search candidates of the targets in `which` (bit t: target t), around each
    // target's current best: the reference's windows max(W,H) / 2^wi, at most `shots` draws
    // per window, the first in-source draw kept unless it is the centre.  Task j = (4 + wi)
    // * kWaveTargets + t.  A lane takes the (wi, t) pairs lw + 64 q; the first draws of all
    // its pairs are formed and their source-mask bytes loaded together, and only draws that
    // miss the source continue the pair's sequence (regenerated from its key): the same
    // candidates as drawing each pair to its end in turn.  Appends to the list from nv.
    auto gen_random = [&](uint64_t which, int nv) {
        constexpr int kQ = (kMaxR * kWaveTargets + 63) / 64;
        const int npairs = R * kWaveTargets;
        uint32_t first[kQ];
        uint32_t first_in = 0;   // bit q: the first draw of pair q lies on a source-mask pixel
#pragma unroll
        for (int q = 0; q < kQ; q++) {
            const int jp = q * 64 + lw;
            first[q] = 0xFFFFFFFFu;
            if (jp < npairs && a.shots > 0) {
                const int t = jp % kWaveTargets, wi = jp / kWaveTargets;
                if (((which >> t) & 1ull) && wp.bd[t] >= 0.0f) {
                    int x, y;
                    target_coords(t, x, y);
                    const int2 ctr = wp.best[t];
                    Rng rng = make_rng(a.seed, a.stream * 32u + (uint32_t)wi, (uint32_t)((size_t)a.tW * y + x));
                    const int ws = a.maxw >> wi;
                    const int x_min = max(ctr.x - ws, 0), y_min = max(ctr.y - ws, 0);
                    const int x_max = min(ctr.x + ws + 1, a.sW), y_max = min(ctr.y + ws + 1, a.sH);
                    const int cx = x_min + rng.below(x_max - x_min);
                    const int cy = y_min + rng.below(y_max - y_min);
                    first[q] = pack_xy(cx, cy);
                    if ((unsigned)cx < (unsigned)a.sW && (unsigned)cy < (unsigned)a.sH &&
                        a.smask[(size_t)a.sW * cy + cx])
                        first_in |= 1u << q;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kQ; q++) {
            const int jp = q * 64 + lw;
            bool ok = false;
            if (first[q] != 0xFFFFFFFFu) {
                const int t = jp % kWaveTargets, wi = jp / kWaveTargets;
                const int2 ctr = wp.best[t];
                const int2 c0 = unpack_xy(first[q]);
                int cx = c0.x, cy = c0.y;
                bool found = cx >= r && cy >= r && cx < a.sW - r && cy < a.sH - r && ((first_in >> q) & 1u);
                if (!found) {
                    int x, y;
                    target_coords(t, x, y);
                    Rng rng = make_rng(a.seed, a.stream * 32u + (uint32_t)wi, (uint32_t)((size_t)a.tW * y + x));
                    const int ws = a.maxw >> wi;
                    const int x_min = max(ctr.x - ws, 0), y_min = max(ctr.y - ws, 0);
                    const int x_max = min(ctr.x + ws + 1, a.sW), y_max = min(ctr.y + ws + 1, a.sH);
                    rng.below(x_max - x_min);   // the first draw, already tried
                    rng.below(y_max - y_min);
                    for (int k = 1; k < a.shots && !found; k++) {
                        cx = x_min + rng.below(x_max - x_min);
                        cy = y_min + rng.below(y_max - y_min);
                        found = is_source(a, cx, cy);
                    }
                }
                if (found) {
                    ok = cx != ctr.x || cy != ctr.y;
                    if (ok) wp.cand[4 * kWaveTargets + jp] = pack_xy(cx, cy);
                }
            }
            const uint64_t m = __ballot(ok);   // pairs jp >= npairs are never ok
            if (ok) wp.list[nv + lane_rank(m)] = (uint16_t)(4 * kWaveTargets + jp);
            nv += __popcll(m);
        }
        return nv;
    };

    // Phases 1 and 2 in one pool.  Phase 1 (propagation): the 4 neighbours at the
    // jump-flood step propose their shifted matches, task order k (left, right, up, down).
    // Phase 2 (random search) belongs around the phase-1 best; its candidates are drawn
    // around the current best and evaluated in the same pool, with their own running
    // minimum (keys[kWaveTargets + t]).  Where propagation improves a target, those draws
    // are discarded and the target's random search runs again around its new best.  So
    // every target ends with exactly the candidates and choices of the two phases in turn,
    // while the pool rarely waits for the propagation tasks alone: in the inpainting loop
    // propagation seldom improves a target.
    int nvalid = gen_random(~0ull, 0);
    for (int base = 0; base < 4 * kWaveTargets; base += 64) {
        const int j = base + lw;
        const int t = j % kWaveTargets, k = j / kWaveTargets;
        bool ok = false;
        if (wp.bd[t] >= 0.0f) {
            int x, y;
            target_coords(t, x, y);
            const int st = a.step;
            const int ox = k == 0 ? -st : (k == 1 ? st : 0);
            const int oy = k == 2 ? -st : (k == 3 ? st : 0);
            const int nx = x + ox, ny = y + oy;
            // the neighbour's mask and match are loaded together (is_target's test)
            const bool inb = nx >= r && ny >= r && nx < a.tW - r && ny < a.tH - r;
            const size_t q = inb ? (size_t)a.tW * ny + nx : 0;
            const uint8_t nm = a.tmask[q];
            int2 e = a.nnf_in[q];
            if (inb && nm) {
                e.x -= ox;
                e.y -= oy;
                const int2 cur = wp.best[t];
                if ((e.x != cur.x || e.y != cur.y) && is_source(a, e.x, e.y)) {
                    wp.cand[j] = pack_xy(e.x, e.y);
                    ok = true;
                }
            }
        }
        const uint64_t m = __ballot(ok);
        if (ok) wp.list[nvalid + lane_rank(m)] = (uint16_t)j;
        nvalid += __popcll(m);
    }
    wave_sync();
    // the list is evaluated from its end: propagation first, then the windows smallest first
    run_pool<KIND, C, P, FP, true, TPB, WPR>(a, wl, tiles, LW, tile_floats, wave, wp, nvalid);
    wave_sync();
    bool again = false;
    if (lw < kWaveTargets) {
        const uint64_t kp = wp.keys[lw], kr = wp.keys[kWaveTargets + lw];
        if (wp.bd[lw] >= 0.0f) {
            if (key_dist(kp) < wp.bd[lw]) {   // propagation improved: search again around it
                wp.best[lw] = unpack_xy(wp.cand[(int)(uint32_t)kp * kWaveTargets + lw]);
                wp.bd[lw] = key_dist(kp);
                again = true;
            } else if (key_dist(kr) < wp.bd[lw]) {
                wp.best[lw] = unpack_xy(wp.cand[(int)(uint32_t)kr * kWaveTargets + lw]);
                wp.bd[lw] = key_dist(kr);
            }
        }
        wp.keys[lw] = kNoKey;
        wp.keys[kWaveTargets + lw] = kNoKey;
    }
    const uint64_t redo = __ballot(again);
    wave_sync();
    if (redo) {
        nvalid = gen_random(redo, 0);
        wave_sync();
        run_pool<KIND, C, P, FP, true, TPB, WPR>(a, wl, tiles, LW, tile_floats, wave, wp, nvalid);
        wave_sync();
        if (lw < kWaveTargets) {
            const uint64_t kr = wp.keys[kWaveTargets + lw];
            if (wp.bd[lw] >= 0.0f && key_dist(kr) < wp.bd[lw]) {
                wp.best[lw] = unpack_xy(wp.cand[(int)(uint32_t)kr * kWaveTargets + lw]);
                wp.bd[lw] = key_dist(kr);
            }
            wp.keys[kWaveTargets + lw] = kNoKey;
        }
        wave_sync();
    }

    // phase 3 (row scan): the reference's sequential scan carries an improvement along a
    // whole row within one sweep (patch_match.cpp:197-232: each target first tries the
    // match of its left neighbour, already updated in this sweep; the right one on odd
    // sweeps).  Within each 16-target tile row of this wave, the targets repeatedly try
    // the shifted match of their scan-side neighbour whenever that neighbour improved in
    // the previous round, until no target improves: an improvement travels the whole tile
    // row in one sweep (the rows lie in one wave, so no workgroup barrier is needed; the
    // host never runs the split-row arrangement with the scan on).
    if (WPR == 1 && a.row_scan) {
        const int sh = (a.stream & 1u) ? 1 : -1;   // the neighbour at x + sh proposes its match - sh
        uint64_t changed = ~0ull;                  // round 0: every neighbour proposes
        for (int round = 0; round < TX; round++) {
            bool ok = false;
            if (lw < kWaveTargets) {
                const int col = lw % TX + sh, tn = lw + sh;
                if (col >= 0 && col < TX && ((changed >> tn) & 1ull) && wp.bd[lw] >= 0.0f && wp.bd[tn] >= 0.0f) {
                    const int2 e = make_int2(wp.best[tn].x - sh, wp.best[tn].y);
                    const int2 cur = wp.best[lw];
                    if ((e.x != cur.x || e.y != cur.y) && is_source(a, e.x, e.y)) {
                        wp.cand[lw] = pack_xy(e.x, e.y);   // task lw: target lw, order 0
                        ok = true;
                    }
                }
            }
            const uint64_t m = __ballot(ok);
            if (m == 0) break;
            if (ok) wp.list[lane_rank(m)] = (uint16_t)lw;
            wave_sync();
            run_pool<KIND, C, P, FP, false, TPB, WPR>(a, wl, tiles, LW, tile_floats, wave, wp, __popcll(m));
            wave_sync();
            bool improved = false;
            if (lw < kWaveTargets) {
                const uint64_t k = wp.keys[lw];
                if (ok && key_dist(k) < wp.bd[lw]) {
                    wp.best[lw] = unpack_xy(wp.cand[lw]);
                    wp.bd[lw] = key_dist(k);
                    improved = true;
                }
                wp.keys[lw] = kNoKey;
            }
            changed = __ballot(improved);
            wave_sync();
        }
    }

    if (lw < kWaveTargets && my_valid) {
        int x, y;
        target_coords(lw, x, y);
        const size_t p = (size_t)a.tW * y + x;
        a.nnf_out[p] = wp.best[lw];
        a.dist_out[p] = wp.bd[lw];
    }
}

// Blocks are dealt to the 8 XCDs round-robin: block b is given slot (b mod 8) * per + b / 8,
// so each XCD gets a contiguous band of tiles and neighbouring targets share its L2.
__device__ inline int xcd_slot()
{
    const int per = (int)(gridDim.x / kXCDs);
    return (int)(blockIdx.x % kXCDs) * per + (int)(blockIdx.x / kXCDs);
}

template <int KIND, int C, int PS, int TPB, int WPR>
__global__ __launch_bounds__(kBlock * WPR) void k_pm_sweep_pool(PMArgs a)
{
    sweep_pool_body<KIND, C, PS, TPB, WPR>(a, xcd_slot());
}

// batched sweep k: start[i] is the first slot (block of TPB tiles) of image i; the slots of
// all images are dealt to the XCDs in bands as above
template <int KIND, int C, int PS, int TPB>
__global__ __launch_bounds__(kBlock) void k_pm_sweep_pool_batch(const PMArgs *__restrict__ args,
                                                                const int *__restrict__ start, int nimg, int k)
{
    const int slot = xcd_slot();
    if (slot >= start[nimg]) return;
    const int img = batch_image(start, nimg, slot);
    PMArgs a = args[img];
    if (a.active && *a.active == 0) return;
    // selected, not indexed by k: an indexed member would put the whole copy in scratch
    const bool odd = (k & 1) != 0;
    a.nnf_in = odd ? a.nnf_pp[1] : a.nnf_pp[0];
    a.dist_in = odd ? a.dist_pp[1] : a.dist_pp[0];
    a.nnf_out = odd ? a.nnf_pp[0] : a.nnf_pp[1];
    a.dist_out = odd ? a.dist_pp[0] : a.dist_pp[1];
    a.step = a.s0 >> (k % (a.levels + 1));
    a.stream = (uint32_t)k;
    sweep_pool_body<KIND, C, PS, TPB, 1>(a, slot - start[img]);
}

template <int KIND, int C, int PS>
__global__ __launch_bounds__(kBlock) void k_dist_field(PMArgs a, const int2 *nnf)
{
    __shared__ float wl[15 * 15];
    for (int i = threadIdx.x; i < a.P * a.P; i += blockDim.x) wl[i] = a.w[i];
    __syncthreads();
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= a.tW) return;
    size_t p = (size_t)a.tW * y + x;
    float out = FLT_MAX;
    if (is_target(a, x, y)) {
        int2 c = nnf[p];
        if (c.x >= 0 && (unsigned)c.x < (unsigned)a.sW && (unsigned)c.y < (unsigned)a.sH) {
            int r = a.P / 2;
            if (c.x >= r && c.y >= r && c.x < a.sW - r && c.y < a.sH - r) {
                TgtGlobal<C> t{a.tgt, a.tgrad, a.tW, x - r, y - r};
                out = patch_distance<KIND, C, PS>(a, wl, c.x, c.y, t, FLT_MAX);
            }
        }
    }
    a.dist_out[p] = out;
}

// ---- dispatch over (kind, channels) ----
// Kernel variants: the patch side is a compile-time constant for the common
// (distance, channels, patch) triples and a runtime value otherwise.
#define NLI_DISPATCH_RT(KIND, C, CALL)                                                        \
    switch (KIND) {                                                                           \
    case DIST_L2:                                                                             \
        switch (C) {                                                                          \
        case 1: CALL(DIST_L2, 1, 0); break;                                                   \
        case 2: CALL(DIST_L2, 2, 0); break;                                                   \
        case 3: CALL(DIST_L2, 3, 0); break;                                                   \
        case 4: CALL(DIST_L2, 4, 0); break;                                                   \
        default: throw std::invalid_argument("patch distance: 1 to 4 channels");              \
        }                                                                                     \
        break;                                                                                \
    case DIST_L1:                                                                             \
        switch (C) {                                                                          \
        case 1: CALL(DIST_L1, 1, 0); break;                                                   \
        case 2: CALL(DIST_L1, 2, 0); break;                                                   \
        case 3: CALL(DIST_L1, 3, 0); break;                                                   \
        case 4: CALL(DIST_L1, 4, 0); break;                                                   \
        default: throw std::invalid_argument("patch distance: 1 to 4 channels");              \
        }                                                                                     \
        break;                                                                                \
    case DIST_L2COMBINED:                                                                     \
        switch (C) {                                                                          \
        case 1: CALL(DIST_L2COMBINED, 1, 0); break;                                           \
        case 2: CALL(DIST_L2COMBINED, 2, 0); break;                                           \
        case 3: CALL(DIST_L2COMBINED, 3, 0); break;                                           \
        case 4: CALL(DIST_L2COMBINED, 4, 0); break;                                           \
        default: throw std::invalid_argument("patch distance: 1 to 4 channels");              \
        }                                                                                     \
        break;                                                                                \
    case DIST_L2RGBD:                                                                         \
        if (C != 4) throw std::invalid_argument("L2 RGBD Distance: Channels do not match.");  \
        CALL(DIST_L2RGBD, 4, 0);                                                              \
        break;                                                                                \
    default: throw std::invalid_argument("unknown patch distance kind");                      \
    }

#define NLI_SPECIAL(KIND, C, P, K, CC, PP, CALL)                                              \
    if ((KIND) == (K) && (C) == (CC) && (P) == (PP)) {                                        \
        CALL(K, CC, PP);                                                                      \
    } else

#define NLI_DISPATCH(KIND, C, P, CALL)                                                        \
    NLI_SPECIAL(KIND, C, P, DIST_L2COMBINED, 3, 9, CALL)                                      \
    NLI_SPECIAL(KIND, C, P, DIST_L2COMBINED, 3, 7, CALL)                                      \
    NLI_SPECIAL(KIND, C, P, DIST_L2COMBINED, 3, 11, CALL)                                     \
    NLI_SPECIAL(KIND, C, P, DIST_L2, 3, 9, CALL)                                              \
    NLI_SPECIAL(KIND, C, P, DIST_L2, 3, 7, CALL)                                              \
    NLI_SPECIAL(KIND, C, P, DIST_L1, 3, 9, CALL)                                              \
    NLI_SPECIAL(KIND, C, P, DIST_L1, 3, 7, CALL)                                              \
    NLI_SPECIAL(KIND, C, P, DIST_L2RGBD, 4, 9, CALL)                                          \
    {                                                                                         \
        NLI_DISPATCH_RT(KIND, C, CALL)                                                        \
    }

struct Prepared {
    PMArgs a;
    int ntiles = 0;
    int targets = 0;   // read back only when kernel timing is enabled
};

Prepared prepare(const DistDesc &d, const float *src, const uint8_t *smask, int sW, int sH, const float *tgt,
                 const uint8_t *tmask, int tW, int tH, int C, PMWork &w, hipStream_t s, bool need_tiles)
{
    if (d.patch < 1 || d.patch > 15) throw std::invalid_argument("patch side must be in [1, 15]");
    if (C < 1 || C > 4) throw std::invalid_argument("1 to 4 channels supported");
    Prepared pr;
    PMArgs &a = pr.a;
    a = PMArgs{};
    a.src = src;
    a.smask = smask;
    a.sW = sW;
    a.sH = sH;
    a.tgt = tgt;
    a.tmask = tmask;
    a.tW = tW;
    a.tH = tH;
    a.P = d.patch;
    a.lambda = d.lambda;
    a.lambda_d = d.lambda_d;
    if (w.w_patch != d.patch || !(w.w_sigma == d.sigma)) {
        std::vector<float> wh((size_t)d.patch * d.patch);
        gaussian_weights(d.patch, d.patch, d.sigma, d.sigma, wh.data());
        w.weights.ensure(wh.size());
        NLI_HIP(hipMemcpyAsync(w.weights.p, wh.data(), wh.size() * sizeof(float), hipMemcpyHostToDevice, s));
        NLI_HIP(hipStreamSynchronize(s));   // wh is pageable and local
        w.w_patch = d.patch;
        w.w_sigma = d.sigma;
    }
    a.w = w.weights.p;
    if (d.kind == DIST_L2COMBINED || d.kind == DIST_L2RGBD) {
        w.sgrad.ensure((size_t)sW * sH * C * 2);
        gradient(src, sW, sH, C, w.sgrad.p, s);
        a.sgrad = w.sgrad.p;
        if (tgt == src) {
            a.tgrad = w.sgrad.p;
        } else {
            w.tgrad.ensure((size_t)tW * tH * C * 2);
            gradient(tgt, tW, tH, C, w.tgrad.p, s);
            a.tgrad = w.tgrad.p;
        }
    }
    const int ntx = (tW + TX - 1) / TX, nty = (tH + TY - 1) / TY;
    const bool cached = w.tiles_mask == tmask && w.tiles_W == tW && w.tiles_H == tH && w.tiles_epoch == w.epoch;
    if (need_tiles && cached) {
        pr.ntiles = w.ntiles;
        pr.targets = w.targets;
        a.tiles = w.tiles.p;
        a.ntiles_x = ntx;
        a.ntiles = pr.ntiles;
    } else if (need_tiles) {
        int nt = ntx * nty;
        w.tflags.ensure(nt);
        w.total.ensure(2);
        NLI_HIP(hipMemsetAsync(w.total.p + 1, 0, sizeof(int), s));
        k_tile_flags<<<(nt + kBlock - 1) / kBlock, kBlock, 0, s>>>(tmask, tW, tH, ntx, nty, w.tflags.p,
                                                                    w.total.p + 1);
        NLI_LAUNCH_CHECK();
        w.tiles.ensure(nt);
        w.counts.ensure(nt / 1024 + 2);
        pr.ntiles = compact_indices(w.tflags.p, nt, w.tiles.p, w.counts.p, w.total.p, s);
        NLI_HIP(hipMemcpyAsync(&pr.targets, w.total.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        NLI_HIP(hipStreamSynchronize(s));
        a.tiles = w.tiles.p;
        a.ntiles_x = ntx;
        a.ntiles = pr.ntiles;
        w.tiles_mask = tmask;
        w.tiles_W = tW;
        w.tiles_H = tH;
        w.tiles_epoch = w.epoch;
        w.ntiles = pr.ntiles;
        w.targets = pr.targets;
    }
    return pr;
}

// The pool sweep packs source coordinates into 16 bits each and holds at most kMaxR
// random-search windows.
bool pool_sweep_fits(int sW, int sH, int R) { return sW <= 65535 && sH <= 65535 && R <= kMaxR; }

int auto_first_step(int W, int H)
{
    int m = std::max(W, H) / 32;
    int s = 1;
    while (s * 2 <= m && s < 32) s *= 2;
    return s;
}

}  // namespace

void gradient(const float *img, int W, int H, int C, float *grad, hipStream_t s)
{
    dim3 g((W + kBlock - 1) / kBlock, H);
    k_gradient<<<g, kBlock, 0, s>>>(img, W, H, C, grad);
    NLI_LAUNCH_CHECK();
}

namespace {

// ---- autotuning of the pool sweep's tile arrangement ----
// Choices are kept per process and device for (kind, channels, patch, log2 tile count).
std::mutex g_tune_mu;
std::map<std::pair<int, uint64_t>, int> g_tuned;
std::atomic<int> g_forced_tpb{0};   // forced tile arrangement (0: autotuned)
std::atomic<int> g_row_scan{0};   // 0: automatic (RGB-D distances), 1: on, 2: off

int current_device()
{
    int dev = 0;
    NLI_HIP(hipGetDevice(&dev));
    return dev;
}

// compute units of the current device (looked up once per device)
int device_cus()
{
    static std::mutex mu;
    static std::map<int, int> cus;
    const int dev = current_device();
    std::lock_guard<std::mutex> g(mu);
    auto it = cus.find(dev);
    if (it != cus.end()) return it->second;
    int n = 0;
    NLI_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    return cus[dev] = n;
}

// The tile arrangement of a call: the tuned one, or 1 (one tile per workgroup) until the
// key's kTuneCall-th call, which returns 0: tune on its first sweep.  The first calls at a
// new size start from a random or coarse-scale field, whose sweeps improve many targets
// and favour two tiles per workgroup; the converged sweeps that make up most calls of an
// inpainting run favour one (measured on the 1024x1024 workload: at the third call the
// 813x813 scale still tuned to two tiles, 14 % slower in its later sweeps).
constexpr int kTuneCall = 8;
std::map<std::pair<int, uint64_t>, int> g_tune_calls;
int tuned_arrangement(uint64_t key)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    const auto k = std::make_pair(current_device(), key);
    auto it = g_tuned.find(k);
    if (it != g_tuned.end()) return it->second;
    return ++g_tune_calls[k] >= kTuneCall ? 0 : 1;
}

// Times sweeps in each of the n arrangements arr[] (the sweep is deterministic, so running
// it again on the same input writes the same output) and keeps the fastest: each
// arrangement twice, in mirrored order (1 2 2 1, 1 2 3 3 2 1, ...), so that none gains
// from running after another (caches, clocks); the smallest total wins.
template <class Launch>
int tune_arrangement(uint64_t key, const int *arr, int n, Launch &&launch, hipStream_t s)
{
    if (n < 1 || n > kArrangements) throw std::logic_error("tile arrangements: 1 to 3");
    const int runs = 2 * n;
    int order[2 * kArrangements];
    for (int i = 0; i < n; i++) order[i] = order[runs - 1 - i] = arr[i];
    hipEvent_t ev[2 * kArrangements + 1];
    for (int i = 0; i <= runs; i++) NLI_HIP(hipEventCreate(&ev[i]));
    NLI_HIP(hipEventRecord(ev[0], s));
    for (int i = 0; i < runs; i++) {
        launch(order[i]);
        NLI_HIP(hipEventRecord(ev[i + 1], s));
    }
    NLI_HIP(hipEventSynchronize(ev[runs]));
    float total[kArrangements + 1] = {};
    for (int i = 0; i < runs; i++) {
        float ms = 0.0f;
        NLI_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        total[order[i]] += ms;
    }
    int best = arr[0];
    for (int i = 1; i < n; i++)
        if (total[arr[i]] < total[best]) best = arr[i];
    for (int i = 0; i <= runs; i++) NLI_HIP(hipEventDestroy(ev[i]));
    std::lock_guard<std::mutex> g(g_tune_mu);
    return g_tuned.emplace(std::make_pair(current_device(), key), best).first->second;
}

}  // namespace

void set_pm_tiles_per_block(int arrangement)
{
    if (arrangement < 0 || arrangement > kArrangements)
        throw std::invalid_argument("tile arrangement: 0 (auto), 1 (one tile per workgroup), 2 (two tiles per "
                                    "workgroup) or 3 (one tile, two waves per tile row)");
    g_forced_tpb = arrangement;
}
int pm_tiles_per_block() { return g_forced_tpb.load(); }

void set_pm_row_scan(int mode)
{
    if (mode < 0 || mode > 2) throw std::invalid_argument("row scan: 0 (automatic), 1 (on) or 2 (off)");
    g_row_scan = mode;
}
int pm_row_scan() { return g_row_scan.load(); }
// Automatic: on for L2RGBD, whose depth-gradient term makes the sweep's propagation
// range matter most; the colour and gradient distances reach the serial reference's
// energy without it.
int pm_row_scan_for(int kind)
{
    const int m = g_row_scan.load();
    return m == 1 || (m == 0 && kind == DIST_L2RGBD) ? 1 : 0;
}

void patchmatch(const DistDesc &d, const PMConfig &cfg, const float *src, const uint8_t *smask, int sW, int sH,
                const float *tgt, const uint8_t *tmask, int tW, int tH, int C, const int32_t *init_nnf,
                uint64_t seed, int32_t *out_nnf, float *out_dist, PMWork &w, hipStream_t s)
{
    Prepared pr = prepare(d, src, smask, sW, sH, tgt, tmask, tW, tH, C, w, s, true);
    PMArgs a = pr.a;
    size_t tn = (size_t)tW * tH;
    float *dA = out_dist ? out_dist : w.dist0.ensure(tn);
    int32_t *nB = w.nnf1.ensure(2 * tn);
    float *dB = w.dist1.ensure(tn);
    // the final sweep must land in (out_nnf, dA)
    int K = std::max(cfg.iterations, 0);
    int2 *n_cur = (int2 *)((K % 2 == 0) ? out_nnf : nB);
    float *d_cur = (K % 2 == 0) ? dA : dB;
    int2 *n_oth = (int2 *)((K % 2 == 0) ? nB : out_nnf);
    float *d_oth = (K % 2 == 0) ? dB : dA;
    fill_fields((int32_t *)n_cur, d_cur, K > 0 ? (int32_t *)n_oth : nullptr, K > 0 ? d_oth : nullptr, tn, s);
    if (pr.ntiles == 0) {
        if (K % 2 == 1) {   // result must still be in out_nnf / dA
            NLI_HIP(hipMemcpyAsync(out_nnf, n_cur, 2 * tn * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
            NLI_HIP(hipMemcpyAsync(dA, d_cur, tn * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        return;
    }
    a.shots = cfg.shots;
    a.maxw = (cfg.window != -1) ? cfg.window : std::max(sW, sH);
    a.seed = seed;
    a.init = (const int2 *)init_nnf;
    a.nnf_out = n_cur;
    a.dist_out = d_cur;
    a.stream = 0xFFFFu;
    a.row_scan = pm_row_scan_for(d.kind);
    const int init_r = d.patch / 2;
    const int init_F = (d.kind == DIST_L2COMBINED || d.kind == DIST_L2RGBD) ? 3 * C : C;
    const size_t init_lds =
        ((size_t)init_F * (TX + 2 * init_r) * (TY + 2 * init_r) + (size_t)d.patch * d.patch) * sizeof(float);
    prof::begin(TK_PM_INIT, s);
#define NLI_INIT(KD, CC, PP) k_pm_init<KD, CC, PP><<<pr.ntiles, TX * TY, init_lds, s>>>(a)
    NLI_DISPATCH(d.kind, C, d.patch, NLI_INIT);
#undef NLI_INIT
    NLI_LAUNCH_CHECK();
    prof::end(TK_PM_INIT, s, 0.0, pr.targets);
    const int r = d.patch / 2;
    const int F = (d.kind == DIST_L2COMBINED || d.kind == DIST_L2RGBD) ? 3 * C : C;
    const size_t lds = ((size_t)F * (TX + 2 * r) * (TY + 2 * r) + (size_t)d.patch * d.patch) * sizeof(float);
    const int sweep_blocks = (pr.ntiles + kXCDs - 1) / kXCDs * kXCDs;
    const int s0 = cfg.first_step > 0 ? cfg.first_step : auto_first_step(tW, tH);
    int levels = 0;
    while ((s0 >> levels) > 1) levels++;
    // algorithmic flops of one sweep: T targets x (2 + R) candidates x P^2 pixels x
    // (9C + 4) (gradient distances) or (3C + 2) (colour-only distances)
    int R = 0;
    for (int ws = a.maxw; ws >= 1; ws /= 2) R++;
    // L2RGBD at lambda_rgb 1, lambda_d 0 runs the variant without the zero-weighted terms:
    // 19 flops a pixel (3 colour differences, squares and sums; 2 depth forward
    // differences, 2 differences, squares and a sum; 2 weighted accumulations)
    const bool rgbd_cd = d.kind == DIST_L2RGBD && d.lambda == 1.0f && d.lambda_d == 0.0f;
    const double per_pixel =
        rgbd_cd ? 19.0 : (d.kind == DIST_L2COMBINED || d.kind == DIST_L2RGBD) ? 9.0 * C + 4 : 3.0 * C + 2;
    const double sweep_flops = (double)pr.targets * (2 + R) * d.patch * d.patch * per_pixel;
    // candidate-pool sweep for the specialised gradient distances (up to kMaxR windows)
    const bool pool = pool_sweep_fits(sW, sH, R) && ((d.kind == DIST_L2COMBINED && C == 3 &&
                                      (d.patch == 7 || d.patch == 9 || d.patch == 11)) ||
                                     (d.kind == DIST_L2RGBD && C == 4 && d.patch == 9) ||
                                     ((d.kind == DIST_L2 || d.kind == DIST_L1) && C == 3 &&
                                      (d.patch == 7 || d.patch == 9)));
    // launches one candidate-pool sweep in tile arrangement `arr` (1, 2 or 3)
    auto launch_pool = [&](int arr) {
        const int tpb = arr == 2 ? 2 : 1, wpr = arr == 3 ? 2 : 1;
        const size_t lds_pool =
            pool_tables_offset(F, d.patch, tpb) + (size_t)kWaves * wpr * pool_wave_bytes(R, tpb, wpr);
        const int blocks = ((pr.ntiles + tpb - 1) / tpb + kXCDs - 1) / kXCDs * kXCDs;
#define NLI_POOL(KD, CC, PP)                                                                          \
    (arr == 2 ? k_pm_sweep_pool<KD, CC, PP, 2, 1>                                                     \
              : (arr == 3 ? k_pm_sweep_pool<KD, CC, PP, 1, 2> : k_pm_sweep_pool<KD, CC, PP, 1, 1>))     \
        <<<blocks, kBlock * wpr, lds_pool, s>>>(a)
        if (rgbd_cd) NLI_POOL(DIST_L2RGBD_CD, 4, 9);
        else if (d.kind == DIST_L2RGBD) NLI_POOL(DIST_L2RGBD, 4, 9);
        else if (d.kind == DIST_L2 && d.patch == 7) NLI_POOL(DIST_L2, 3, 7);
        else if (d.kind == DIST_L2) NLI_POOL(DIST_L2, 3, 9);
        else if (d.kind == DIST_L1 && d.patch == 7) NLI_POOL(DIST_L1, 3, 7);
        else if (d.kind == DIST_L1) NLI_POOL(DIST_L1, 3, 9);
        else if (d.patch == 7) NLI_POOL(DIST_L2COMBINED, 3, 7);
        else if (d.patch == 9) NLI_POOL(DIST_L2COMBINED, 3, 9);
        else NLI_POOL(DIST_L2COMBINED, 3, 11);
#undef NLI_POOL
        NLI_LAUNCH_CHECK();
    };
    // tile arrangement: forced, tuned earlier for this key (kind, the L2RGBD lambda (1, 0)
    // variant, channels, patch, row scan, whether the split-row arrangement is offered,
    // size class), the default before tuning, or tuned now on the first sweep (every
    // arrangement gives identical fields).  The split-row arrangement 3 needs the row scan
    // off (forced 3 with the scan on runs arrangement 1), and it pays only where the tiles
    // leave compute units idle (1024x1024 workload: 7-12 % faster sweeps at 72-144 tiles,
    // slower from 200 tiles, 26 % slower at 3008), so the tuner is offered it up to 3/4
    // of the compute units
    const bool split_ok = !a.row_scan && (int64_t)pr.ntiles * 4 <= (int64_t)device_cus() * 3;
    int arr_list[kArrangements] = {1, 2}, n_arr = 2;
    if (split_ok) arr_list[n_arr++] = 3;
    int arr = pool ? pm_tiles_per_block() : 0;
    if (arr == 3 && a.row_scan) arr = 1;
    const uint64_t tune_key = ((uint64_t)d.kind << 56) | ((uint64_t)C << 48) | ((uint64_t)d.patch << 40) |
                              ((uint64_t)(rgbd_cd ? 1 : 0) << 35) | ((uint64_t)(split_ok ? 1 : 0) << 34) |
                              ((uint64_t)(a.row_scan ? 1 : 0) << 32) |
                              (uint64_t)(32 - __builtin_clz((unsigned)pr.ntiles | 1u));
    if (pool && arr == 0) arr = tuned_arrangement(tune_key);
    auto set_sweep = [&](int k) {
        a.nnf_in = n_cur;
        a.dist_in = d_cur;
        a.nnf_out = n_oth;
        a.dist_out = d_oth;
        a.step = s0 >> (k % (levels + 1));
        a.stream = (uint32_t)k;
    };
    // tuning runs on the first sweep's input (its outputs are overwritten by that sweep),
    // outside the timed span
    if (K > 0 && pool && arr == 0) {
        set_sweep(0);
        arr = tune_arrangement(tune_key, arr_list, n_arr, launch_pool, s);
    }
    // the K sweeps are timed as one span (one event pair: a pair per sweep would add its
    // markers' dispatch to every sweep)
    if (K > 0) prof::begin(TK_PM_SWEEP, s);
    for (int k = 0; k < K; k++) {
        set_sweep(k);
        if (pool) {
            launch_pool(arr);
        } else {
#define NLI_SWEEP(KD, CC, PP) k_pm_sweep<KD, CC, PP><<<sweep_blocks, kBlock, lds, s>>>(a)
            NLI_DISPATCH(d.kind, C, d.patch, NLI_SWEEP);
#undef NLI_SWEEP
            NLI_LAUNCH_CHECK();
        }
        std::swap(n_cur, n_oth);
        std::swap(d_cur, d_oth);
    }
    if (K > 0) prof::end(TK_PM_SWEEP, s, sweep_flops * K, (double)pr.targets * K, K);
}

// ---- several PatchMatch calls in lockstep ----
bool pm_batch_supported(const DistDesc &d, int C, int sW, int sH, const PMConfig &cfg)
{
    const int maxw = (cfg.window != -1) ? cfg.window : std::max(sW, sH);
    int R = 0;
    for (int ws = maxw; ws >= 1; ws /= 2) R++;
    return pool_sweep_fits(sW, sH, R) &&
           ((d.kind == DIST_L2COMBINED && C == 3 && (d.patch == 7 || d.patch == 9 || d.patch == 11)) ||
            (d.kind == DIST_L2RGBD && C == 4 && d.patch == 9) ||
            ((d.kind == DIST_L2 || d.kind == DIST_L1) && C == 3 && (d.patch == 7 || d.patch == 9)));
}

void patchmatch_batch(const PMJob *jobs, int n, PMBatchWork &bw, hipStream_t s)
{
    if (n <= 0) return;
    const DistDesc &d0 = jobs[0].d;
    const int C = jobs[0].C, K = std::max(jobs[0].cfg.iterations, 0);
    for (int i = 0; i < n; i++) {
        const PMJob &j = jobs[i];
        if (j.d.kind != d0.kind || j.d.patch != d0.patch || j.C != C || std::max(j.cfg.iterations, 0) != K ||
            !pm_batch_supported(j.d, j.C, j.sW, j.sH, j.cfg))
            throw std::invalid_argument("PatchMatch batch: jobs must share a supported distance, channels, patch "
                                        "and sweep count");
    }
    // per job: source gradient and target tiles (prepare), the output and ping-pong
    // fields reset, the arguments; then one launch per phase for all jobs
    const size_t args_bytes = sizeof(PMArgs) * n;
    PMArgs *A = static_cast<PMArgs *>(bw.args.fill(args_bytes));
    int *st = static_cast<int *>(bw.starts.fill(sizeof(int) * 3 * (n + 1)));
    int *st_init = st, *st1 = st + (n + 1), *st2 = st + 2 * (n + 1);
    st_init[0] = st1[0] = st2[0] = 0;
    int maxR = 0;
    double sweep_flops = 0, targets = 0;
    for (int i = 0; i < n; i++) {
        const PMJob &j = jobs[i];
        Prepared pr = prepare(j.d, j.src, j.smask, j.sW, j.sH, j.tgt, j.tmask, j.tW, j.tH, C, *j.w, s, true);
        PMArgs a = pr.a;
        const size_t tn = (size_t)j.tW * j.tH;
        float *dA = j.out_dist ? j.out_dist : j.w->dist0.ensure(tn);
        int32_t *nB = j.w->nnf1.ensure(2 * tn);
        float *dB = j.w->dist1.ensure(tn);
        // the final sweep must land in (out_nnf, dA): sweep k reads pair[k & 1]
        int2 *p0 = (int2 *)((K % 2 == 0) ? j.out_nnf : nB), *p1 = (int2 *)((K % 2 == 0) ? nB : j.out_nnf);
        float *q0 = (K % 2 == 0) ? dA : dB, *q1 = (K % 2 == 0) ? dB : dA;
        fill_fields((int32_t *)p0, q0, K > 0 ? (int32_t *)p1 : nullptr, K > 0 ? q1 : nullptr, tn, s);
        a.shots = j.cfg.shots;
        a.maxw = (j.cfg.window != -1) ? j.cfg.window : std::max(j.sW, j.sH);
        a.seed = j.seed;
        a.init = (const int2 *)j.init_nnf;
        a.stream = 0xFFFFu;
        a.row_scan = pm_row_scan_for(j.d.kind);
        a.nnf_pp[0] = p0;
        a.nnf_pp[1] = p1;
        a.dist_pp[0] = q0;
        a.dist_pp[1] = q1;
        a.s0 = j.cfg.first_step > 0 ? j.cfg.first_step : auto_first_step(j.tW, j.tH);
        a.levels = 0;
        while ((a.s0 >> a.levels) > 1) a.levels++;
        a.active = j.active;
        A[i] = a;
        int R = 0;
        for (int ws = a.maxw; ws >= 1; ws /= 2) R++;
        maxR = std::max(maxR, R);
        st_init[i + 1] = st_init[i] + pr.ntiles;
        st1[i + 1] = st1[i] + pr.ntiles;
        st2[i + 1] = st2[i] + (pr.ntiles + 1) / 2;
        const double per_pixel = (j.d.kind == DIST_L2COMBINED || j.d.kind == DIST_L2RGBD) ? 9.0 * C + 4 : 3.0 * C + 2;
        sweep_flops += (double)pr.targets * (2 + R) * j.d.patch * j.d.patch * per_pixel;
        targets += pr.targets;
    }
    const PMArgs *dargs = static_cast<const PMArgs *>(bw.args.upload(args_bytes, s));
    const int *dst = static_cast<const int *>(bw.starts.upload(sizeof(int) * 3 * (n + 1), s));
    const int total_tiles = st_init[n];
    if (total_tiles == 0) return;
    const int P = d0.patch, r = P / 2;
    const int F = (d0.kind == DIST_L2COMBINED || d0.kind == DIST_L2RGBD) ? 3 * C : C;
    const size_t init_lds = ((size_t)F * (TX + 2 * r) * (TY + 2 * r) + (size_t)P * P) * sizeof(float);
    prof::begin(TK_PM_INIT, s);
#define NLI_INIT_B(KD, CC, PP) k_pm_init_batch<KD, CC, PP><<<total_tiles, TX * TY, init_lds, s>>>(dargs, dst, n)
    NLI_DISPATCH(d0.kind, C, P, NLI_INIT_B);
#undef NLI_INIT_B
    NLI_LAUNCH_CHECK();
    prof::end(TK_PM_INIT, s, 0.0, targets);
    auto launch = [&](int tpb, int k) {
        const int *start = tpb == 1 ? dst + (n + 1) : dst + 2 * (n + 1);
        const int slots = tpb == 1 ? st1[n] : st2[n];
        const size_t lds_pool = pool_tables_offset(F, P, tpb) + kWaves * pool_wave_bytes(maxR, tpb, 1);
        const int blocks = (slots + kXCDs - 1) / kXCDs * kXCDs;
#define NLI_POOL_B(KD, CC, PP)                                                                        \
    (tpb == 1 ? k_pm_sweep_pool_batch<KD, CC, PP, 1> : k_pm_sweep_pool_batch<KD, CC, PP, 2>)            \
        <<<blocks, kBlock, lds_pool, s>>>(dargs, start, n, k)
        if (d0.kind == DIST_L2RGBD) NLI_POOL_B(DIST_L2RGBD, 4, 9);
        else if (d0.kind == DIST_L2 && P == 7) NLI_POOL_B(DIST_L2, 3, 7);
        else if (d0.kind == DIST_L2) NLI_POOL_B(DIST_L2, 3, 9);
        else if (d0.kind == DIST_L1 && P == 7) NLI_POOL_B(DIST_L1, 3, 7);
        else if (d0.kind == DIST_L1) NLI_POOL_B(DIST_L1, 3, 9);
        else if (P == 7) NLI_POOL_B(DIST_L2COMBINED, 3, 7);
        else if (P == 9) NLI_POOL_B(DIST_L2COMBINED, 3, 9);
        else NLI_POOL_B(DIST_L2COMBINED, 3, 11);
#undef NLI_POOL_B
        NLI_LAUNCH_CHECK();
    };
    // tile arrangement 1 or 2: the lockstep kernels take whole tile rows per wave (a forced
    // split-row arrangement 3 runs as 1); batched calls tune under keys of their own
    int tpb = pm_tiles_per_block();
    if (tpb > 2) tpb = 1;
    const uint64_t tune_key = ((uint64_t)d0.kind << 56) | ((uint64_t)C << 48) | ((uint64_t)P << 40) |
                              (1ull << 33) | (uint64_t)(32 - __builtin_clz((unsigned)total_tiles | 1u));
    if (tpb == 0) tpb = tuned_arrangement(tune_key);
    static constexpr int kBatchArrangements[2] = {1, 2};
    if (K > 0 && tpb == 0) tpb = tune_arrangement(tune_key, kBatchArrangements, 2, [&](int t) { launch(t, 0); }, s);
    if (K > 0) prof::begin(TK_PM_SWEEP, s);   // the K sweeps as one timed span
    for (int k = 0; k < K; k++) launch(tpb, k);
    if (K > 0) prof::end(TK_PM_SWEEP, s, sweep_flops * K, targets * K, K);
}

// ---- fixed-field distances without allocation ----
// Device memory of the allocation-free entry point: the source and target gradients and
// the patch weights.  Every GPU holds its own copy until the process ends; every call on
// a GPU uses that GPU's copy, so calls that overlap on one GPU (several streams) race.
__device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` on each device
}  // namespace

void noalloc_scratch_init()
{
    const int dev = current_device();
    if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[dev]) return;
    void *p = nullptr;
    NLI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<float *>(p);
}

void patch_distance_field_noalloc(const DistDesc &d, const float *src, int sW, int sH, const float *tgt, int tW,
                                  int tH, int C, const int32_t *nnf, const uint8_t *tmask, float *out, hipStream_t s)
{
    if (d.patch < 1 || d.patch > 15) throw std::invalid_argument("patch side must be in [1, 15]");
    if (C < 1 || C > 4) throw std::invalid_argument("1 to 4 channels supported");
    const int dev = current_device();
    float *base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        base = (dev >= 0 && dev < kMaxDevices) ? g_scratch_ptr[dev] : nullptr;
    }
    if (!base) throw std::logic_error("no-allocation scratch not set up on this device (create a context first)");
    const bool grad = d.kind == DIST_L2COMBINED || d.kind == DIST_L2RGBD;
    const size_t wfl = 256;   // weights region (P^2 <= 225 floats), keeps the gradients 1 KiB aligned
    const size_t sg = grad ? (size_t)sW * sH * C * 2 : 0, tg = (grad && tgt != src) ? (size_t)tW * tH * C * 2 : 0;
    if (wfl + sg + tg > kNoAllocScratchFloats)
        throw std::invalid_argument("image too large for the no-allocation scratch (" +
                                    std::to_string(kNoAllocScratchFloats * sizeof(float)) +
                                    " bytes per device): 256 + 2 C (source + distinct target pixels) floats needed");
    float wh[15 * 15];
    gaussian_weights(d.patch, d.patch, d.sigma, d.sigma, wh);
    NLI_HIP(hipMemcpyAsync(base, wh, sizeof(float) * d.patch * d.patch, hipMemcpyHostToDevice, s));
    NLI_HIP(hipStreamSynchronize(s));   // wh is on the host stack
    PMArgs a{};
    a.src = src;
    a.smask = tmask;   // not read for a fixed field
    a.sW = sW;
    a.sH = sH;
    a.tgt = tgt;
    a.tmask = tmask;
    a.tW = tW;
    a.tH = tH;
    a.P = d.patch;
    a.lambda = d.lambda;
    a.lambda_d = d.lambda_d;
    a.w = base;
    if (grad) {
        a.sgrad = base + wfl;
        gradient(src, sW, sH, C, base + wfl, s);
        a.tgrad = a.sgrad;
        if (tgt != src) {
            a.tgrad = base + wfl + sg;
            gradient(tgt, tW, tH, C, base + wfl + sg, s);
        }
    }
    a.dist_out = out;
    dim3 g((tW + kBlock - 1) / kBlock, tH);
#define NLI_FIELD(KD, CC, PP) k_dist_field<KD, CC, PP><<<g, kBlock, 0, s>>>(a, (const int2 *)nnf)
    NLI_DISPATCH(d.kind, C, d.patch, NLI_FIELD);
#undef NLI_FIELD
    NLI_LAUNCH_CHECK();
}

void patch_distance_field(const DistDesc &d, const float *src, int sW, int sH, const float *tgt, int tW, int tH,
                          int C, const int32_t *nnf, const uint8_t *tmask, float *out, PMWork &w, hipStream_t s)
{
    // the source mask is not needed for a fixed field; pass the target mask (unused)
    Prepared pr = prepare(d, src, tmask, sW, sH, tgt, tmask, tW, tH, C, w, s, false);
    PMArgs a = pr.a;
    a.dist_out = out;
    dim3 g((tW + kBlock - 1) / kBlock, tH);
#define NLI_FIELD(KD, CC, PP) k_dist_field<KD, CC, PP><<<g, kBlock, 0, s>>>(a, (const int2 *)nnf)
    NLI_DISPATCH(d.kind, C, d.patch, NLI_FIELD);
#undef NLI_FIELD
    NLI_LAUNCH_CHECK();
}

}  // namespace nli
