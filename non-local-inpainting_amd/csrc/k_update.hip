// k_update.hip -- the non-local image updates (gfx950): NL-means, NL-medians,
// NL-Poisson (coefficients + conjugate gradient) and NL-Depth (two NL-Poisson
// solves on channel subsets).
//
// Reference: patch_non_local_means.cpp:23-100, patch_non_local_medians.cpp:23-126,
// patch_non_local_poisson.cpp:36-477, patch_non_local_depth.cpp:36-55.
//
// * Means/medians: one thread per hole pixel gathers its clipped P x P window in the
//   reference's order (x outer, y inner), float sums as in the reference.  New values
//   go to a side buffer and are written back afterwards, so the result does not depend
//   on thread order (the reference updates in place in scan order; for NNF entries that
//   are source points the two are identical, contributors never being hole pixels).
// * Poisson coefficients: the reference scatters from every extended-domain centre p
//   into the P x P pixels q around it.  Here each pixel q gathers from the centres p
//   around it, visiting them in row-major order -- the order in which the scatter
//   reaches q -- so the double sums a1, a2, f1, F1, F2 are bit-exact.
// * CG: the reference recurrence (r recomputed as b - A x each iteration, stop on
//   |r|^2 < tol |r0|^2 or iteration >= limit).  Unknowns are the hole pixels in
//   row-major order; A x = a1 x - L x with the reference's anisotropic 5-point
//   stencil.  Two kernels per iteration for all channels at once; the direction and
//   iterate updates of the neighbours are recomputed on the fly (double-buffered).
//   Each kernel stores block partials of its dot product and the next kernel sums
//   them itself, in a fixed order, so the run is deterministic and needs no grid-wide
//   fence or atomic.  Scalars (alpha, beta, stop flags) stay on the device.
// * Mixed precision (precision = 1): the same recurrence with FP32 vectors and FP32
//   stencil arithmetic, FP64 dot-product accumulation, and the residual recomputed
//   from an FP64 evaluation of b - A x every `refine_every` iterations.
#include "nli_internal.h"
#include "nli_introsort.h"

#include <atomic>
#include <chrono>
#include <cstring>
#include <cfloat>
#include <cmath>
#include <type_traits>
#include <vector>

namespace nli {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxNC = 4;
constexpr int kPollSlot = kMaxNC + 1;   // a CG poll slot: kMaxNC stop flags, then its stamp

// The update's kernels are short dependent chains (the CG above all) that share the CUs
// with other streams' PatchMatch sweeps, which keep the vector units busy.  Their waves
// run at raised issue priority so that they are not starved there (batch of 8 images,
// mean of 5 runs: 4.49 -> 4.61 MPix/s).  A wave priority is local to the kernel's waves.
__device__ inline void raise_wave_priority() { __builtin_amdgcn_s_setprio(3); }
constexpr int kMaxP = 15;

// scalar slots per channel
// Per-update scalars read back at the end of an update, in one copy: the total squared
// change, then per channel the CG's stopping iteration, |r_0|^2 and its final |r|^2.
constexpr int kSumTotal = 0, kSumDone = 1, kSumNr0 = 1 + kMaxNC, kSumNr = 1 + 2 * kMaxNC;
constexpr int kSummary = 16;   // doubles; the CG history follows

// ---------------- deterministic block reductions ----------------
// Butterfly sum over the 64 lanes of a wave: every lane ends with the same bits
// (each step adds the same two values in both partner lanes).
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block totals of v[] in a fixed order (lanes by butterfly, then the waves in order);
// every thread of the block receives them.  All threads of the block must call it.
template <int NC>
__device__ inline void block_sum(double (&v)[NC])
{
    constexpr int kWaves = kBlock / 64;
    __shared__ double sh[kWaves][NC];
    const int w = (int)threadIdx.x / 64;
#pragma unroll
    for (int c = 0; c < NC; c++) v[c] = wave_sum(v[c]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < NC; c++) sh[w][c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; c++) {
        double t = sh[0][c];
        for (int i = 1; i < kWaves; i++) t += sh[i][c];
        v[c] = t;
    }
    __syncthreads();
}

template <int NC>
__device__ inline void block_reduce_store_at(double (&v)[NC], double *partials, int nb, int blk)
{
    block_sum<NC>(v);
    if (threadIdx.x == 0)
        for (int c = 0; c < NC; c++) partials[(size_t)c * nb + blk] = v[c];
}
template <int NC>
__device__ inline void block_reduce_store(double (&v)[NC], double *partials, int nb)
{
    block_reduce_store_at<NC>(v, partials, nb, (int)blockIdx.x);
}

// The job of a block in a batched launch: start[i] is the first block of job i (start[n] =
// the total), so the job is the last i with start[i] <= b.
__device__ inline int batch_job(const int *__restrict__ start, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Total of the nb block partials of a one-value reduction, in a fixed order: thread t
// sums partials t, t + kBlock, ..., then a tree over the threads.  One block; it runs
// after the kernel that stored the partials, so no grid-wide fence is needed.  With a
// CG history (hist != nullptr) it also writes the CG part of the update's summary: per
// channel the stopping iteration, |r_0|^2 and the final |r|^2.
// The summary also goes to pinned host memory (host_out, when given), followed by a
// stamp that the host waits for instead of a copy and a stream synchronisation: each
// writer fences at system scope before the barrier after which thread 0 writes the stamp.
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *__restrict__ partials, int nb,
                                                         double *__restrict__ out, const double *__restrict__ hist,
                                                         int hist_len, const int *__restrict__ done_at, int nc,
                                                         double *host_out, unsigned *host_stamp, unsigned stamp)
{
    raise_wave_priority();
    __shared__ double sh[kBlock];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += partials[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[kSumTotal] = sh[0];
        if (host_out) host_out[kSumTotal] = sh[0];
    }
    const int c = threadIdx.x;
    if (hist && c < nc) {
        const int d = done_at[c];
        const double v[3] = {(double)d, hist[(size_t)c * hist_len], hist[(size_t)c * hist_len + (d > 0 ? d : 0)]};
        out[kSumDone + c] = v[0];
        out[kSumNr0 + c] = v[1];
        out[kSumNr + c] = v[2];
        if (host_out) {
            host_out[kSumDone + c] = v[0];
            host_out[kSumNr0 + c] = v[1];
            host_out[kSumNr + c] = v[2];
        }
    }
    if (host_out) {
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0) *reinterpret_cast<volatile unsigned *>(host_stamp) = stamp;
    }
}

// ---------------- 16 lanes per hole pixel ----------------
// A wavefront takes 4 hole pixels, 16 lanes each.  The lanes of a pixel gather its
// window's contributors together (element e = lane + 16 s, in the reference's order: x
// outer, y inner) and compact them in that order into the pixel's LDS arrays with
// ballots.  The order-dependent float sums then run over the LDS copy.  A pixel costs one
// or two dependent memory trips per channel instead of one per window element, and the
// sequential parts run for 4 pixels at once.
constexpr int kNLWaves = kBlock / 64;
constexpr int kNLG = 16;               // lanes per pixel
constexpr int kNLPix = 64 / kNLG;      // pixels per wavefront

__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// contributors of hole pixel p (live = false: none): source index, weight and gather
// position (-1: not a contributor); returns their count m and writes the weights to
// lw[pos].  Every lane of the wavefront calls it.
template <int SLOTS>
__device__ __attribute__((always_inline)) inline int nl_gather(bool live, int p, int W, int H, int P, const uint8_t *__restrict__ hole,
                                const int2 *__restrict__ nnf, const float *__restrict__ conf, const float *wl,
                                float *lw, int (&src)[SLOTS], float (&wt)[SLOTS], int (&pos)[SLOTS])
{
    const int lane = (int)threadIdx.x & 63, gl = lane % kNLG, grp = lane / kNLG;
    const int x = p % W, y = p / W, h = P / 2;
    const int x_a = max(x - h, 0), x_b = min(x + h, W - 1);
    const int y_a = max(y - h, 0), y_b = min(y + h, H - 1);
    const int nh = y_b - y_a + 1, cnt = live ? (x_b - x_a + 1) * nh : 0;
    // all loads first, from window pixels (the window's corner for elements past the
    // window), so they travel together instead of one dependent trip per element
    int ii[SLOTS], jj[SLOTS];
    int2 c[SLOTS];
    float cf[SLOTS];
    bool hf[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int e = gl + kNLG * s;
        const bool in = e < cnt;
        ii[s] = x_a + (in ? e / nh : 0);
        jj[s] = y_a + (in ? e % nh : 0);
        const size_t q = (size_t)W * jj[s] + ii[s];
        c[s] = nnf[q];
        hf[s] = conf && hole[q];
        cf[s] = conf ? conf[q] : 1.0f;
    }
    int m = 0;
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int e = gl + kNLG * s;
        const int i = ii[s], j = jj[s];
        int2 cc = c[s];
        bool ok = e < cnt && cc.x != -1;
        cc.x += x - i;
        cc.y += y - j;
        ok = ok && cc.x >= 0 && cc.y >= 0 && cc.x < W && cc.y < H;
        float v = 0.0f;
        if (ok) {
            v = wl[P * (j - y_a) + (i - x_a)];
            if (hf[s]) v *= cf[s];
        }
        const uint32_t gb = (uint32_t)(__ballot(ok) >> (kNLG * grp)) & ((1u << kNLG) - 1);
        pos[s] = ok ? m + __popc(gb & ((1u << gl) - 1)) : -1;
        m += __popc(gb);
        src[s] = ok ? W * cc.y + cc.x : 0;
        wt[s] = v;
        if (ok) lw[pos[s]] = v;
    }
    wave_lds_sync();
    return m;
}

// Order key of a colour for std::sort's comparison: the float order as unsigned integers
// (sign-magnitude folded), with -0 and +0 one key, as the float comparison has them equal.
__device__ inline uint32_t order_key(float x)
{
    const uint32_t b = __float_as_uint(x + 0.0f);   // -0 + 0 = +0
    return b ^ ((uint32_t)((int)b >> 31) | 0x80000000u);
}

// NL-means (patch_non_local_means.cpp:23-100): weighted average in gather order
template <int SLOTS>
__global__ __launch_bounds__(kBlock) void k_nlmeans_w(const float *__restrict__ img, int Cs, int c0, int nc,
                                                      const int *__restrict__ pts, int n,
                                                      const uint8_t *__restrict__ hole,
                                                      const int2 *__restrict__ nnf, const float *__restrict__ conf,
                                                      const float *__restrict__ w, int P, int W, int H,
                                                      float *__restrict__ newv, uint8_t *__restrict__ has)
{
    constexpr int kE = kNLG * SLOTS;
    __shared__ float wl[kMaxP * kMaxP];
    __shared__ __align__(16) float lw[kNLWaves][kNLPix][kE];
    __shared__ __align__(16) float lv[kNLWaves][kNLPix][kMaxNC][kE];
    for (int i = threadIdx.x; i < P * P; i += blockDim.x) wl[i] = w[i];
    __syncthreads();
    const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x & 63;
    const int gl = lane % kNLG, grp = lane / kNLG;
    float *mw = lw[wave][grp];
    for (int base = (blockIdx.x * kNLWaves + wave) * kNLPix; base < n; base += gridDim.x * kNLWaves * kNLPix) {
        const int k = base + grp;                  // the loop bound is uniform per wavefront
        const bool live = k < n;
        int src[SLOTS], pos[SLOTS];
        float wt[SLOTS];
        const int m = nl_gather<SLOTS>(live, pts[live ? k : base], W, H, P, hole, nnf, conf, wl, mw, src, wt, pos);
        // every slot loads (non-contributors read pixel 0), then the contributors store
        float val[SLOTS][kMaxNC];
#pragma unroll
        for (int s = 0; s < SLOTS; s++)
#pragma unroll
            for (int ch = 0; ch < kMaxNC; ch++) val[s][ch] = ch < nc ? img[(size_t)src[s] * Cs + c0 + ch] : 0.0f;
#pragma unroll
        for (int s = 0; s < SLOTS; s++)
            if (pos[s] >= 0)
#pragma unroll
                for (int ch = 0; ch < kMaxNC; ch++)
                    if (ch < nc) lv[wave][grp][ch][pos[s]] = val[s][ch];
        wave_lds_sync();
        // the float sums in gather order, formed by every lane of the pixel alike, four
        // contributors per LDS read
        float tw = 0.0f, tv[kMaxNC] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int f0 = 0; f0 < m; f0 += 4) {
            const float4 w4 = *reinterpret_cast<const float4 *>(mw + f0);
            float4 v4[kMaxNC];
#pragma unroll
            for (int ch = 0; ch < kMaxNC; ch++)
                if (ch < nc) v4[ch] = *reinterpret_cast<const float4 *>(&lv[wave][grp][ch][f0]);
            const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (f0 + j >= m) break;
                const float v = wv[j];
                tw += v;
#pragma unroll
                for (int ch = 0; ch < kMaxNC; ch++) {
                    const float u[4] = {v4[ch].x, v4[ch].y, v4[ch].z, v4[ch].w};
                    if (ch < nc) tv[ch] += u[j] * v;
                }
            }
        }
        if (live && gl == 0) {
            has[k] = tw > 0.0;
            if (tw > 0.0)
                for (int ch = 0; ch < nc; ch++) newv[(size_t)k * nc + ch] = tv[ch] / tw;
        }
        wave_lds_sync();
    }
}

// NL-medians (patch_non_local_medians.cpp:23-126): per channel the contributors are put in
// std::sort's order by value and the weights subtracted from the total in that order until
// it drops below half.  Without equal values that order is the ranking by value; with
// every value equal the median is that value whatever the order; otherwise (equal values
// among different ones) the ranking gives the same median as every other order of equal
// values unless rounding of the running total could move the crossing to another value
// (the bound of median_order_free, checked during the walk), and only then lane 0 of the
// pixel sorts the (value, weight) pairs in gather order with the restatement of std::sort
// (nli_introsort.h).
template <int SLOTS>
__global__ __launch_bounds__(kBlock) void k_nlmedians_w(const float *__restrict__ img, int Cs, int c0, int nc,
                                                        const int *__restrict__ pts, int n,
                                                        const uint8_t *__restrict__ hole,
                                                        const int2 *__restrict__ nnf,
                                                        const float *__restrict__ conf,
                                                        const float *__restrict__ w, int P, int W, int H,
                                                        float *__restrict__ newv, uint8_t *__restrict__ has)
{
    constexpr int kE = kNLG * SLOTS;
    static_assert(kE % 4 == 0, "contributor arrays are read four at a time");
    __shared__ float wl[kMaxP * kMaxP];
    __shared__ __align__(16) float lw[kNLWaves][kNLPix][kE];
    __shared__ __align__(16) float lv[kNLWaves][kNLPix][kE];
    __shared__ __align__(16) float sv[kNLWaves][kNLPix][kE];
    __shared__ __align__(16) float sw[kNLWaves][kNLPix][kE];
    __shared__ int sstk[kNLWaves][kNLPix][3 * kIntrosortStack];
    for (int i = threadIdx.x; i < P * P; i += blockDim.x) wl[i] = w[i];
    __syncthreads();
    const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x & 63;
    const int gl = lane % kNLG, grp = lane / kNLG;
    float *mw = lw[wave][grp], *mv = lv[wave][grp], *ms = sv[wave][grp], *mo = sw[wave][grp];
    for (int base = (blockIdx.x * kNLWaves + wave) * kNLPix; base < n; base += gridDim.x * kNLWaves * kNLPix) {
        const int k = base + grp;
        const bool live = k < n;
        int src[SLOTS], pos[SLOTS];
        float wt[SLOTS];
        const int m = nl_gather<SLOTS>(live, pts[live ? k : base], W, H, P, hole, nnf, conf, wl, mw, src, wt, pos);
        float tw = 0.0f;
        for (int f0 = 0; f0 < m; f0 += 4) {
            const float4 w4 = *reinterpret_cast<const float4 *>(mw + f0);
            const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (f0 + j < m) tw += wv[j];
        }
        const float half = tw / 2;
        const int m4 = (m + 3) & ~3;
        for (int ch = 0; ch < nc; ch++) {
            float v[SLOTS];
#pragma unroll
            for (int s = 0; s < SLOTS; s++) v[s] = img[(size_t)src[s] * Cs + c0 + ch];   // pixel 0 if not a contributor
#pragma unroll
            for (int s = 0; s < SLOTS; s++)
                if (pos[s] >= 0) mv[pos[s]] = v[s];
            if (gl < m4 - m) mv[m + gl] = INFINITY;   // padding: ranks below no finite colour
            wave_lds_sync();
            // rank = the number of contributors before this one in std::sort's comparison
            // (colour, then gather position): one 64-bit key comparison per pair
            int rank[SLOTS];
            uint64_t key[SLOTS];
#pragma unroll
            for (int s = 0; s < SLOTS; s++) {
                rank[s] = 0;
                key[s] = ((uint64_t)order_key(v[s]) << 32) | (uint32_t)pos[s];
            }
            for (int f0 = 0; f0 < m4; f0 += 4) {
                const float4 u4 = *reinterpret_cast<const float4 *>(mv + f0);
                const float u[4] = {u4.x, u4.y, u4.z, u4.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint64_t kf = ((uint64_t)order_key(u[j]) << 32) | (uint32_t)(f0 + j);
#pragma unroll
                    for (int s = 0; s < SLOTS; s++) rank[s] += kf < key[s] ? 1 : 0;
                }
            }
#pragma unroll
            for (int s = 0; s < SLOTS; s++)
                if (pos[s] >= 0) {
                    ms[rank[s]] = v[s];
                    mo[rank[s]] = wt[s];
                }
            wave_lds_sync();
            // the walk in ranked order (equal values in gather order), with the exact totals
            // at the ends of the groups of equal values alongside (median_order_free's bound):
            // the ranked order gives the reference's median if no values tie up to the
            // crossing, or if every order of equal values crosses in the same group
            float total = tw, med = 0.0f;
            const double unit = (double)tw * (0x1p-24 + 0x1p-50), halfd = (double)half;
            double S = (double)tw, S_prev = (double)tw;
            int j_prev = 0, e = m;
            bool tie = false, crossed = false;
            float prev = 0.0f;
            // four sorted pairs per LDS read; the running totals step one pair at a time
            for (int e0 = 0; e0 < m && !crossed; e0 += 4) {
                const float4 a4 = *reinterpret_cast<const float4 *>(ms + e0);
                const float4 b4 = *reinterpret_cast<const float4 *>(mo + e0);
                const float av[4] = {a4.x, a4.y, a4.z, a4.w}, bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int ee = e0 + j;
                    if (ee < m && !crossed) {
                        if (ee > 0 && av[j] != prev) {
                            S_prev = S;
                            j_prev = ee;
                        } else if (ee > 0) {
                            tie = true;
                        }
                        total -= bv[j];
                        S -= (double)bv[j];
                        if (total < half) {
                            med = av[j];
                            crossed = true;
                            e = ee;
                        }
                        prev = av[j];
                    }
                }
            }
            bool ok = !tie;
            if (crossed) {
                int j = e + 1;
                for (; j < m && ms[j] == ms[e]; j++) {
                    S -= (double)mo[j];
                    tie = true;
                }
                ok = !tie || (tw > 0.0f && S_prev - j_prev * unit >= halfd && S + j * unit < halfd);
                // equal colours that differ in their bits (-0.0 and +0.0) leave the
                // median's bits to the order within the group: then std::sort's order
                if (ok && tie) {
                    const uint32_t bits = __float_as_uint(ms[e]);
                    for (int i = j_prev; i < j && ok; i++) ok = __float_as_uint(ms[i]) == bits;
                }
            }
            // otherwise std::sort's order itself (the same decision in every lane of the pixel)
            if (!ok && gl == 0) {
                for (int f = 0; f < m; f++) {
                    ms[f] = mv[f];
                    mo[f] = mw[f];
                }
                introsort_pairs(ms, mo, m, sstk[wave][grp]);
            }
            wave_lds_sync();
            if (!ok) {
                total = tw;
                med = 0.0f;
                for (int f = 0; f < m; f++) {
                    total -= mo[f];
                    if (total < half) {
                        med = ms[f];
                        break;
                    }
                }
            }
            if (live && gl == 0 && m > 0) newv[(size_t)k * nc + ch] = med;
            wave_lds_sync();
        }
        if (live && gl == 0) has[k] = m > 0;
    }
}


// write new values into the hole pixels that received some, and sum (prev - new)^2
__global__ void k_write_back(float *__restrict__ img, int Cs, int c0, int nc, const int *__restrict__ pts, int n,
                             const float *__restrict__ newv, const uint8_t *__restrict__ has,
                             double *__restrict__ partials)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (k < n && has[k]) {
        float *px = img + (size_t)pts[k] * Cs + c0;
        for (int ch = 0; ch < nc; ch++) {
            float v = newv[(size_t)k * nc + ch];
            float d = px[ch] - v;
            acc[0] += (double)(d * d);
            px[ch] = v;
        }
    }
    block_reduce_store<1>(acc, partials, (int)gridDim.x);
}

// ---------------- NL-Poisson coefficients ----------------
// a1, a2: W*H; f1, F1, F2: nc planes of W*H (double)
// Coefficients at pixel q: gather from the centres p around q (row-major, the order in
// which the reference's scatter reaches q).  The source gradient is formed from the
// image with gradient.cpp's float forward differences (0 past the last column / row).
// The loads of CH centres of a patch row are issued together: first their
// extended-domain flags, NNF entries and confidences, then every valid centre's source
// samples (colour, right and lower neighbour); then the centres are accumulated in
// row-major order.  A row costs two dependent memory trips instead of three per centre.
// The kernel bodies below are shared by the single-image and the batched kernels; with
// two callers the compiler would otherwise keep the large ones as real calls (248 VGPRs
// and scratch for the coefficient kernels instead of 117-197 VGPRs inlined).
template <int NC, int CH>
__device__ __attribute__((always_inline)) inline void poisson_coef_rows(int qx, int qy, const float *__restrict__ img, int Cs, int c0,
                                         const uint8_t *__restrict__ ext, const int2 *__restrict__ nnf,
                                         const float *__restrict__ conf, const float *__restrict__ wl, int P,
                                         float lambda, int W, int H, double *__restrict__ a1,
                                         double *__restrict__ a2, double *__restrict__ f1, double *__restrict__ F1,
                                         double *__restrict__ F2)
{
    const int h = P / 2;
    const size_t plane = (size_t)W * H;
    const size_t q = (size_t)W * qy + qx;
    const double lam = (double)lambda;
    const double lam1 = (double)(1 - lambda);
    double A1 = 0.0, A2 = 0.0;
    double v1[NC], g1[NC], g2[NC];
#pragma unroll
    for (int ch = 0; ch < NC; ch++) v1[ch] = g1[ch] = g2[ch] = 0.0;
    const int x_lo = max(qx - h, 0), x_hi = min(qx + h, W - 1);
    for (int py = max(qy - h, 0); py <= min(qy + h, H - 1); py++) {
        const size_t row = (size_t)W * py;
        for (int px0 = x_lo; px0 <= x_hi; px0 += CH) {
            bool ok[CH];
            int2 nb[CH];
            float cf[CH];
            uint8_t ex[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) {   // every flag, entry and confidence of the row first
                const int px = px0 + j;
                const size_t pidx = row + (px <= x_hi ? px : x_lo);
                ex[j] = ext[pidx];
                nb[j] = nnf[pidx];
                cf[j] = conf ? conf[pidx] : 1.0f;
            }
            float v[CH][NC], gx[CH][NC], gy[CH][NC];
#pragma unroll
            for (int j = 0; j < CH; j++) {
                const int px = px0 + j;
                const int n_x = nb[j].x + qx - px, n_y = nb[j].y + qy - py;
                ok[j] = px <= x_hi && ex[j] && n_x >= 0 && n_y >= 0 && n_x < W && n_y < H;
                const bool right = n_x + 1 < W, below = n_y + 1 < H;
                const float *iv = img + (ok[j] ? (size_t)W * n_y + n_x : 0) * Cs + c0;
                // in-bounds loads for every centre (the sample itself stands in for a
                // missing neighbour)
                const size_t ro = (ok[j] && right) ? (size_t)Cs : 0, bo = (ok[j] && below) ? (size_t)W * Cs : 0;
#pragma unroll
                for (int ch = 0; ch < NC; ch++) {
                    const float c = iv[ch];
                    const float r = (ok[j] && right) ? iv[ch + ro] : 0.0f;
                    const float d = (ok[j] && below) ? iv[ch + bo] : 0.0f;
                    v[j][ch] = c;
                    gx[j][ch] = right ? r - c : 0.0f;
                    gy[j][ch] = below ? d - c : 0.0f;
                }
            }
#pragma unroll
            for (int j = 0; j < CH; j++) {
                if (!ok[j]) continue;
                const int px = px0 + j;
                const double weight = (double)(cf[j] * wl[P * (qy - py + h) + (qx - px + h)]);
                A1 += weight * lam;
                A2 += weight * lam1;
#pragma unroll
                for (int ch = 0; ch < NC; ch++) {
                    v1[ch] += weight * (double)v[j][ch] * lam;
                    g1[ch] += weight * (double)gx[j][ch] * lam1;
                    g2[ch] += weight * (double)gy[j][ch] * lam1;
                }
            }
        }
    }
    a1[q] = A1;
    a2[q] = A2;
#pragma unroll
    for (int ch = 0; ch < NC; ch++) {
        f1[ch * plane + q] = v1[ch];
        F1[ch * plane + q] = g1[ch];
        F2[ch * plane + q] = g2[ch];
    }
}

// Coefficients where the solve reads them: every hole point (blockIdx.y 0), and its left
// (1) and top (2) neighbours (the stencil's a2 and the divergence's F there).  A
// neighbour that is itself a hole point is left to its own thread; one that is the
// neighbour of two hole points is written twice with the same values.
template <int NC, int CH>
__device__ __attribute__((always_inline)) inline void poisson_coef_body(const int *__restrict__ pts, int n, const int *__restrict__ lidx,
                                         const float *__restrict__ img, int Cs, int c0,
                                         const uint8_t *__restrict__ ext, const int2 *__restrict__ nnf,
                                         const float *__restrict__ conf, const float *__restrict__ w, int P,
                                         float lambda, int W, int H, double *__restrict__ a1,
                                         double *__restrict__ a2, double *__restrict__ f1, double *__restrict__ F1,
                                         double *__restrict__ F2, int bx, int row)
{
    raise_wave_priority();
    __shared__ float wl[kMaxP * kMaxP];
    for (int i = threadIdx.x; i < P * P; i += blockDim.x) wl[i] = w[i];
    __syncthreads();
    const int k = bx * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = pts[k];
    int x = p % W, y = p / W;
    if (row == 1) {
        if (x == 0 || lidx[p - 1] >= 0) return;
        x -= 1;
    } else if (row == 2) {
        if (y == 0 || lidx[p - W] >= 0) return;
        y -= 1;
    }
    poisson_coef_rows<NC, CH>(x, y, img, Cs, c0, ext, nnf, conf, wl, P, lambda, W, H, a1, a2, f1, F1, F2);
}
template <int NC, int CH>
__global__ __launch_bounds__(kBlock) void k_poisson_coef(const int *__restrict__ pts, int n, const int *__restrict__ lidx,
                               const float *__restrict__ img, int Cs, int c0, const uint8_t *__restrict__ ext,
                               const int2 *__restrict__ nnf, const float *__restrict__ conf,
                               const float *__restrict__ w, int P, float lambda, int W, int H,
                               double *__restrict__ a1, double *__restrict__ a2, double *__restrict__ f1,
                               double *__restrict__ F1, double *__restrict__ F2)
{
    poisson_coef_body<NC, CH>(pts, n, lidx, img, Cs, c0, ext, nnf, conf, w, P, lambda, W, H, a1, a2, f1, F1, F2,
                              (int)blockIdx.x, (int)blockIdx.y);
}

// The same coefficients with 16 lanes per point (4 points per wavefront): the lanes
// gather the point's whole window at once (element e = lane + 16 s, row-major over the
// window, the scatter's order), compact the contributing centres in that order into LDS
// (weight c(p) g, colour and forward differences of the source sample), and then lane j
// forms accumulator j (a1, a2, f1[c], F1[c], F2[c]) by adding its terms in that order:
// the FP64 sums of poisson_coef_rows, with one dependent memory trip per window instead
// of two per patch row.  Grid row 0: hole points, 1 and 2: their left and top
// neighbours that are not hole points.
template <int NC, int SLOTS>
__device__ __attribute__((always_inline)) inline void poisson_coef16_body(const int *__restrict__ pts, int n, const int *__restrict__ lidx,
                                           const float *__restrict__ img, int Cs, int c0,
                                           const uint8_t *__restrict__ ext, const int2 *__restrict__ nnf,
                                           const float *__restrict__ conf, const float *__restrict__ w, int P,
                                           float lambda, int W, int H, double *__restrict__ a1,
                                           double *__restrict__ a2, double *__restrict__ f1,
                                           double *__restrict__ F1, double *__restrict__ F2, int bx, int row)
{
    raise_wave_priority();
    static_assert(2 + 3 * NC <= kNLG, "one lane per accumulator");
    constexpr int kE = kNLG * SLOTS;
    __shared__ float wl[kMaxP * kMaxP];
    __shared__ float lw[kNLPix][kE];
    __shared__ float lv[kNLPix][3 * NC][kE];   // colour, x difference, y difference per channel
    for (int i = threadIdx.x; i < P * P; i += blockDim.x) wl[i] = w[i];
    __syncthreads();
    const int lane = (int)threadIdx.x, gl = lane % kNLG, grp = lane / kNLG;
    const int k = bx * kNLPix + grp;
    bool live = k < n;
    int qx = 0, qy = 0;
    if (live) {
        const int p = pts[k];
        qx = p % W;
        qy = p / W;
        if (row == 1) {
            live = qx > 0 && lidx[p - 1] < 0;
            qx -= 1;
        } else if (row == 2) {
            live = qy > 0 && lidx[p - W] < 0;
            qy -= 1;
        }
    }
    const int h = P / 2;
    const int x_lo = max(qx - h, 0), x_hi = min(qx + h, W - 1);
    const int y_lo = max(qy - h, 0), y_hi = min(qy + h, H - 1);
    const int nw = x_hi - x_lo + 1, cnt = live ? nw * (y_hi - y_lo + 1) : 0;
    // every slot's centre data first, then every slot's source samples: two memory trips
    bool ok[SLOTS];
    int2 nb[SLOTS];
    float cf[SLOTS];
    int wi[SLOTS];
    uint8_t ex[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int e = gl + kNLG * s;
        const bool in = e < cnt;
        const int py = y_lo + (in ? e / nw : 0), px = x_lo + (in ? e % nw : 0);
        const size_t pidx = (size_t)W * py + px;
        ex[s] = ext[pidx];
        ok[s] = in;
        nb[s] = nnf[pidx];
        cf[s] = conf ? conf[pidx] : 1.0f;
        nb[s].x += qx - px;
        nb[s].y += qy - py;
        wi[s] = P * (qy - py + h) + (qx - px + h);
    }
    float v[SLOTS][3 * NC];
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const int n_x = nb[s].x, n_y = nb[s].y;
        ok[s] = ok[s] && ex[s] && n_x >= 0 && n_y >= 0 && n_x < W && n_y < H;
        const bool right = n_x + 1 < W, below = n_y + 1 < H;
        const float *iv = img + (ok[s] ? (size_t)W * n_y + n_x : 0) * Cs + c0;
        // in-bounds loads for every slot (the sample itself stands in for a missing neighbour)
        const size_t ro = (ok[s] && right) ? (size_t)Cs : 0, bo = (ok[s] && below) ? (size_t)W * Cs : 0;
#pragma unroll
        for (int ch = 0; ch < NC; ch++) {
            const float c = iv[ch];
            const float r = (ok[s] && right) ? iv[ch + ro] : 0.0f;
            const float d = (ok[s] && below) ? iv[ch + bo] : 0.0f;
            v[s][ch] = c;
            v[s][NC + ch] = right ? r - c : 0.0f;
            v[s][2 * NC + ch] = below ? d - c : 0.0f;
        }
    }
    int m = 0;
#pragma unroll
    for (int s = 0; s < SLOTS; s++) {
        const uint32_t gb = (uint32_t)(__ballot(ok[s]) >> (kNLG * grp)) & ((1u << kNLG) - 1);
        const int pos = m + __popc(gb & ((1u << gl) - 1));
        m += __popc(gb);
        if (ok[s]) {
            lw[grp][pos] = cf[s] * wl[wi[s]];
#pragma unroll
            for (int j = 0; j < 3 * NC; j++) lv[grp][j][pos] = v[s][j];
        }
    }
    wave_lds_sync();
    if (!live || gl >= 2 + 3 * NC) return;
    const double lam = (double)lambda;
    const double lam1 = (double)(1 - lambda);
    double acc = 0.0;
    if (gl < 2) {
        const double f = gl == 0 ? lam : lam1;
        for (int i = 0; i < m; i++) acc += (double)lw[grp][i] * f;
    } else {
        const int j = gl - 2;                    // colour c, then x and y differences
        const double f = j < NC ? lam : lam1;
        const float *vals = lv[grp][j];
        for (int i = 0; i < m; i++) acc += (double)lw[grp][i] * (double)vals[i] * f;
    }
    const size_t plane = (size_t)W * H;
    const size_t q = (size_t)W * qy + qx;
    if (gl == 0) a1[q] = acc;
    else if (gl == 1) a2[q] = acc;
    else if (gl < 2 + NC) f1[(gl - 2) * plane + q] = acc;
    else if (gl < 2 + 2 * NC) F1[(gl - 2 - NC) * plane + q] = acc;
    else F2[(gl - 2 - 2 * NC) * plane + q] = acc;
}
template <int NC, int SLOTS>
__global__ __launch_bounds__(64) void k_poisson_coef16(const int *__restrict__ pts, int n,
                                                       const int *__restrict__ lidx, const float *__restrict__ img,
                                                       int Cs, int c0, const uint8_t *__restrict__ ext,
                                                       const int2 *__restrict__ nnf, const float *__restrict__ conf,
                                                       const float *__restrict__ w, int P, float lambda, int W, int H,
                                                       double *__restrict__ a1, double *__restrict__ a2,
                                                       double *__restrict__ f1, double *__restrict__ F1,
                                                       double *__restrict__ F2)
{
    poisson_coef16_body<NC, SLOTS>(pts, n, lidx, img, Cs, c0, ext, nnf, conf, w, P, lambda, W, H, a1, a2, f1, F1,
                                   F2, (int)blockIdx.x, (int)blockIdx.y);
}

template <int NC>
void launch_poisson_coef(int P, dim3 grid, hipStream_t s, const int *pts, int n, const int *lidx, const float *img,
                         int Cs, int c0, const uint8_t *ext, const int2 *nnf, const float *conf, const float *w,
                         float lambda, int W, int H, double *a1, double *a2, double *f1, double *F1, double *F2)
{
    // Up to 32 k points the 16-lanes-per-point kernel finishes sooner (two memory trips
    // per point); above that its LDS (one window per 16 lanes) limits the points in
    // flight, and the thread-per-point kernel is faster.  Both give identical values.
    if (n < 32768) {
        const dim3 g16((unsigned)((n + kNLPix - 1) / kNLPix), 3);
#define NLI_COEF16(S)                                                                                           \
    k_poisson_coef16<NC, S><<<g16, 64, 0, s>>>(pts, n, lidx, img, Cs, c0, ext, nnf, conf, w, P, lambda, W, H, a1, \
                                               a2, f1, F1, F2)
        const int pp = P * P;
        if (pp <= 32) NLI_COEF16(2);
        else if (pp <= 64) NLI_COEF16(4);
        else if (pp <= 96) NLI_COEF16(6);
        else if (pp <= 128) NLI_COEF16(8);
        else NLI_COEF16(15);
#undef NLI_COEF16
        return;
    }
#define NLI_COEF(CH)                                                                                             \
    k_poisson_coef<NC, CH><<<grid, kBlock, 0, s>>>(pts, n, lidx, img, Cs, c0, ext, nnf, conf, w, P, lambda, W, H, \
                                                   a1, a2, f1, F1, F2)
    switch (P) {
    case 5: NLI_COEF(5); break;
    case 7: NLI_COEF(7); break;
    case 9: NLI_COEF(9); break;
    case 11: NLI_COEF(11); break;
    default: NLI_COEF(4); break;
    }
#undef NLI_COEF
}

// ---------------- CG set-up ----------------
// stencil per hole point: a_t, a_l, a_b, a_r, sum, a1  (reference coefficient choice:
// top/left use a2 of that neighbour, bottom/right a2 of the centre; 0 outside the image)
struct Stencil {
    double at, al, ab, ar, sum, a1;
};

// list index of every hole pixel
__global__ void k_lidx(const int *pts, int n, int *lidx)
{
    raise_wave_priority();
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) lidx[pts[k]] = k;
}

// anisotropic Laplacian of a list-space vector (0 outside the hole) at point k
__device__ inline double lap(const Stencil &s, const int4 &nb, const double *v, double vc)
{
    double ut = nb.x >= 0 ? v[nb.x] : 0.0;
    double ul = nb.y >= 0 ? v[nb.y] : 0.0;
    double ub = nb.z >= 0 ? v[nb.z] : 0.0;
    double ur = nb.w >= 0 ? v[nb.w] : 0.0;
    return ut * s.at + ul * s.al + ub * s.ab + ur * s.ar - vc * s.sum;
}


// The solve's set-up at hole point k, in one pass: the backward divergence f2 of (F1, F2)
// at the point (patch_non_local_poisson.cpp:235-285, with its boundary cases), the
// stencil and neighbour indices (reference coefficient choice: top/left use a2 of that
// neighbour, bottom/right a2 of the centre; 0 outside the image), the right-hand side
// b = f1 - f2 - a1 u0 + L(u0) with u0 over the full image, and x0 = image - u0.
__device__ __attribute__((always_inline)) inline void cg_prepare_body(const int *__restrict__ pts, int n, const int *__restrict__ lidx,
                                       const float *__restrict__ img, const float *__restrict__ orig, int Cs, int c0,
                                       int nc, const double *__restrict__ a1, const double *__restrict__ a2,
                                       const double *__restrict__ f1, const double *__restrict__ F1,
                                       const double *__restrict__ F2, int div_active, int W, int H,
                                       Stencil *__restrict__ st, int4 *__restrict__ nbr, double *__restrict__ b,
                                       double *__restrict__ x0, int *__restrict__ done_at, int blk)
{
    raise_wave_priority();
    const int k = blk * blockDim.x + threadIdx.x;
    if (blk == 0 && threadIdx.x < kMaxNC) done_at[threadIdx.x] = -1;   // the CG's stop flags
    if (k >= n) return;
    const int p = pts[k];
    const int x = p % W, y = p / W;
    const size_t plane = (size_t)W * H;
    Stencil s;
    int4 nb;
    if (y != 0) { s.at = a2[p - W]; nb.x = lidx[p - W]; } else { s.at = 0; nb.x = -1; }
    if (x != 0) { s.al = a2[p - 1]; nb.y = lidx[p - 1]; } else { s.al = 0; nb.y = -1; }
    if (y != H - 1) { s.ab = a2[p]; nb.z = lidx[p + W]; } else { s.ab = 0; nb.z = -1; }
    if (x != W - 1) { s.ar = a2[p]; nb.w = lidx[p + 1]; } else { s.ar = 0; nb.w = -1; }
    s.sum = s.at + s.al + s.ab + s.ar;
    s.a1 = a1[p];
    st[k] = s;
    nbr[k] = nb;
    for (int ch = 0; ch < nc; ch++) {
        const double *fx = F1 + ch * plane, *fy = F2 + ch * plane;
        double dv = 0.0;
        if (div_active) {
            if (x == 0) dv = fx[p];
            else if (x == W - 1) dv = -fx[p - 1];
            else dv = fx[p] - fx[p - 1];
            if (y == 0) dv += fy[p];
            else if (y == H - 1) dv += -fy[p - W];
            else dv += fy[p] - fy[p - W];
        }
        auto U = [&](int q) { return (double)orig[(size_t)q * Cs + c0 + ch]; };
        const double ut = (y != 0) ? U(p - W) : 0.0;
        const double ul = (x != 0) ? U(p - 1) : 0.0;
        const double ub = (y != H - 1) ? U(p + W) : 0.0;
        const double ur = (x != W - 1) ? U(p + 1) : 0.0;
        const double u0 = U(p);
        const double L = ut * s.at + ul * s.al + ub * s.ab + ur * s.ar - u0 * (s.at + s.al + s.ab + s.ar);
        b[(size_t)ch * n + k] = f1[ch * plane + p] + (-dv - s.a1 * u0 + L);
        x0[(size_t)ch * n + k] = (double)img[(size_t)p * Cs + c0 + ch] - u0;
    }
}
__global__ void k_cg_prepare(const int *__restrict__ pts, int n, const int *__restrict__ lidx,
                             const float *__restrict__ img, const float *__restrict__ orig, int Cs, int c0, int nc,
                             const double *__restrict__ a1, const double *__restrict__ a2,
                             const double *__restrict__ f1, const double *__restrict__ F1,
                             const double *__restrict__ F2, int div_active, int W, int H,
                             Stencil *__restrict__ st, int4 *__restrict__ nbr, double *__restrict__ b,
                             double *__restrict__ x0, int *__restrict__ done_at)
{
    cg_prepare_body(pts, n, lidx, img, orig, Cs, c0, nc, a1, a2, f1, F1, F2, div_active, W, H, st, nbr, b, x0,
                    done_at, (int)blockIdx.x);
}

// ---------------- CG ----------------
// Device state per channel c:
//   hist[c * hist_len + it] = |r_it|^2, written by block 0 of phase A of iteration it
//   done_at[c]  = the iteration count at which channel c stopped (-1 while running),
//                 written by block 0 of the phase A that detects the stop
//   prr, prq    = block partial sums of r.r and r.(A d) ([c * nb + block])
// Every kernel first sums the previous kernel's partials itself, in block order, so
// all blocks see identical scalars and no grid-wide fence or atomic is needed.  No
// kernel writes the partial array it reads (phase A reads prr and writes prq, phase B
// the reverse): a block that starts late must still find the previous kernel's values.
template <typename T>
struct CGArgs {
    const Stencil *st;
    const int4 *nbr;
    const double *b;   // right-hand side (FP64)
    int n, nb, hist_len;
    T *x[2];
    T *d[2];
    T *r;
    double *prr, *prq;
    double *hist;
    int *done_at;
    int *poll_host;    // pinned host slots [2][kPollSlot]: stop state after every poll_every iterations
    int poll_every;
    unsigned poll_base;   // stamp of this solve's first poll (chunk k is announced as poll_base + k)
    float tol;
    int limit;
    int refine_every;  // T = float: residual recomputed in FP64 every k iterations (0 = never)
    // batched launches: the FP64 initial guess, and the image's flag (0: its stopping rule
    // ended its iterations; the kernels then only announce its polls as finished)
    const double *x0d;
    const int *active;
};

// Grid totals of the previous kernel's block partials part[c * nb + block]; every
// thread of every block forms the same sums in the same order.  The first 4 partials of
// each thread and channel are loaded together, in bounds (partial 0 stands in for those
// past nb and is then discarded), so that they travel with the caller's other loads in
// one memory round trip; grids of more than 4 * kBlock blocks add the rest afterwards.
template <int NC>
__device__ inline void load_partials(const double *part, int nb, double (&tot)[NC])
{
    double v[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int b = (int)threadIdx.x + j * kBlock;
            v[c][j] = part[(size_t)c * nb + (b < nb ? b : 0)];
        }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        double u[4];
#pragma unroll
        for (int j = 0; j < 4; j++) u[j] = (int)threadIdx.x + j * kBlock < nb ? v[c][j] : 0.0;
        tot[c] = ((u[0] + u[1]) + u[2]) + u[3];
    }
    if (nb > 4 * kBlock) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const double *pc = part + (size_t)c * nb;
            double t = tot[c];
            for (int b = (int)threadIdx.x + 4 * kBlock; b < nb; b += kBlock) t += pc[b];
            tot[c] = t;
        }
    }
}

template <typename T>
struct StencilOf {
    T at, al, ab, ar, sum, a1;
    __device__ explicit StencilOf(const Stencil &s)
        : at((T)s.at), al((T)s.al), ab((T)s.ab), ar((T)s.ar), sum((T)s.sum), a1((T)s.a1) {}
};

// r0 = b - A x0 (FP64), d0 = r0, partials of |r0|^2
template <int NC, typename T>
__device__ __attribute__((always_inline)) inline void cg_init_body(const CGArgs<T> &a, const double *x0, int blk)
{
    raise_wave_priority();
    int k = blk * blockDim.x + threadIdx.x;
    double acc[NC];
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    if (k < a.n) {
        Stencil s = a.st[k];
        int4 nb = a.nbr[k];
        for (int c = 0; c < NC; c++) {
            size_t o = (size_t)c * a.n;
            const double *xv = x0 + o;
            double xc = xv[k];
            double rv = a.b[o + k] - s.a1 * xc + lap(s, nb, xv, xc);
            a.x[0][o + k] = (T)xc;
            a.r[o + k] = (T)rv;
            a.d[0][o + k] = (T)rv;
            acc[c] = rv * rv;
        }
    }
    block_reduce_store_at<NC>(acc, a.prr, a.nb, blk);
}
template <int NC, typename T>
__global__ void k_cg_init(CGArgs<T> a, const double *x0)
{
    cg_init_body<NC, T>(a, x0, (int)blockIdx.x);
}

// The per-channel scalars of an iteration (the stop test on |r_it|^2, beta, alpha, and
// whether a channel still runs) are formed identically by every thread of every block.
// Block 0 of k_cg_a records |r_it|^2 and the stop; a block that reads done_at after
// block 0 set it in the same kernel sees the channel stopped, which is what its own
// stop test says.

// Loads of one unknown and its four stencil neighbours from two vectors (u, v), all
// issued together (a missing neighbour reads the unknown itself and is then zeroed), before
// the grid totals are formed so that the memory round trips overlap.
template <int NC, typename T>
struct Gather5 {
    T u[NC][5], v[NC][5];
    __device__ void load(const T *uv, const T *vv, int n, int k, const int4 &nb, bool with_u)
    {
        const int idx[5] = {k, nb.x, nb.y, nb.z, nb.w};
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const size_t o = (size_t)c * n;
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const size_t q = o + (idx[j] >= 0 ? idx[j] : k);
                v[c][j] = vv[q];
                u[c][j] = with_u ? uv[q] : (T)0;   // with_u is uniform
            }
        }
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int j = 1; j < 5; j++)
                if (idx[j] < 0) {
                    u[c][j] = (T)0;
                    v[c][j] = (T)0;
                }
    }
};

// phase A of iteration it: stop test on |r_it|^2, d_it = r_it + beta d_{it-1} (it > 0),
// partials of <r_it, A d_it>
template <int NC, typename T>
__device__ __attribute__((always_inline)) inline void cg_a_body(const CGArgs<T> &a, int it, int blk)
{
    raise_wave_priority();
    const int k = blk * blockDim.x + threadIdx.x;
    const bool live = k < a.n;
    const T *dp = a.d[it & 1];
    T *dn = a.d[(it + 1) & 1];
    // first memory round trip: this unknown's neighbours and stencil, the partials and the
    // CG history (in-bounds indices everywhere, values of the stand-ins discarded)
    const int kc = live ? k : 0;
    const int4 nb0 = a.nbr[kc];
    const Stencil s64 = a.st[kc];
    const int4 nb = live ? nb0 : make_int4(-1, -1, -1, -1);
    int done[NC];
    double h0[NC], hprev[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        done[c] = a.done_at[c];
        const double *h = a.hist + (size_t)c * a.hist_len;
        h0[c] = h[0];
        hprev[c] = h[it > 0 ? it - 1 : 0];
    }
    double tot[NC];
    load_partials<NC>(a.prr, a.nb, tot);
    // second: the neighbours' values
    Gather5<NC, T> g;   // u = r, v = d_{it-1}
    if (live) g.load(a.r, dp, a.n, k, nb, it > 0);
    block_sum<NC>(tot);

    bool any = false;
    T beta[NC];
    bool run[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        run[c] = done[c] < 0;
        beta[c] = (T)0;
        if (run[c]) {
            const double nr = tot[c];
            const double nr0 = (it == 0) ? nr : h0[c];
            const bool stop = (it == 0) ? !(nr0 > 0) : ((nr < (double)a.tol * nr0) || (it - 1 >= a.limit));
            if (blk == 0 && threadIdx.x == 0) {
                a.hist[(size_t)c * a.hist_len + it] = nr;
                if (stop) a.done_at[c] = it;
            }
            beta[c] = (it == 0 || stop) ? (T)0 : (T)(nr / hprev[c]);
            run[c] = !stop;
        }
        any = any || run[c];
    }
    // the stop state for the host's convergence poll, written straight to pinned host
    // memory (no copy in the stream); slot parity follows the poll count
    if ((it + 1) % a.poll_every == 0 && blk == 0 && threadIdx.x == 0 && a.poll_host) {
        const int chunk = (it + 1) / a.poll_every - 1;
        int *slot = a.poll_host + (chunk & 1) * kPollSlot;
#pragma unroll
        for (int c = 0; c < NC; c++) slot[c] = run[c] ? -1 : (done[c] >= 0 ? done[c] : it);
        __threadfence_system();   // the flags reach host memory before the stamp that announces them
        reinterpret_cast<volatile unsigned *>(slot)[kMaxNC] = a.poll_base + (unsigned)chunk;
    }
    if (!any) return;   // uniform over the grid
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    if (live) {
        const StencilOf<T> s(s64);
        const bool has[5] = {true, nb.x >= 0, nb.y >= 0, nb.z >= 0, nb.w >= 0};
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (!run[c]) continue;
            T D[5];
#pragma unroll
            for (int j = 0; j < 5; j++)
                D[j] = !has[j] ? (T)0 : (it == 0 ? g.v[c][j] : g.u[c][j] + beta[c] * g.v[c][j]);
            const T L = D[1] * s.at + D[2] * s.al + D[3] * s.ab + D[4] * s.ar - D[0] * s.sum;
            dn[(size_t)c * a.n + k] = D[0];
            const T rk = it == 0 ? g.v[c][0] : g.u[c][0];   // r_0 = d_0 (same bits)
            acc[c] = (double)rk * (double)(s.a1 * D[0] - L);
        }
    }
    block_reduce_store_at<NC>(acc, a.prq, a.nb, blk);
}
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_a(CGArgs<T> a, int it)
{
    cg_a_body<NC, T>(a, it, (int)blockIdx.x);
}

// phase B of iteration it: alpha = |r_it|^2 / <r_it, A d_it>, x_{it+1} = x_it + alpha d_it,
// r_{it+1} = b - A x_{it+1}, partials of |r_{it+1}|^2
template <int NC, typename T>
__device__ __attribute__((always_inline)) inline void cg_b_body(const CGArgs<T> &a, int it, int blk)
{
    raise_wave_priority();
    const int k = blk * blockDim.x + threadIdx.x;
    const bool live = k < a.n;
    const T *xp = a.x[it & 1];
    T *xn = a.x[(it + 1) & 1];
    const T *dcur = a.d[(it + 1) & 1];
    // first memory round trip: neighbours, stencil, right-hand side, partials, history
    const int kc = live ? k : 0;
    const int4 nb0 = a.nbr[kc];
    const Stencil s64 = a.st[kc];
    const int4 nb = live ? nb0 : make_int4(-1, -1, -1, -1);
    int done[NC];
    double hit[NC], bk[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        done[c] = a.done_at[c];
        hit[c] = a.hist[(size_t)c * a.hist_len + it];
        bk[c] = a.b[(size_t)c * a.n + kc];   // (unused unless live)
    }
    double tot[NC];
    load_partials<NC>(a.prq, a.nb, tot);
    Gather5<NC, T> g;   // u = x_it, v = d_it
    if (live) g.load(xp, dcur, a.n, k, nb, true);
    block_sum<NC>(tot);

    bool any = false;
    bool run[NC];
    T alpha[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        run[c] = done[c] < 0;
        alpha[c] = run[c] ? (T)(hit[c] / tot[c]) : (T)0;
        any = any || run[c];
    }
    if (!any) return;   // uniform over the grid
    const bool refine = std::is_same<T, float>::value && a.refine_every > 0 && ((it + 1) % a.refine_every == 0);
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    if (live) {
        const StencilOf<T> s(s64);
        const bool has[5] = {true, nb.x >= 0, nb.y >= 0, nb.z >= 0, nb.w >= 0};
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (!run[c]) continue;
            T X[5];
#pragma unroll
            for (int j = 0; j < 5; j++) X[j] = has[j] ? g.u[c][j] + alpha[c] * g.v[c][j] : (T)0;
            double rv;
            if (refine) {   // residual replacement: b - A x evaluated in FP64
                const double L = (double)X[1] * s64.at + (double)X[2] * s64.al + (double)X[3] * s64.ab +
                                 (double)X[4] * s64.ar - (double)X[0] * s64.sum;
                rv = bk[c] - s64.a1 * (double)X[0] + L;
            } else {
                const T L = X[1] * s.at + X[2] * s.al + X[3] * s.ab + X[4] * s.ar - X[0] * s.sum;
                rv = (double)((T)bk[c] - s.a1 * X[0] + L);
            }
            xn[(size_t)c * a.n + k] = X[0];
            a.r[(size_t)c * a.n + k] = (T)rv;
            acc[c] = rv * rv;
        }
    }
    block_reduce_store_at<NC>(acc, a.prr, a.nb, blk);
}
template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_b(CGArgs<T> a, int it)
{
    cg_b_body<NC, T>(a, it, (int)blockIdx.x);
}

// final colours: image = float(u0 + x_final) on hole points; sum of (prev - new)^2
template <typename T>
__device__ __attribute__((always_inline)) inline void poisson_write_body(float *__restrict__ img, const float *__restrict__ orig, int Cs, int c0,
                                          int nc, const int *__restrict__ pts, int n, const T *x0, const T *x1,
                                          const int *__restrict__ done_at, double *__restrict__ partials, int nb,
                                          int blk)
{
    raise_wave_priority();
    int k = blk * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (k < n) {
        size_t p = (size_t)pts[k];
        for (int ch = 0; ch < nc; ch++) {
            const T *xf = (done_at[ch] & 1) ? x1 : x0;
            double u0 = (double)orig[p * Cs + c0 + ch];
            float v = (float)(u0 + (double)xf[(size_t)ch * n + k]);
            float prev = img[p * Cs + c0 + ch];
            float d = prev - v;
            acc[0] += (double)(d * d);
            img[p * Cs + c0 + ch] = v;
        }
    }
    block_reduce_store_at<1>(acc, partials, nb, blk);
}
template <typename T>
__global__ void k_poisson_write(float *__restrict__ img, const float *__restrict__ orig, int Cs, int c0, int nc,
                                const int *__restrict__ pts, int n, T *x0, T *x1, const int *__restrict__ done_at,
                                double *__restrict__ partials)
{
    poisson_write_body<T>(img, orig, Cs, c0, nc, pts, n, x0, x1, done_at, partials, (int)gridDim.x,
                          (int)blockIdx.x);
}

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void upload_weights(UpdWork &w, int P, float sigma, hipStream_t s)
{
    if (w.w_patch == P && w.w_sigma == sigma) return;
    std::vector<float> wh((size_t)P * P);
    gaussian_weights(P, P, sigma, sigma, wh.data());
    w.weights.ensure(wh.size());
    NLI_HIP(hipMemcpyAsync(w.weights.p, wh.data(), wh.size() * sizeof(float), hipMemcpyHostToDevice, s));
    NLI_HIP(hipStreamSynchronize(s));
    w.w_patch = P;
    w.w_sigma = sigma;
}

// Waits until *st (pinned host memory a kernel writes after its data) holds `stamp`.  If
// it has not come after two seconds, the stream is drained and it must be there then.
inline void wait_stamp(const volatile unsigned *st, unsigned stamp, hipStream_t s, const char *what)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spins = 1; *st != stamp; spins++) {
        __builtin_ia32_pause();
        if ((spins & 4095) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
            NLI_HIP(hipStreamSynchronize(s));
            if (*st != stamp) throw std::runtime_error(std::string(what) + ": the device did not write its result");
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
}

// Waits until the poll slot holds the given stamp (written by k_cg_a after the flags)
// and returns the slot.
inline const volatile int *wait_poll(UpdWork &w, int slot_index, unsigned stamp, hipStream_t s)
{
    const volatile int *slot = w.done_host + slot_index * kPollSlot;
    wait_stamp(reinterpret_cast<const volatile unsigned *>(slot + kMaxNC), stamp, s, "CG poll");
    return slot;
}

// The update's summary (k_sum_partials' scalars) from pinned memory: the kernel was
// launched with sum_args(w) and stamp w.sum_seq.
struct SumArgs {
    double *host_out;
    unsigned *host_stamp;
    unsigned stamp;
};
inline SumArgs sum_args(UpdWork &w)
{
    if (!w.sum_host) {
        NLI_HIP(hipHostMalloc((void **)&w.sum_host, sizeof(double) * (kSummary + 1),
                              hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(w.sum_host, 0, sizeof(double) * (kSummary + 1));   // stamp 0: nothing written yet
        NLI_HIP(hipHostGetDevicePointer((void **)&w.sum_dev, w.sum_host, 0));
    }
    return SumArgs{w.sum_dev, reinterpret_cast<unsigned *>(w.sum_dev + kSummary), ++w.sum_seq};
}
inline const volatile double *wait_summary(UpdWork &w, const SumArgs &a, hipStream_t s)
{
    wait_stamp(reinterpret_cast<const volatile unsigned *>(w.sum_host + kSummary), a.stamp, s, "update summary");
    return w.sum_host;
}

// Runs the CG on n unknowns x NC channels; returns with x in a.x[done_at & 1].
template <int NC, typename T>
void run_cg(const CGParams &cg, int n, UpdWork &w, CGArgs<T> a, hipStream_t s)
{
    const int nb = (int)grid_for(n);
    k_cg_init<NC, T><<<nb, kBlock, 0, s>>>(a, w.x0.p);
    NLI_LAUNCH_CHECK();
    // Convergence is polled every a.poll_every iterations without draining the stream
    // and without markers in it: after each chunk, k_cg_a writes the stop flags and then
    // the chunk's stamp to pinned host memory, and the host waits for the previous
    // chunk's stamp while the current chunk runs (kernels of a converged channel return
    // at once).
    if (!a.poll_host || a.poll_every < 1) throw std::logic_error("CG poll slots not set");
    const int poll = a.poll_every;
    a.poll_base = w.poll_seq;
    w.poll_seq += (unsigned)((cg.limit + 2) / poll + 2);
    // algorithmic bytes of one iteration of the reference recurrence (two matvecs):
    // 2 x 13 words per unknown and channel
    const double it_bytes = 2.0 * 13 * sizeof(T) * NC * (double)n;
    int chunk = 0;
    for (int it = 0; it <= cg.limit + 1; it++) {
        prof::begin(TK_CG, s);
        k_cg_a<NC, T><<<nb, kBlock, 0, s>>>(a, it);
        if (it <= cg.limit) k_cg_b<NC, T><<<nb, kBlock, 0, s>>>(a, it);
        NLI_LAUNCH_CHECK();
        prof::end(TK_CG, s, it_bytes, (double)NC * n);
        if ((it + 1) % poll == 0 && it <= cg.limit) {
            // k_cg_a of this iteration writes the stop state to slot chunk & 1
            if (chunk > 0) {
                const int prev = (chunk - 1) & 1;
                const volatile int *slot = wait_poll(w, prev, a.poll_base + (unsigned)(chunk - 1), s);
                bool all = true;
                for (int c = 0; c < NC; c++) all = all && slot[c] >= 0;
                if (all) break;
            }
            chunk++;
        }
    }
}

template <typename T>
CGArgs<T> cg_args(const CGParams &cg, int n, int nc, UpdWork &w, const Stencil *st, const int4 *nbr, T *x0, T *x1,
                  T *d0, T *d1, T *r)
{
    CGArgs<T> a;
    a.st = st;
    a.nbr = nbr;
    a.b = w.b.p;
    a.n = n;
    a.nb = (int)grid_for(n);
    a.hist_len = cg.limit + 2;
    a.x[0] = x0;
    a.x[1] = x1;
    a.d[0] = d0;
    a.d[1] = d1;
    a.r = r;
    a.prr = w.partials.p;
    a.prq = w.partials.p + (size_t)kMaxNC * a.nb;
    a.hist = w.scalars.p + kSummary;
    a.done_at = w.flag.p + 2;
    a.poll_host = w.done_dev;
    a.poll_every = 2;   // measured with stamped polls: 2 (or 1) beats 4 and 8: fewer iterations run past convergence
    a.tol = cg.tol;
    a.limit = cg.limit;
    a.refine_every = cg.refine_every;
    (void)nc;
    return a;
}

double poisson(int patch, float sigma, float lambda, const CGParams &cg, float *image, const float *orig, int Cs,
               int c0, int nc, const int *hole_pts, int n, const uint8_t *ext, const int32_t *nnf,
               const float *conf, int W, int H, UpdWork &w, CGStats *stats, hipStream_t s)
{
    const size_t np = (size_t)W * H;
    upload_weights(w, patch, sigma, s);
    w.a1.ensure(np);
    w.a2.ensure(np);
    w.f1.ensure(np * nc);
    w.F1.ensure(np * nc);
    w.F2.ensure(np * nc);
    const int nb = (int)grid_for(std::max(n, 1));
    w.scalars.ensure(kSummary + (size_t)kMaxNC * (cg.limit + 2));
    if (!w.done_host) {   // CG poll slots: coherent pinned memory the kernels write directly
        NLI_HIP(hipHostMalloc((void **)&w.done_host, sizeof(int) * 2 * kPollSlot,
                              hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(w.done_host, 0, sizeof(int) * 2 * kPollSlot);   // stamp 0: no poll yet
        NLI_HIP(hipHostGetDevicePointer((void **)&w.done_dev, w.done_host, 0));
    }
    w.partials.ensure((size_t)2 * kMaxNC * nb + 8);
    w.flag.ensure(2 + kMaxNC + 2);
    if (n == 0) {
        if (stats)
            for (int c = 0; c < nc; c++) { stats->iterations[c] = 0; stats->nr0[c] = 0; stats->nr[c] = 0; }
        return 0.0;
    }
    // the list-index map is rebuilt only when the hole point list changes (points_changed)
    w.lidx.ensure(np);
    if (w.lidx_version != w.pts_version || w.lidx_of != w.lidx.p || w.lidx_pts != hole_pts || w.lidx_n != n ||
        w.lidx_np != np) {
        fill_i32(w.lidx.p, -1, np, s);
        k_lidx<<<grid_for(n), kBlock, 0, s>>>(hole_pts, n, w.lidx.p);
        NLI_LAUNCH_CHECK();
        w.lidx_version = w.pts_version;
        w.lidx_of = w.lidx.p;
        w.lidx_pts = hole_pts;
        w.lidx_n = n;
        w.lidx_np = np;
    }
    prof::begin(TK_POISSON_COEF, s);
    {
        const dim3 grid(grid_for(n), 3);
        auto coef = [&](auto nc_tag) {
            constexpr int NC = decltype(nc_tag)::value;
            launch_poisson_coef<NC>(patch, grid, s, hole_pts, n, w.lidx.p, image, Cs, c0, ext, (const int2 *)nnf,
                                    conf, w.weights.p, lambda, W, H, w.a1.p, w.a2.p, w.f1.p, w.F1.p, w.F2.p);
        };
        switch (nc) {
        case 1: coef(std::integral_constant<int, 1>{}); break;
        case 2: coef(std::integral_constant<int, 2>{}); break;
        case 3: coef(std::integral_constant<int, 3>{}); break;
        default: coef(std::integral_constant<int, 4>{}); break;
        }
    }
    NLI_LAUNCH_CHECK();
    prof::end(TK_POISSON_COEF, s, 0.0, (double)n);
    w.stencil.ensure((size_t)n * (sizeof(Stencil) / sizeof(double)));
    w.nbr.ensure((size_t)n * 4);
    Stencil *st = (Stencil *)w.stencil.p;
    int4 *nbr = (int4 *)w.nbr.p;
    w.b.ensure((size_t)n * nc);
    w.x0.ensure((size_t)n * nc);
    k_cg_prepare<<<grid_for(n), kBlock, 0, s>>>(hole_pts, n, w.lidx.p, image, orig, Cs, c0, nc, w.a1.p, w.a2.p,
                                                w.f1.p, w.F1.p, w.F2.p, lambda != 1 ? 1 : 0, W, H, st, nbr, w.b.p,
                                                w.x0.p, w.flag.p + 2);
    NLI_LAUNCH_CHECK();
    if (cg.precision == 0) {
        w.x1.ensure((size_t)n * nc);
        w.d0.ensure((size_t)n * nc);
        w.d1.ensure((size_t)n * nc);
        w.r.ensure((size_t)n * nc);
    } else {
        w.xs0.ensure((size_t)n * nc);
        w.xs1.ensure((size_t)n * nc);
        w.ds0.ensure((size_t)n * nc);
        w.ds1.ensure((size_t)n * nc);
        w.rs.ensure((size_t)n * nc);
    }
    const int *done_at = w.flag.p + 2;
    if (cg.precision == 0) {
        CGArgs<double> a = cg_args<double>(cg, n, nc, w, st, nbr, w.x0.p, w.x1.p, w.d0.p, w.d1.p, w.r.p);
        switch (nc) {
        case 1: run_cg<1, double>(cg, n, w, a, s); break;
        case 2: run_cg<2, double>(cg, n, w, a, s); break;
        case 3: run_cg<3, double>(cg, n, w, a, s); break;
        case 4: run_cg<4, double>(cg, n, w, a, s); break;
        default: throw std::invalid_argument("poisson: 1 to 4 channels");
        }
        k_poisson_write<double><<<nb, kBlock, 0, s>>>(image, orig, Cs, c0, nc, hole_pts, n, w.x0.p, w.x1.p, done_at,
                                                      w.partials.p);
    } else {
        CGArgs<float> a = cg_args<float>(cg, n, nc, w, st, nbr, w.xs0.p, w.xs1.p, w.ds0.p, w.ds1.p, w.rs.p);
        switch (nc) {
        case 1: run_cg<1, float>(cg, n, w, a, s); break;
        case 2: run_cg<2, float>(cg, n, w, a, s); break;
        case 3: run_cg<3, float>(cg, n, w, a, s); break;
        case 4: run_cg<4, float>(cg, n, w, a, s); break;
        default: throw std::invalid_argument("poisson: 1 to 4 channels");
        }
        k_poisson_write<float><<<nb, kBlock, 0, s>>>(image, orig, Cs, c0, nc, hole_pts, n, w.xs0.p, w.xs1.p, done_at,
                                                     w.partials.p);
    }
    const SumArgs sa = sum_args(w);
    k_sum_partials<<<1, kBlock, 0, s>>>(w.partials.p, nb, w.scalars.p, w.scalars.p + kSummary, cg.limit + 2,
                                        done_at, nc, sa.host_out, sa.host_stamp, sa.stamp);
    NLI_LAUNCH_CHECK();
    const volatile double *sum = wait_summary(w, sa, s);
    if (stats)
        for (int c = 0; c < nc; c++) {
            stats->iterations[c] = (int)sum[kSumDone + c];
            stats->nr0[c] = sum[kSumNr0 + c];
            stats->nr[c] = sum[kSumNr + c];
        }
    return sum[kSumTotal];
}

}  // namespace

void points_changed(UpdWork &w) { w.pts_version++; }

double update(int kind, int patch, float sigma, float lambda, const CGParams &cg, float *image, const float *orig,
              int Cs, int c0, int nc, const uint8_t *hole, const int *hole_pts, int n_hole, const uint8_t *ext,
              const int32_t *nnf, const float *conf, int W, int H, UpdWork &w, CGStats *stats, hipStream_t s)
{
    if (patch < 1 || patch > kMaxP) throw std::invalid_argument("patch side must be in [1, 15]");
    if (nc < 1 || nc > kMaxNC || c0 < 0 || c0 + nc > Cs) throw std::invalid_argument("bad channel range");
    if (kind == UPD_POISSON)
        return poisson(patch, sigma, lambda, cg, image, orig, Cs, c0, nc, hole_pts, n_hole, ext, nnf, conf, W, H, w,
                       stats, s);
    if (kind == UPD_MEANS || kind == UPD_MEDIANS) {
        upload_weights(w, patch, sigma, s);
        const int n = n_hole;
        if (n == 0) return 0.0;
        w.x1.ensure(((size_t)n * nc + 1) / 2 + 1);   // new values (floats) live in x1's storage
        float *newv = (float *)w.x1.p;
        w.pflag.ensure(n);
        prof::begin(TK_NL_GATHER, s);
        // 16 lanes per hole pixel, SLOTS window elements per lane
        const int pp = patch * patch;
        const int slots = pp <= 32 ? 2 : (pp <= 64 ? 4 : (pp <= 96 ? 6 : (pp <= 128 ? 8 : 15)));
        const size_t per_block = (size_t)kNLWaves * kNLPix;
        const unsigned blocks = (unsigned)std::min<size_t>(((size_t)n + per_block - 1) / per_block, 16384);
#define NLI_NL(KERNEL, S)                                                                                     \
    KERNEL<S><<<blocks, kBlock, 0, s>>>(image, Cs, c0, nc, hole_pts, n, hole, (const int2 *)nnf, conf, w.weights.p, \
                                        patch, W, H, newv, w.pflag.p)
        if (kind == UPD_MEANS) {
            switch (slots) {
            case 2: NLI_NL(k_nlmeans_w, 2); break;
            case 4: NLI_NL(k_nlmeans_w, 4); break;
            case 6: NLI_NL(k_nlmeans_w, 6); break;
            case 8: NLI_NL(k_nlmeans_w, 8); break;
            default: NLI_NL(k_nlmeans_w, 15); break;
            }
        } else {
            switch (slots) {
            case 2: NLI_NL(k_nlmedians_w, 2); break;
            case 4: NLI_NL(k_nlmedians_w, 4); break;
            case 6: NLI_NL(k_nlmedians_w, 6); break;
            case 8: NLI_NL(k_nlmedians_w, 8); break;
            default: NLI_NL(k_nlmedians_w, 15); break;
            }
        }
#undef NLI_NL
        NLI_LAUNCH_CHECK();
        prof::end(TK_NL_GATHER, s, 0.0, (double)n);
        const int nb = (int)grid_for(n);
        w.scalars.ensure(kSummary);
        w.partials.ensure((size_t)kMaxNC * nb + 8);
        k_write_back<<<nb, kBlock, 0, s>>>(image, Cs, c0, nc, hole_pts, n, newv, w.pflag.p, w.partials.p);
        const SumArgs sa = sum_args(w);
        k_sum_partials<<<1, kBlock, 0, s>>>(w.partials.p, nb, w.scalars.p, nullptr, 0, nullptr, 0, sa.host_out,
                                            sa.host_stamp, sa.stamp);
        NLI_LAUNCH_CHECK();
        return wait_summary(w, sa, s)[kSumTotal];
    }
    throw std::invalid_argument("unknown image updating kind");
}

}  // namespace nli

// ======================= updates in lockstep =======================
// One launch per kernel for several NL-Poisson updates: every kernel of poisson() above,
// with each block belonging to one job (start[] tables), running the same device bodies on
// the job's own buffers, so that each job's results equal a single update's.  A job whose
// image has stopped iterating (*active == 0) is skipped; its CG polls are announced as
// finished.  The last kernel applies the inpainting loop's stopping rule on the device.
namespace nli {
namespace {

template <typename T>
struct PoissonJob {
    const int *active;
    const int *pts;
    int n;
    const int *lidx;
    float *img;
    const float *orig;
    int Cs, c0, nc;
    const uint8_t *ext;
    const int2 *nnf;
    const float *conf;
    const float *w;
    int P;
    float lambda;
    int W, H;
    double *a1, *a2, *f1, *F1, *F2;
    Stencil *st;
    int4 *nbr;
    double *b, *x0;
    int *done_at;
    int div_active;
    CGArgs<T> cg;
    double *wpart;     // block partials of the total change
    double *scalars;   // the update's summary (kSummary doubles)
    StopRule stop;
};

__device__ inline bool job_active(const int *active) { return !active || *active != 0; }

template <int NC, int CH, typename T>
__global__ __launch_bounds__(kBlock) void k_poisson_coef_batch(const PoissonJob<T> *__restrict__ jobs,
                                                               const int *__restrict__ start, int n)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    if (!job_active(J.active)) return;
    const int lb = b - start[j], nbj = (J.n + kBlock - 1) / kBlock;
    poisson_coef_body<NC, CH>(J.pts, J.n, J.lidx, J.img, J.Cs, J.c0, J.ext, J.nnf, J.conf, J.w, J.P, J.lambda, J.W,
                              J.H, J.a1, J.a2, J.f1, J.F1, J.F2, lb % nbj, lb / nbj);
}

template <int NC, int SLOTS, typename T>
__global__ __launch_bounds__(64) void k_poisson_coef16_batch(const PoissonJob<T> *__restrict__ jobs,
                                                             const int *__restrict__ start, int n)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    if (!job_active(J.active)) return;
    const int lb = b - start[j], nbj = (J.n + kNLPix - 1) / kNLPix;
    poisson_coef16_body<NC, SLOTS>(J.pts, J.n, J.lidx, J.img, J.Cs, J.c0, J.ext, J.nnf, J.conf, J.w, J.P,
                                   J.lambda, J.W, J.H, J.a1, J.a2, J.f1, J.F1, J.F2, lb % nbj, lb / nbj);
}

template <typename T>
__global__ void k_cg_prepare_batch(const PoissonJob<T> *__restrict__ jobs, const int *__restrict__ start, int n)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    if (!job_active(J.active)) return;
    cg_prepare_body(J.pts, J.n, J.lidx, J.img, J.orig, J.Cs, J.c0, J.nc, J.a1, J.a2, J.f1, J.F1, J.F2, J.div_active,
                    J.W, J.H, J.st, J.nbr, J.b, J.x0, J.done_at, b - start[j]);
}

template <int NC, typename T>
__global__ void k_cg_init_batch(const PoissonJob<T> *__restrict__ jobs, const int *__restrict__ start, int n)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    if (!job_active(J.active)) return;
    cg_init_body<NC, T>(J.cg, J.cg.x0d, b - start[j]);
}

template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_a_batch(const PoissonJob<T> *__restrict__ jobs,
                                                       const int *__restrict__ start, int n, int it)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    const int lb = b - start[j];
    if (!job_active(J.active)) {
        // the image has stopped iterating: its polls report every channel finished
        const CGArgs<T> &a = J.cg;
        if (lb == 0 && threadIdx.x == 0 && a.poll_host && (it + 1) % a.poll_every == 0) {
            const int chunk = (it + 1) / a.poll_every - 1;
            int *slot = a.poll_host + (chunk & 1) * kPollSlot;
            for (int c = 0; c < kMaxNC; c++) slot[c] = 0;
            __threadfence_system();
            reinterpret_cast<volatile unsigned *>(slot)[kMaxNC] = a.poll_base + (unsigned)chunk;
        }
        return;
    }
    cg_a_body<NC, T>(J.cg, it, lb);
}

template <int NC, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_b_batch(const PoissonJob<T> *__restrict__ jobs,
                                                       const int *__restrict__ start, int n, int it)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    if (!job_active(J.active)) return;
    cg_b_body<NC, T>(J.cg, it, b - start[j]);
}

template <typename T>
__global__ void k_poisson_write_batch(const PoissonJob<T> *__restrict__ jobs, const int *__restrict__ start, int n)
{
    const int b = (int)blockIdx.x;
    if (b >= start[n]) return;
    const int j = batch_job(start, n, b);
    const PoissonJob<T> &J = jobs[j];
    if (!job_active(J.active)) return;
    poisson_write_body<T>(J.img, J.orig, J.Cs, J.c0, J.nc, J.pts, J.n, J.cg.x[0], J.cg.x[1], J.done_at, J.wpart,
                          J.cg.nb, b - start[j]);
}

// Block j: job j's total change (the partials summed as k_sum_partials sums them), its CG
// summary, and the stopping rule.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sum_batch(const PoissonJob<T> *__restrict__ jobs)
{
    raise_wave_priority();
    const PoissonJob<T> &J = jobs[blockIdx.x];
    const StopRule &R = J.stop;
    if (!job_active(J.active)) {
        if (R.host && threadIdx.x == 0) {
            reinterpret_cast<volatile int *>(R.host)[0] = 0;
            __threadfence_system();
            reinterpret_cast<volatile unsigned *>(R.host)[1] = R.stamp;
        }
        return;
    }
    __shared__ double sh[kBlock];
    double s = 0.0;
    for (int b = threadIdx.x; b < J.cg.nb; b += kBlock) s += J.wpart[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double total = sh[0];
    J.scalars[kSumTotal] = total;
    int cg_it = 0;
    for (int c = 0; c < J.nc && J.n > 0; c++) {
        const int d = J.done_at[c];
        const double *h = J.cg.hist + (size_t)c * J.cg.hist_len;
        J.scalars[kSumDone + c] = (double)d;
        J.scalars[kSumNr0 + c] = h[0];
        J.scalars[kSumNr + c] = h[d > 0 ? d : 0];
        cg_it += d;
    }
    if (R.cg_total) *R.cg_total += cg_it;
    if (!R.active) return;
    if (R.part == 1) {
        *R.partial = total;
        return;
    }
    const double t = R.part == 2 ? *R.partial + total : total;
    *R.total = t;
    const int cnt = *R.count + 1;
    *R.count = cnt;
    const int act = (t > (double)R.tol && cnt < R.iters) ? 1 : 0;
    *R.active = act;
    if (R.host) {
        reinterpret_cast<volatile int *>(R.host)[0] = act;
        __threadfence_system();
        reinterpret_cast<volatile unsigned *>(R.host)[1] = R.stamp;
    }
}

template <int NC, typename T>
void poisson_batch(const UpdJob *jobs, int n, UpdBatchWork &bw, hipStream_t s)
{
    const int P = jobs[0].patch;
    const CGParams &cg0 = jobs[0].cg;
    const size_t jb = sizeof(PoissonJob<T>) * n;
    PoissonJob<T> *J = static_cast<PoissonJob<T> *>(bw.jobs.fill(jb));
    int *st = static_cast<int *>(bw.starts.fill(sizeof(int) * 3 * (n + 1)));
    int *st_coef = st, *st_c16 = st + (n + 1), *st_nb = st + 2 * (n + 1);
    st_coef[0] = st_c16[0] = st_nb[0] = 0;
    const int poll = 2;
    const unsigned poll_base = bw.poll_seq;
    bw.poll_seq += (unsigned)((cg0.limit + 2) / poll + 2);
    for (int i = 0; i < n; i++) {
        const UpdJob &u = jobs[i];
        UpdWork &w = *u.w;
        const size_t np = (size_t)u.W * u.H;
        const int m = u.n_hole;
        upload_weights(w, u.patch, u.sigma, s);
        w.a1.ensure(np);
        w.a2.ensure(np);
        w.f1.ensure(np * NC);
        w.F1.ensure(np * NC);
        w.F2.ensure(np * NC);
        const int nb = (int)grid_for(std::max(m, 1));
        w.scalars.ensure(kSummary + (size_t)kMaxNC * (u.cg.limit + 2));
        if (!w.done_host) {
            NLI_HIP(hipHostMalloc((void **)&w.done_host, sizeof(int) * 2 * kPollSlot,
                                  hipHostMallocCoherent | hipHostMallocMapped));
            std::memset(w.done_host, 0, sizeof(int) * 2 * kPollSlot);
            NLI_HIP(hipHostGetDevicePointer((void **)&w.done_dev, w.done_host, 0));
        }
        w.partials.ensure((size_t)2 * kMaxNC * nb + 8);
        w.flag.ensure(2 + kMaxNC + 2);
        w.lidx.ensure(np);
        if (m > 0 && (w.lidx_version != w.pts_version || w.lidx_of != w.lidx.p || w.lidx_pts != u.hole_pts ||
                      w.lidx_n != m || w.lidx_np != np)) {
            fill_i32(w.lidx.p, -1, np, s);
            k_lidx<<<grid_for(m), kBlock, 0, s>>>(u.hole_pts, m, w.lidx.p);
            NLI_LAUNCH_CHECK();
            w.lidx_version = w.pts_version;
            w.lidx_of = w.lidx.p;
            w.lidx_pts = u.hole_pts;
            w.lidx_n = m;
            w.lidx_np = np;
        }
        w.stencil.ensure((size_t)std::max(m, 1) * (sizeof(Stencil) / sizeof(double)));
        w.nbr.ensure((size_t)std::max(m, 1) * 4);
        w.b.ensure((size_t)std::max(m, 1) * NC);
        w.x0.ensure((size_t)std::max(m, 1) * NC);
        CGArgs<T> a;
        const size_t vn = (size_t)std::max(m, 1) * NC;
        if (std::is_same<T, double>::value) {
            w.x1.ensure(vn);
            w.d0.ensure(vn);
            w.d1.ensure(vn);
            w.r.ensure(vn);
            a = cg_args<T>(u.cg, m, NC, w, (Stencil *)w.stencil.p, (int4 *)w.nbr.p, (T *)w.x0.p, (T *)w.x1.p,
                           (T *)w.d0.p, (T *)w.d1.p, (T *)w.r.p);
        } else {
            w.xs0.ensure(vn);
            w.xs1.ensure(vn);
            w.ds0.ensure(vn);
            w.ds1.ensure(vn);
            w.rs.ensure(vn);
            a = cg_args<T>(u.cg, m, NC, w, (Stencil *)w.stencil.p, (int4 *)w.nbr.p, (T *)w.xs0.p, (T *)w.xs1.p,
                           (T *)w.ds0.p, (T *)w.ds1.p, (T *)w.rs.p);
        }
        a.poll_base = poll_base;
        a.poll_every = poll;
        a.x0d = w.x0.p;
        a.active = u.active;
        PoissonJob<T> &q = J[i];
        q.active = u.active;
        q.pts = u.hole_pts;
        q.n = m;
        q.lidx = w.lidx.p;
        q.img = u.image;
        q.orig = u.orig;
        q.Cs = u.Cs;
        q.c0 = u.c0;
        q.nc = NC;
        q.ext = u.ext;
        q.nnf = (const int2 *)u.nnf;
        q.conf = u.conf;
        q.w = w.weights.p;
        q.P = u.patch;
        q.lambda = u.lambda;
        q.W = u.W;
        q.H = u.H;
        q.a1 = w.a1.p;
        q.a2 = w.a2.p;
        q.f1 = w.f1.p;
        q.F1 = w.F1.p;
        q.F2 = w.F2.p;
        q.st = (Stencil *)w.stencil.p;
        q.nbr = (int4 *)w.nbr.p;
        q.b = w.b.p;
        q.x0 = w.x0.p;
        q.done_at = w.flag.p + 2;
        q.div_active = u.lambda != 1 ? 1 : 0;
        q.cg = a;
        q.wpart = w.partials.p;
        q.scalars = w.scalars.p;
        q.stop = u.stop;
        const bool c16 = m < 32768;
        st_coef[i + 1] = st_coef[i] + (m > 0 && !c16 ? 3 * (int)grid_for(m) : 0);
        st_c16[i + 1] = st_c16[i] + (m > 0 && c16 ? 3 * ((m + kNLPix - 1) / kNLPix) : 0);
        st_nb[i + 1] = st_nb[i] + (m > 0 ? (int)grid_for(m) : 0);
    }
    const PoissonJob<T> *dj = static_cast<const PoissonJob<T> *>(bw.jobs.upload(jb, s));
    const int *ds = static_cast<const int *>(bw.starts.upload(sizeof(int) * 3 * (n + 1), s));
    const int *d_coef = ds, *d_c16 = ds + (n + 1), *d_nb = ds + 2 * (n + 1);
    prof::begin(TK_POISSON_COEF, s);
    if (st_c16[n] > 0) {
#define NLI_C16B(S) k_poisson_coef16_batch<NC, S, T><<<st_c16[n], 64, 0, s>>>(dj, d_c16, n)
        const int pp = P * P;
        if (pp <= 32) NLI_C16B(2);
        else if (pp <= 64) NLI_C16B(4);
        else if (pp <= 96) NLI_C16B(6);
        else if (pp <= 128) NLI_C16B(8);
        else NLI_C16B(15);
#undef NLI_C16B
    }
    if (st_coef[n] > 0) {
#define NLI_CB(CH) k_poisson_coef_batch<NC, CH, T><<<st_coef[n], kBlock, 0, s>>>(dj, d_coef, n)
        switch (P) {
        case 5: NLI_CB(5); break;
        case 7: NLI_CB(7); break;
        case 9: NLI_CB(9); break;
        case 11: NLI_CB(11); break;
        default: NLI_CB(4); break;
        }
#undef NLI_CB
    }
    NLI_LAUNCH_CHECK();
    prof::end(TK_POISSON_COEF, s, 0.0, 0.0);
    const int nbt = st_nb[n];
    if (nbt > 0) {
        k_cg_prepare_batch<T><<<nbt, kBlock, 0, s>>>(dj, d_nb, n);
        k_cg_init_batch<NC, T><<<nbt, kBlock, 0, s>>>(dj, d_nb, n);
        NLI_LAUNCH_CHECK();
        double it_bytes = 0;
        for (int i = 0; i < n; i++) it_bytes += 2.0 * 13 * sizeof(T) * NC * (double)jobs[i].n_hole;
        int chunk = 0;
        for (int it = 0; it <= cg0.limit + 1; it++) {
            prof::begin(TK_CG, s);
            k_cg_a_batch<NC, T><<<nbt, kBlock, 0, s>>>(dj, d_nb, n, it);
            if (it <= cg0.limit) k_cg_b_batch<NC, T><<<nbt, kBlock, 0, s>>>(dj, d_nb, n, it);
            NLI_LAUNCH_CHECK();
            prof::end(TK_CG, s, it_bytes, 0.0);
            if ((it + 1) % poll == 0 && it <= cg0.limit) {
                if (chunk > 0) {
                    bool all = true;
                    for (int i = 0; i < n && all; i++) {
                        if (jobs[i].n_hole == 0) continue;
                        const volatile int *slot =
                            wait_poll(*jobs[i].w, (chunk - 1) & 1, poll_base + (unsigned)(chunk - 1), s);
                        for (int c = 0; c < NC; c++) all = all && slot[c] >= 0;
                    }
                    if (all) break;
                }
                chunk++;
            }
        }
        k_poisson_write_batch<T><<<nbt, kBlock, 0, s>>>(dj, d_nb, n);
        NLI_LAUNCH_CHECK();
    }
    k_sum_batch<T><<<n, kBlock, 0, s>>>(dj);
    NLI_LAUNCH_CHECK();
}

}  // namespace

void update_batch(const UpdJob *jobs, int n, UpdBatchWork &bw, hipStream_t s)
{
    if (n <= 0) return;
    const UpdJob &j0 = jobs[0];
    for (int i = 0; i < n; i++) {
        const UpdJob &u = jobs[i];
        if (u.kind != j0.kind || u.patch != j0.patch || u.nc != j0.nc || u.cg.precision != j0.cg.precision ||
            u.cg.limit != j0.cg.limit)
            throw std::invalid_argument("update batch: jobs must share the kind, patch, channels and CG settings");
        if (u.patch < 1 || u.patch > kMaxP) throw std::invalid_argument("patch side must be in [1, 15]");
        if (u.nc < 1 || u.nc > kMaxNC || u.c0 < 0 || u.c0 + u.nc > u.Cs)
            throw std::invalid_argument("bad channel range");
    }
    if (j0.kind != UPD_POISSON) throw std::invalid_argument("update batch: NL-Poisson updates only");
    auto run = [&](auto nc_tag) {
        constexpr int NC = decltype(nc_tag)::value;
        if (j0.cg.precision == 0) poisson_batch<NC, double>(jobs, n, bw, s);
        else poisson_batch<NC, float>(jobs, n, bw, s);
    };
    switch (j0.nc) {
    case 1: run(std::integral_constant<int, 1>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 3: run(std::integral_constant<int, 3>{}); break;
    default: run(std::integral_constant<int, 4>{}); break;
    }
}

}  // namespace nli
