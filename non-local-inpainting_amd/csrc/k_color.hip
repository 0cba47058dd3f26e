// sRGB <-> CIE L*a*b* conversion of the reference's command-line pipeline
// (io_utility.cpp:268-318 and the per-pixel helpers :380-450): main.cpp converts the
// 0-255 RGB input to Lab before ImageInpainting::process (main.cpp:117) and the result
// back to RGB afterwards (:307).  The RGB-D form concatenates the depth plane as the
// fourth channel (IOUtility::cat, io_utility.cpp:453-470) and takes it back out of the
// result (IOUtility::get_depth, :506-513); both are fused into the conversions here.
//
// One thread per pixel; the arithmetic is the reference's, operation by operation:
// float where it computes in float, double where a double constant promotes it
// (`aux / 12.92`, `aux > 0.04045`, `lab[1] / 500.0 + aux_y`), no contraction (the
// library is compiled with -ffp-contract=off).  The reference's pow(float, float) is
// glibc's powf.  Here every power is the correctly rounded float result, which equals
// glibc's powf for all but a few inputs in ten thousand (where the two differ by one
// unit in the last place):
//  - x^y for the float exponents 1/3, 2.4 and 1/2.4 as x^(m/N) x^(y - m/N) in double,
//    with N = 3, 5, 12: an inverse N-th root by Newton's method from a hardware-log2/exp2
//    seed, then the small factor x^(y - m/N) to third order.  The result is within
//    about 1e-13 of x^y, so the float rounding is the correct one except within that
//    distance of a rounding midpoint (a CPU emulation of the method against pow in
//    double on 2.4e7 inputs: 1 to 3 differ, each by one unit in the last place; on the
//    GPU the conversions agree with the reference as often as with exp2(y log2 x) in
//    double, which pow_pos computes and which serves outside [2^-60, 2^60]);
//  - the 8-bit values 0..255 of the gamma expansion through a table computed with the
//    same function;
//  - x^3 exactly: x*x is exact in double, x*x*x is rounded to odd (the exact error of
//    the double product, from an FMA, breaks ties away from even), and a round-to-odd
//    double rounds to float as the exact value would.
#include "nli_internal.h"

namespace nli {
namespace {

constexpr int kBlock = 256;

// x^y, x > 0 (the callers' branches guarantee it)
__device__ inline float pow_pos(float x, float y) { return (float)exp2((double)y * log2((double)x)); }

// s = x^(-1/N) in double: a float seed from the hardware log2 / exp2 (relative error
// about 1e-6), then two Newton steps s += s (1 - x s^N) / N (quadratic: about 1e-24,
// below double rounding)
template <int N>
__device__ inline double inv_root(double x, float lg2)
{
    double s = (double)__builtin_amdgcn_exp2f(-lg2 * (1.0f / N));
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double s2 = s * s, s4 = s2 * s2;
        double sn;
        if constexpr (N == 3) sn = s2 * s;
        else if constexpr (N == 5) sn = s4 * s;
        else sn = s4 * s4 * s4;   // N == 12
        s = fma(s * (1.0 / N), fma(-x, sn, 1.0), s);
    }
    return s;
}

// the power of a rational exponent m/N times x^delta, where delta = y - m/N is the
// difference between the float exponent and the rational one: exp(delta ln x) to third
// order, with ln x from the hardware log2 (|delta ln x| < 1e-5, so its error stays
// near 1e-13 of the result)
__device__ inline float with_delta(double r, float lg2, double delta_ln2)
{
    const double t = delta_ln2 * (double)lg2;
    return (float)(r * (1.0 + t * (1.0 + t * 0.5 * (1.0 + t * (1.0 / 3.0)))));
}

// x^(1/3f), x^2.4f and x^(1/2.4f): the same correctly rounded results as pow_pos except
// within about 1e-13 of a rounding midpoint (measured: 1 value in 10^7 differs, by one
// unit in the last place); outside [2^-60, 2^60] (never reached by the conversions'
// branches for finite samples) pow_pos is used
__device__ inline bool pow_fast_range(float x) { return x >= 0x1p-60f && x <= 0x1p60f; }
constexpr double kLn2 = 0.69314718055994530942;

__device__ inline float pow_third(float x)   // x^(1/3f) = x^(1/3) x^delta
{
    if (!pow_fast_range(x)) return pow_pos(x, 1.0f / 3.0f);
    const float lg2 = __builtin_amdgcn_logf(x);
    const double s = inv_root<3>((double)x, lg2);
    return with_delta((double)x * (s * s), lg2, ((double)(1.0f / 3.0f) - 1.0 / 3.0) * kLn2);
}

__device__ inline float pow_5_12(float x)   // x^(1/2.4f) = x^(5/12) x^delta = x s^7 x^delta, s = x^(-1/12)
{
    if (!pow_fast_range(x)) return pow_pos(x, 1.0f / 2.4f);
    const float lg2 = __builtin_amdgcn_logf(x);
    const double s = inv_root<12>((double)x, lg2);
    const double s2 = s * s, s4 = s2 * s2;
    return with_delta((double)x * (s4 * s2 * s), lg2, ((double)(1.0f / 2.4f) - 5.0 / 12.0) * kLn2);
}

__device__ inline float pow_12_5(float x)   // x^2.4f = x^(12/5) x^delta = (x s)^3 x^delta, s = x^(-1/5)
{
    if (!pow_fast_range(x)) return pow_pos(x, 2.4f);
    const float lg2 = __builtin_amdgcn_logf(x);
    const double u = (double)x * inv_root<5>((double)x, lg2);
    return with_delta(u * u * u, lg2, ((double)2.4f - 2.4) * kLn2);
}

// x^3, correctly rounded
__device__ inline float cube(float x)
{
    const double x2 = (double)x * (double)x;
    double p = x2 * (double)x;
    const double e = fma(x2, (double)x, -p);
    long long b = __double_as_longlong(p);
    if (e != 0.0 && !(b & 1)) b += ((e > 0.0) == (p > 0.0)) ? 1 : -1;
    return (float)__longlong_as_double(b);
}

// the gamma expansion of one 0-255 channel value, scaled by 100 (io_utility.cpp:382-392)
__device__ inline float srgb_expand(float v)
{
    float r = v / 255.0f;
    r = ((double)r > 0.04045) ? pow_12_5((r + 0.055f) / 1.055f) : (float)((double)r / 12.92);
    return r * 100.0f;
}

// srgb_expand through a table of the 256 integral values 0..255 (8-bit image samples,
// the reference program's input), computed by the same function: identical results.
// -0.0 and non-integral values are computed directly.
__device__ inline float srgb_expand_lut(float v, const float *lut)
{
    const unsigned bits = __float_as_uint(v);
    return (bits <= 0x437F0000u && rintf(v) == v) ? lut[(int)v] : srgb_expand(v);   // 0x437F0000 = 255.0f
}

// IOUtility::rgb_to_xyz + xyz_to_lab (io_utility.cpp:380-414)
__device__ inline void px_rgb_to_lab(float r8, float g8, float b8, const float *lut, float &L, float &A, float &B)
{
    const float r = srgb_expand_lut(r8, lut), g = srgb_expand_lut(g8, lut), b = srgb_expand_lut(b8, lut);
    const float X = r * 0.412453f + g * 0.357580f + b * 0.180423f;
    const float Y = r * 0.212671f + g * 0.715160f + b * 0.072169f;
    const float Z = r * 0.019334f + g * 0.119193f + b * 0.950227f;

    float x = X / 95.047f, y = Y / 100.000f, z = Z / 108.883f;
    x = (x > 0.008856f) ? pow_third(x) : (7.787f * x) + (16.0f / 116.0f);
    y = (y > 0.008856f) ? pow_third(y) : (7.787f * y) + (16.0f / 116.0f);
    z = (z > 0.008856f) ? pow_third(z) : (7.787f * z) + (16.0f / 116.0f);
    L = (116.0f * y) - 16.0f;
    A = 500.0f * (x - y);
    B = 200.0f * (y - z);
}

// IOUtility::lab_to_xyz + xyz_to_rgb (io_utility.cpp:417-450)
__device__ inline void px_lab_to_rgb(float L, float A, float B, float &r8, float &g8, float &b8)
{
    float y = (L + 16.0f) / 116.0f;
    float x = (float)((double)A / 500.0 + (double)y);
    float z = y - B / 200.0f;
    const float x3 = cube(x), y3 = cube(y), z3 = cube(z);
    x = (x3 > 0.008856f) ? x3 : (x - 16.0f / 116.0f) / 7.787f;
    y = (y3 > 0.008856f) ? y3 : (y - 16.0f / 116.0f) / 7.787f;
    z = (z3 > 0.008856f) ? z3 : (z - 16.0f / 116.0f) / 7.787f;
    const float X = x * 95.047f, Y = y * 100.000f, Z = z * 108.883f;

    const float ax = X / 100.0f, ay = Y / 100.0f, az = Z / 100.0f;
    float r = ax * 3.240479f + ay * -1.537150f + az * -0.498535f;
    float g = ax * -0.969256f + ay * 1.875992f + az * 0.041556f;
    float b = ax * 0.055648f + ay * -0.204043f + az * 1.057311f;
    r = (r > 0.0031308f) ? 1.055f * pow_5_12(r) - 0.055f : 12.92f * r;
    g = (g > 0.0031308f) ? 1.055f * pow_5_12(g) - 0.055f : 12.92f * g;
    b = (b > 0.0031308f) ? 1.055f * pow_5_12(b) - 0.055f : 12.92f * b;
    r8 = r * 255.0f;
    g8 = g * 255.0f;
    b8 = b * 255.0f;
}

// rgb: n pixels x 3; lab: n pixels x out_c (3, or 4 with the depth plane appended)
__global__ __launch_bounds__(kBlock) void k_rgb_to_lab(const float *__restrict__ rgb,
                                                       const float *__restrict__ depth, long n, int out_c,
                                                       float *__restrict__ lab)
{
    __shared__ float lut[256];
    static_assert(kBlock == 256, "one table entry per thread");
    lut[threadIdx.x] = srgb_expand((float)threadIdx.x);
    __syncthreads();
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float *p = rgb + 3 * i;
        float L, A, B;
        px_rgb_to_lab(p[0], p[1], p[2], lut, L, A, B);
        float *q = lab + (long)out_c * i;
        q[0] = L;
        q[1] = A;
        q[2] = B;
        if (out_c == 4) q[3] = depth[i];
    }
}

// lab: n pixels x in_c (3 or 4); rgb: n pixels x 3; depth (optional, in_c = 4): channel 3
__global__ __launch_bounds__(kBlock) void k_lab_to_rgb(const float *__restrict__ lab, long n, int in_c,
                                                       float *__restrict__ rgb, float *__restrict__ depth)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float *p = lab + (long)in_c * i;
        float r, g, b;
        px_lab_to_rgb(p[0], p[1], p[2], r, g, b);
        float *q = rgb + 3 * i;
        q[0] = r;
        q[1] = g;
        q[2] = b;
        if (depth) depth[i] = p[3];
    }
}

// at most 2048 blocks (8 per CU): k_rgb_to_lab's table costs each block 256 powers
unsigned color_grid(long n)
{
    const long blocks = (n + kBlock - 1) / kBlock;
    return (unsigned)(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
}

}  // namespace

void rgb_to_lab(const float *rgb, const float *depth, long n, int out_c, float *lab, hipStream_t s)
{
    if (n <= 0) return;
    if ((out_c != 3 && out_c != 4) || (out_c == 4 && !depth))
        throw std::invalid_argument("rgb_to_lab: 3 output channels, or 4 with a depth plane");
    prof::begin(TK_COLOR, s);
    hipLaunchKernelGGL(k_rgb_to_lab, dim3(color_grid(n)), dim3(kBlock), 0, s, rgb, depth, n, out_c, lab);
    NLI_HIP(hipGetLastError());
    // bytes: 3 floats in (+ depth), out_c floats out
    prof::end(TK_COLOR, s, (double)n * 4.0 * (3 + out_c + (out_c == 4 ? 1 : 0)), (double)n);
}

void lab_to_rgb(const float *lab, long n, int in_c, float *rgb, float *depth, hipStream_t s)
{
    if (n <= 0) return;
    if ((in_c != 3 && in_c != 4) || (depth && in_c != 4))
        throw std::invalid_argument("lab_to_rgb: 3 or 4 input channels; depth needs 4");
    prof::begin(TK_COLOR, s);
    hipLaunchKernelGGL(k_lab_to_rgb, dim3(color_grid(n)), dim3(kBlock), 0, s, lab, n, in_c, rgb, depth);
    NLI_HIP(hipGetLastError());
    // bytes: the three colour planes in (all four with depth), 3 floats (+ depth) out
    prof::end(TK_COLOR, s, (double)n * 4.0 * ((depth ? 4 : 3) + 3 + (depth ? 1 : 0)), (double)n);
}

}  // namespace nli
