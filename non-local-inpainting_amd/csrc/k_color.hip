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
//  - x^y for x > 0 and the constant exponents 2.4, 1/3 and 1/2.4 as exp2(y log2 x) in
//    double: its relative error (about 2^-50) decides the float rounding except within
//    2^-50 of a rounding midpoint;
//  - x^3 exactly: x*x is exact in double, x*x*x is rounded to odd (the exact error of
//    the double product, from an FMA, breaks ties away from even), and a round-to-odd
//    double rounds to float as the exact value would.
#include "nli_internal.h"

namespace nli {
namespace {

constexpr int kBlock = 256;

// x^y, x > 0 (the callers' branches guarantee it)
__device__ inline float pow_pos(float x, float y) { return (float)exp2((double)y * log2((double)x)); }

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

// IOUtility::rgb_to_xyz + xyz_to_lab (io_utility.cpp:380-414)
__device__ inline void px_rgb_to_lab(float r8, float g8, float b8, float &L, float &A, float &B)
{
    float r = r8 / 255.0f, g = g8 / 255.0f, b = b8 / 255.0f;
    r = ((double)r > 0.04045) ? pow_pos((r + 0.055f) / 1.055f, 2.4f) : (float)((double)r / 12.92);
    g = ((double)g > 0.04045) ? pow_pos((g + 0.055f) / 1.055f, 2.4f) : (float)((double)g / 12.92);
    b = ((double)b > 0.04045) ? pow_pos((b + 0.055f) / 1.055f, 2.4f) : (float)((double)b / 12.92);
    r *= 100.0f;
    g *= 100.0f;
    b *= 100.0f;
    const float X = r * 0.412453f + g * 0.357580f + b * 0.180423f;
    const float Y = r * 0.212671f + g * 0.715160f + b * 0.072169f;
    const float Z = r * 0.019334f + g * 0.119193f + b * 0.950227f;

    float x = X / 95.047f, y = Y / 100.000f, z = Z / 108.883f;
    x = (x > 0.008856f) ? pow_pos(x, 1.0f / 3.0f) : (7.787f * x) + (16.0f / 116.0f);
    y = (y > 0.008856f) ? pow_pos(y, 1.0f / 3.0f) : (7.787f * y) + (16.0f / 116.0f);
    z = (z > 0.008856f) ? pow_pos(z, 1.0f / 3.0f) : (7.787f * z) + (16.0f / 116.0f);
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
    r = (r > 0.0031308f) ? 1.055f * pow_pos(r, 1.0f / 2.4f) - 0.055f : 12.92f * r;
    g = (g > 0.0031308f) ? 1.055f * pow_pos(g, 1.0f / 2.4f) - 0.055f : 12.92f * g;
    b = (b > 0.0031308f) ? 1.055f * pow_pos(b, 1.0f / 2.4f) - 0.055f : 12.92f * b;
    r8 = r * 255.0f;
    g8 = g * 255.0f;
    b8 = b * 255.0f;
}

// rgb: n pixels x 3; lab: n pixels x out_c (3, or 4 with the depth plane appended)
__global__ __launch_bounds__(kBlock) void k_rgb_to_lab(const float *__restrict__ rgb,
                                                       const float *__restrict__ depth, long n, int out_c,
                                                       float *__restrict__ lab)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float *p = rgb + 3 * i;
        float L, A, B;
        px_rgb_to_lab(p[0], p[1], p[2], L, A, B);
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

unsigned color_grid(long n)
{
    const long blocks = (n + kBlock - 1) / kBlock;
    return (unsigned)(blocks < 16384 ? (blocks > 0 ? blocks : 1) : 16384);
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
