// Device-side helpers shared by the assembly kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace bh {

template <typename R>
struct cx {
    R r, i;
};

template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b)
{
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kInv4Pi = 1.0 / (4.0 * 3.14159265358979323846);

// Pair classes (doublenumwiltonsauterqstrat.jl:32-69)
enum PairClass : int { kFar = 0, kNear = 1, kCV = 2, kCE = 3, kCF = 4 };
// Strategy modes
enum QsMode : int { kModeWiltonSauter = 0, kModeSauter = 1, kModeDoubleNum = 2 };

// Per-cell geometry record (doubles): v1[3] v2[3] v3[3] centroid[3] area normal[3]
constexpr int kGeo = 16;
constexpr int kGeoC = 9;
constexpr int kGeoA = 12;
constexpr int kGeoN = 13;

// Classification of one (test, trial) pair, bit-exact with the reference:
// squared distances summed left to right without contraction, hits counted on
// sqrt(d2) < 1e3*eps, near test max(dmin2*|gamma|^2, dmin2/(16*area_trial)) < 0.2*0.2.
// Returns the class (or -1 if hits > 3) and the matched vertex pairs packed as
// ti0 | ti1<<2 | ti2<<4 | sj0<<6 | sj1<<8 | sj2<<10.
__device__ __forceinline__ int classify_pair(const double *__restrict__ tv, const double *__restrict__ sv,
                                             double area_s, double k2, int mode, int *packed)
{
#pragma clang fp contract(off)
    const double dtol = 1.0e3 * 2.220446049250313e-16;
    int hits = 0, pk = 0;
    double dmin2 = 1.79769313486231570815e+308;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double dx = tv[3 * i] - sv[3 * j];
            const double dy = tv[3 * i + 1] - sv[3 * j + 1];
            const double dz = tv[3 * i + 2] - sv[3 * j + 2];
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            dmin2 = d2 < dmin2 ? d2 : dmin2;
            if (__builtin_sqrt(d2) < dtol) {
                if (hits < 3) pk |= (i << (2 * hits)) | (j << (6 + 2 * hits));
                ++hits;
            }
        }
    *packed = pk;
    if (mode == kModeDoubleNum) return kFar;   // DoubleNumQStrat: no classification
    if (hits > 3) return -1;
    if (hits == 3) return kCF;
    if (hits == 2) return kCE;
    if (hits == 1) return kCV;
    if (mode == kModeWiltonSauter) {
        const double xtol2 = 0.2 * 0.2;
        const double a = dmin2 * k2;
        const double b = dmin2 / (16.0 * area_s);
        const double m = a > b ? a : b;
        if (m < xtol2) return kNear;
    }
    return kFar;
}

// Wave-wide sum (64 lanes) of a double.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// 1/sqrt(x) in FP64 for x > 0: the hardware estimate y (relative error e below 2^-24)
// refined by one third-order step y (1 + r/2 + 3 r^2/8), r = 1 - x y^2, which leaves
// about 2.5 e^3 relative error before rounding (measured within 1 ulp, as two Newton steps).
__device__ __forceinline__ double rsqrt_f64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double r = __builtin_fma(-(x * y), y, 1.0);
    return __builtin_fma(y * r, __builtin_fma(0.375, r, 0.5), y);
}

// sin and cos of t in FP64: Cody & Waite reduction by pi/2 (q = nearest integer
// to 2t/pi; r = t - q pi/2 with pi/2 split into three doubles: the first FMA is
// exact, since its result is a multiple of 2^-52 below 1 in magnitude, and the
// split leaves q * 1e-49 unaccounted), then Taylor series of sin (to r^15) and
// cos (to r^16) on |r| <= pi/4, where the first omitted terms are below 5e-17.
__device__ __forceinline__ void sincos_f64(double t, double *sp, double *cp)
{
    constexpr double kTwoOverPi = 0.6366197723675814;          // 2/pi
    constexpr double kP1 = 1.5707963267948966;                 // pi/2 = kP1 + kP2 + kP3
    constexpr double kP2 = 6.123233995736766e-17;
    constexpr double kP3 = -1.4973849048591698e-33;
    const double q = __builtin_rint(t * kTwoOverPi);
    double r = __builtin_fma(-q, kP1, t);
    r = __builtin_fma(-q, kP2, r);
    r = __builtin_fma(-q, kP3, r);
    const double r2 = r * r;
    // (-1)^k / (2k+1)!, k = 1..7 and (-1)^k / (2k)!, k = 1..8
    double ps = -1.0 / 1307674368000.0;
    ps = __builtin_fma(ps, r2, 1.0 / 6227020800.0);
    ps = __builtin_fma(ps, r2, -1.0 / 39916800.0);
    ps = __builtin_fma(ps, r2, 1.0 / 362880.0);
    ps = __builtin_fma(ps, r2, -1.0 / 5040.0);
    ps = __builtin_fma(ps, r2, 1.0 / 120.0);
    ps = __builtin_fma(ps, r2, -1.0 / 6.0);
    const double sn = __builtin_fma(ps * r2, r, r);
    double pc = 1.0 / 20922789888000.0;
    pc = __builtin_fma(pc, r2, -1.0 / 87178291200.0);
    pc = __builtin_fma(pc, r2, 1.0 / 479001600.0);
    pc = __builtin_fma(pc, r2, -1.0 / 3628800.0);
    pc = __builtin_fma(pc, r2, 1.0 / 40320.0);
    pc = __builtin_fma(pc, r2, -1.0 / 720.0);
    pc = __builtin_fma(pc, r2, 1.0 / 24.0);
    pc = __builtin_fma(pc, r2, -0.5);
    const double cs = __builtin_fma(pc, r2, 1.0);
    // t = q pi/2 + r: rotate by the quadrant
    const int quad = (int)(q - 4.0 * __builtin_floor(q * 0.25));   // q mod 4, exact for |q| < 2^53
    const double s1 = (quad & 1) ? cs : sn, c1 = (quad & 1) ? sn : cs;
    *sp = (quad & 2) ? -s1 : s1;
    *cp = ((quad + 1) & 2) ? -c1 : c1;
}

template <typename R>
__device__ __forceinline__ void cross3(R *o, const R *a, const R *b)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename R>
__device__ __forceinline__ R dot3(const R *a, const R *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// real and imaginary parts of the dot product of a real and a complex 3-vector
template <typename R>
__device__ __forceinline__ R dot3c_r(const R *a, const cx<R> *b)
{
    return a[0] * b[0].r + a[1] * b[1].r + a[2] * b[2].r;
}

template <typename R>
__device__ __forceinline__ R dot3c_i(const R *a, const cx<R> *b)
{
    return a[0] * b[0].i + a[1] * b[1].i + a[2] * b[2].i;
}

// Request of a batched launch's item: first[r] <= item < first[r + 1] (first[0] = 0,
// first[n] = the number of items), by binary search.
__device__ __forceinline__ int batch_find(const int *__restrict__ first, int n, int item)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace bh
