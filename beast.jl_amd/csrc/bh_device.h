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

}  // namespace bh
