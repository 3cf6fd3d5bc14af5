// Far fields of a surface current on a Raviart-Thomas space: the reference's
//     potential(op, points, coeffs, X; quadstrat = SingleNumQStrat(10))
// (src/postproc.jl:138-150,174-211, farfieldlocal! :273-288) for the far-field
// operators of src/maxwell/farfield.jl:1-120:
//     MWFarField3D                     amplitude (y x (K J)) x y        :61-63
//     MWDoubleLayerFarField3D          amplitude  y x (K J)             :65-67
//     MWDoubleLayerRotatedFarField3D   amplitude  y x ((K J) x n)       :98-100
//     MWFarField3DDropConstant         coeff      (y x (K J)) x y,  K = expm1(.)  :69-75
// with K(y, x) = exp(gamma y.x) (:59) and J(x) = sum_n coeffs[n] f_n(x) the
// current at the rule point x (the rule is a Dunavant triangle rule of the
// SingleNumQStrat degree, src/postproc/farfield.jl:3-7).
//
// The reference loops point -> cell -> rule point -> shape and scatters into ff;
// here the sum is regrouped as ff(y) = amplitude I(y, S(y)) with
//     S(y) = sum over all rule points x of K(y, x) w J(x),
// which is linear in the current, so the current is formed once per rule point
// (k_ff_current) and S is a dense sum over (direction, point) pairs
// (k_ff_partial: directions on lanes, points staged through LDS and broadcast to
// every lane, chunks of points over the grid's second dimension), reduced over
// the chunks in a fixed order (k_ff_finish).  For lossless kernels the phase is
// factored per cell (k_ff_partial_cells): one sincos per (direction, cell) and
// Taylor polynomials of the small in-cell phase.  Everything is FP64.
#include <algorithm>
#include <cmath>

#include "bh_device.h"
#include "bh_internal.h"

namespace bh {

namespace {

constexpr int kFFThreads = 256;    // directions per workgroup, and points per LDS tile
constexpr int kFFPoint = 9;        // doubles per staged point: x[3], Re wJ[3], Im wJ[3]
constexpr int kFFBlocks = 4096;    // target workgroups of k_ff_partial

// One thread per (cell, rule point): the absolute position x and the weighted
// current w J(x), J = sum_r a_r (x - v_r) / (2 A) with a_r = sum over the
// element-to-dof slots (dof, c) of (cell, r) of c coeffs[dof] (the RT shape
// values of src/bases/local/rtlocal.jl and the assemblydata scatter of
// postproc.jl:200-205).  rotated: w J x n (the rotated double layer's J x n).
// relative: the position is stored as x - centroid (for k_ff_partial_cells).
__global__ void k_ff_current(const double *__restrict__ geo, const double *__restrict__ qp, int Q, int64_t nf,
                             const int32_t *__restrict__ e2d, const double *__restrict__ e2c, int kmax,
                             const double *__restrict__ coeffs, int rotated, int relative, double *__restrict__ pts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf * Q) return;
    const int64_t c = i / Q;
    const double *g = geo + c * kGeo;
    const double *q = qp + i * 4;
    const double x[3] = {q[0] + g[kGeoC], q[1] + g[kGeoC + 1], q[2] + g[kGeoC + 2]};
    const double s = q[3] / (2.0 * g[kGeoA]);
    double jr[3] = {0.0, 0.0, 0.0}, ji[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < 3; ++r) {
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < kmax; ++k) {
            const int64_t at = (c * 3 + r) * kmax + k;
            const int dof = e2d[at];
            if (dof < 0) continue;
            ar += e2c[at] * coeffs[2 * (int64_t)dof];
            ai += e2c[at] * coeffs[2 * (int64_t)dof + 1];
        }
        for (int d = 0; d < 3; ++d) {
            const double f = (x[d] - g[3 * r + d]) * s;
            jr[d] += ar * f;
            ji[d] += ai * f;
        }
    }
    if (rotated) {
        const double *nn = g + kGeoN;
        const double tr[3] = {jr[1] * nn[2] - jr[2] * nn[1], jr[2] * nn[0] - jr[0] * nn[2], jr[0] * nn[1] - jr[1] * nn[0]};
        const double ti[3] = {ji[1] * nn[2] - ji[2] * nn[1], ji[2] * nn[0] - ji[0] * nn[2], ji[0] * nn[1] - ji[1] * nn[0]};
        for (int d = 0; d < 3; ++d) { jr[d] = tr[d]; ji[d] = ti[d]; }
    }
    double *o = pts + i * kFFPoint;
    for (int d = 0; d < 3; ++d) {
        o[d] = relative ? q[d] : x[d];
        o[3 + d] = jr[d];
        o[6 + d] = ji[d];
    }
}

// Partial sums S_chunk(y) = sum over the chunk's points of K(y, x) w J(x).
// Grid: x = direction blocks of kFFThreads lanes, y = chunks of `per` points.
// Every lane takes part in the LDS staging and the barriers, also the lanes past
// the last direction.  K = exp(gamma t) or, with expm1, exp(gamma t) - 1 formed as
// (e^a - 1) - 2 e^a sin^2(b/2) + i e^a sin(b) for gamma t = a + i b, so that small
// phases keep their relative accuracy.
__global__ __launch_bounds__(kFFThreads) void k_ff_partial(const double *__restrict__ pts, int64_t M, int64_t per,
                                                           const double *__restrict__ dirs, int64_t P, double gr,
                                                           double gi, int expm1_kernel, double *__restrict__ part)
{
    __shared__ double tile[kFFThreads * kFFPoint];
    const int64_t p = (int64_t)blockIdx.x * kFFThreads + threadIdx.x;
    const bool live = p < P;
    const double y0 = live ? dirs[3 * p] : 0.0, y1 = live ? dirs[3 * p + 1] : 0.0, y2 = live ? dirs[3 * p + 2] : 0.0;
    const int64_t c0 = (int64_t)blockIdx.y * per;
    const int64_t c1 = c0 + per < M ? c0 + per : M;
    double sr[3] = {0.0, 0.0, 0.0}, si[3] = {0.0, 0.0, 0.0};
    for (int64_t t0 = c0; t0 < c1; t0 += kFFThreads) {
        const int cnt = (int)(c1 - t0 < kFFThreads ? c1 - t0 : kFFThreads);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * kFFPoint; e += kFFThreads) tile[e] = pts[t0 * kFFPoint + e];
        __syncthreads();
        if (!live) continue;
        for (int j = 0; j < cnt; ++j) {
            const double *s = tile + j * kFFPoint;
            const double t = y0 * s[0] + y1 * s[1] + y2 * s[2];
            double sn, cs;
            double kr, ki;
            if (!expm1_kernel) {
                sincos(gi * t, &sn, &cs);
                if (gr != 0.0) {
                    const double ea = exp(gr * t);
                    kr = ea * cs;
                    ki = ea * sn;
                } else {
                    kr = cs;
                    ki = sn;
                }
            } else {
                const double b = gi * t;
                const double h = sin(0.5 * b);
                const double em = expm1(gr * t);
                kr = em - 2.0 * (em + 1.0) * h * h;
                ki = (em + 1.0) * sin(b);
            }
            for (int d = 0; d < 3; ++d) {
                sr[d] += kr * s[3 + d] - ki * s[6 + d];
                si[d] += kr * s[6 + d] + ki * s[3 + d];
            }
        }
    }
    if (!live) return;
    double *o = part + ((int64_t)blockIdx.y * P + p) * 6;
    for (int d = 0; d < 3; ++d) {
        o[d] = sr[d];
        o[3 + d] = si[d];
    }
}

// Taylor coefficients of sin and cos in u = t^2: sin t = t sum_k s_k u^k,
// cos t = sum_k c_k u^k, s_k = (-1)^k / (2k+1)!, c_k = (-1)^k / (2k)!.
template <int K> struct TrigSeries {
    double s[K + 1], c[K + 1];
    constexpr TrigSeries() : s(), c() {
        double f = 1.0;   // (2k)!
        for (int k = 0; k <= K; ++k) {
            if (k > 0) f *= (2.0 * k - 1.0) * (2.0 * k);
            c[k] = (k % 2 ? -1.0 : 1.0) / f;
            s[k] = (k % 2 ? -1.0 : 1.0) / (f * (2.0 * k + 1.0));
        }
    }
};

// Partial sums as k_ff_partial for a lossless kernel exp(i gi y.x) (gi real), factored
// per cell: exp(i gi y.x) = exp(i gi y.c) exp(i gi y.(x - c)) with c the cell centroid.
// The first factor is one sincos per (direction, cell); the second has a phase of at
// most theta_max = |gi| max|y| max|x - c|, and its cos and sin are the Taylor
// polynomials of degree 2K and 2K+1, whose remainders theta_max^(2K+2)/(2K+2)! and
// theta_max^(2K+3)/(2K+3)! the host keeps below 1e-17 when it picks K.  Per cell:
// A = sum_q exp(i gi y.(x_q - c)) w J(x_q), then S += exp(i gi y.c) A.
// Tiles hold whole cells (CT = floor(kFFThreads / Q) cells), and chunks whole tiles.
template <int K>
__global__ __launch_bounds__(kFFThreads) void k_ff_partial_cells(const double *__restrict__ pts,
                                                                 const double *__restrict__ geo, int64_t nf, int Q,
                                                                 int64_t per_cells, const double *__restrict__ dirs,
                                                                 int64_t P, double gi, double *__restrict__ part)
{
    constexpr TrigSeries<K> ts{};
    __shared__ double tile[kFFThreads * kFFPoint];
    __shared__ double cen[kFFThreads * 3];
    const int CT = kFFThreads / Q;
    const int64_t p = (int64_t)blockIdx.x * kFFThreads + threadIdx.x;
    const bool live = p < P;
    const double y0 = live ? dirs[3 * p] : 0.0, y1 = live ? dirs[3 * p + 1] : 0.0, y2 = live ? dirs[3 * p + 2] : 0.0;
    const double g0 = gi * y0, g1 = gi * y1, g2 = gi * y2;
    const int64_t c0 = (int64_t)blockIdx.y * per_cells;
    const int64_t c1 = c0 + per_cells < nf ? c0 + per_cells : nf;
    double sr[3] = {0.0, 0.0, 0.0}, si[3] = {0.0, 0.0, 0.0};
    for (int64_t t0 = c0; t0 < c1; t0 += CT) {
        const int nc = (int)(c1 - t0 < CT ? c1 - t0 : CT);
        __syncthreads();
        for (int e = threadIdx.x; e < nc * Q * kFFPoint; e += kFFThreads) tile[e] = pts[t0 * Q * kFFPoint + e];
        for (int e = threadIdx.x; e < nc * 3; e += kFFThreads) cen[e] = geo[(t0 + e / 3) * kGeo + kGeoC + e % 3];
        __syncthreads();
        if (!live) continue;
        for (int j = 0; j < nc; ++j) {
            double ar[3] = {0.0, 0.0, 0.0}, ai[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < Q; ++q) {
                const double *s = tile + (j * Q + q) * kFFPoint;
                const double t = g0 * s[0] + g1 * s[1] + g2 * s[2];
                const double u = t * t;
                double ps = ts.s[K], pc = ts.c[K];
#pragma unroll
                for (int k = K - 1; k >= 0; --k) {
                    ps = ps * u + ts.s[k];
                    pc = pc * u + ts.c[k];
                }
                const double sn = t * ps, cs = pc;
                for (int d = 0; d < 3; ++d) {
                    ar[d] += cs * s[3 + d] - sn * s[6 + d];
                    ai[d] += cs * s[6 + d] + sn * s[3 + d];
                }
            }
            double en, ec;
            sincos(g0 * cen[3 * j] + g1 * cen[3 * j + 1] + g2 * cen[3 * j + 2], &en, &ec);
            for (int d = 0; d < 3; ++d) {
                sr[d] += ec * ar[d] - en * ai[d];
                si[d] += ec * ai[d] + en * ar[d];
            }
        }
    }
    if (!live) return;
    double *o = part + ((int64_t)blockIdx.y * P + p) * 6;
    for (int d = 0; d < 3; ++d) {
        o[d] = sr[d];
        o[3 + d] = si[d];
    }
}

// ff(y) = amplitude I(y, S), S = sum of the chunk partials in chunk order;
// I = (y x S) x y (single layer, drop-constant) or y x S (double layers; the
// rotated one's J x n was taken in k_ff_current).  Output: 3 complex per point.
__global__ void k_ff_finish(const double *__restrict__ part, int64_t nchunk, const double *__restrict__ dirs,
                            int64_t P, int transverse, double ar, double ai, double *__restrict__ ff)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t k = 0; k < nchunk; ++k) {
        const double *q = part + (k * P + p) * 6;
        for (int d = 0; d < 6; ++d) s[d] += q[d];
    }
    const double y[3] = {dirs[3 * p], dirs[3 * p + 1], dirs[3 * p + 2]};
    double v[6];
    for (int h = 0; h < 2; ++h) {   // real and imaginary parts separately (y is real)
        const double *a = s + 3 * h;
        const double c[3] = {y[1] * a[2] - y[2] * a[1], y[2] * a[0] - y[0] * a[2], y[0] * a[1] - y[1] * a[0]};
        if (transverse) {
            v[3 * h + 0] = c[1] * y[2] - c[2] * y[1];
            v[3 * h + 1] = c[2] * y[0] - c[0] * y[2];
            v[3 * h + 2] = c[0] * y[1] - c[1] * y[0];
        } else {
            for (int d = 0; d < 3; ++d) v[3 * h + d] = c[d];
        }
    }
    double *o = ff + p * 6;
    for (int d = 0; d < 3; ++d) {
        o[2 * d] = ar * v[d] - ai * v[3 + d];
        o[2 * d + 1] = ar * v[3 + d] + ai * v[d];
    }
}

void check_ff_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) BH_FAIL(BH_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
}

}  // namespace

}  // namespace bh

using namespace bh;

extern "C" int bh_farfield(bh_ctx *ctx, int kind, const double *gamma, const double *amplitude, const bh_space *X,
                           const double *coeffs, int64_t npoints, const double *points, int rule, double *ff)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && X && gamma && amplitude, "NULL argument");
    BH_REQUIRE(kind == BH_FF_SINGLELAYER || kind == BH_FF_DOUBLELAYER || kind == BH_FF_DOUBLELAYER_ROTATED ||
                   kind == BH_FF_DROPCONSTANT,
               "unknown far-field kind %d", kind);
    BH_REQUIRE(rule >= 1, "SingleNumQStrat: triangle rule degree %d must be at least 1", rule);
    if (rule > kMaxTriDegree)
        BH_FAIL(BH_ERR_UNSUPPORTED, "SingleNumQStrat: no triangle rule of degree %d on the GPU path (1..%d)", rule,
                kMaxTriDegree);
    BH_REQUIRE(npoints >= 0, "npoints < 0");
    BH_REQUIRE(std::isfinite(gamma[0]) && std::isfinite(gamma[1]), "gamma is not finite");
    if (npoints == 0) return BH_OK;
    BH_REQUIRE(points && ff, "points or ff is NULL");
    bh_space *sp = const_cast<bh_space *>(X);
    BH_REQUIRE(sp->ndofs == 0 || coeffs, "coeffs is NULL");
    Mesh &m = *sp->mesh;
    hipStream_t s = ctx->stream;
    int Q = 0;
    const double *qp = (const double *)quad_points(m, rule, &Q, false, s);
    const int64_t M = m.nf * Q;
    const int64_t P = npoints;
    // Lossless kernels (gamma purely imaginary, not the drop-constant one) take the
    // cell-factored path when the in-cell phase bound admits a Taylor degree in {4, 6, 9}.
    int K = 0;
    if (gamma[0] == 0.0 && kind != BH_FF_DROPCONSTANT && m.nf > 0 && Q <= kFFThreads) {
        double rmax = 0.0, ymax = 0.0;
        for (int64_t c = 0; c < m.nf; ++c) {
            const double *g = &m.geo[(size_t)c * kGeo];
            for (int a = 0; a < 3; ++a) {
                const double d0 = g[3 * a] - g[kGeoC], d1 = g[3 * a + 1] - g[kGeoC + 1], d2 = g[3 * a + 2] - g[kGeoC + 2];
                rmax = std::max(rmax, std::sqrt(d0 * d0 + d1 * d1 + d2 * d2));
            }
        }
        for (int64_t p = 0; p < P; ++p) {
            const double *y = points + 3 * p;
            ymax = std::max(ymax, std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]));
        }
        const double th = std::fabs(gamma[1]) * ymax * rmax * (1.0 + 1e-12);
        for (int k : {4, 6, 9}) {
            // remainder of the cos series, theta^(2k+2) / (2k+2)!, bounds the sin one's too
            double r = 1.0;
            for (int j = 1; j <= 2 * k + 2; ++j) r *= th / j;
            if (std::isfinite(th) && r <= 1e-17) { K = k; break; }
        }
    }
    const int64_t nblk = (P + kFFThreads - 1) / kFFThreads;
    const int CT = K ? kFFThreads / Q : 1;                            // cells per tile
    const int64_t tiles = K ? (m.nf + CT - 1) / CT : (M + kFFThreads - 1) / kFFThreads;
    int64_t nchunk = (kFFBlocks + nblk - 1) / nblk;
    if (nchunk > tiles) nchunk = tiles;
    if (nchunk < 1) nchunk = 1;
    // partial sums take 48 bytes per (chunk, direction); keep them under 1 GiB
    const int64_t cap = ((int64_t)1 << 30) / (48 * P);
    if (nchunk > cap) nchunk = cap < 1 ? 1 : cap;
    BH_REQUIRE(nblk <= 0x7fffffff && nchunk <= 65535, "too many far-field directions");
    const int64_t per_tiles = (tiles + nchunk - 1) / nchunk;          // whole tiles per chunk
    DevBuf<double> dcoef(2 * sp->ndofs + 2), dpts(M * kFFPoint + 1), ddir(3 * P), dpart(nchunk * P * 6), dff(6 * P);
    if (sp->ndofs) dcoef.upload(coeffs, 2 * sp->ndofs, s);
    ddir.upload(points, 3 * P, s);
    if (M > 0) {
        ProfScope prof(*ctx, "k_ff_current", (double)M);
        k_ff_current<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(m.d_geo.p, qp, Q, m.nf, sp->d_e2d.p, sp->d_e2c.p,
                                                               sp->kmax, dcoef.p,
                                                               kind == BH_FF_DOUBLELAYER_ROTATED, K != 0, dpts.p);
        check_ff_launch("k_ff_current");
    }
    if (K) {
        ProfScope prof(*ctx, "k_ff_partial_cells", (double)M * (double)P);
        const dim3 grid((unsigned)nblk, (unsigned)nchunk);
        const int64_t per_cells = per_tiles * CT;
        if (K == 4)
            k_ff_partial_cells<4><<<grid, kFFThreads, 0, s>>>(dpts.p, m.d_geo.p, m.nf, Q, per_cells, ddir.p, P,
                                                              gamma[1], dpart.p);
        else if (K == 6)
            k_ff_partial_cells<6><<<grid, kFFThreads, 0, s>>>(dpts.p, m.d_geo.p, m.nf, Q, per_cells, ddir.p, P,
                                                              gamma[1], dpart.p);
        else
            k_ff_partial_cells<9><<<grid, kFFThreads, 0, s>>>(dpts.p, m.d_geo.p, m.nf, Q, per_cells, ddir.p, P,
                                                              gamma[1], dpart.p);
        check_ff_launch("k_ff_partial_cells");
    } else {
        ProfScope prof(*ctx, "k_ff_partial", (double)M * (double)P);
        k_ff_partial<<<dim3((unsigned)nblk, (unsigned)nchunk), kFFThreads, 0, s>>>(
            dpts.p, M, per_tiles * kFFThreads, ddir.p, P, gamma[0], gamma[1], kind == BH_FF_DROPCONSTANT, dpart.p);
        check_ff_launch("k_ff_partial");
    }
    {
        ProfScope prof(*ctx, "k_ff_finish", (double)P);
        k_ff_finish<<<(unsigned)((P + 255) / 256), 256, 0, s>>>(
            dpart.p, nchunk, ddir.p, P, kind == BH_FF_SINGLELAYER || kind == BH_FF_DROPCONSTANT, amplitude[0],
            amplitude[1], dff.p);
        check_ff_launch("k_ff_finish");
    }
    BH_CHECK_HIP(hipMemcpyAsync(ff, dff.p, sizeof(double) * 6 * P, hipMemcpyDeviceToHost, s));
    BH_CHECK_HIP(hipStreamSynchronize(s));
    BH_API_END
}
