// Plane-wave right-hand side: b[m] = sum over the shapes (c, r, a) of dof m of
// a * sum_p w_p f_r(x_p) . T E(x_p), E(x) = amplitude exp(-gamma d.x) pol
// (src/excitation.jl:38-115, src/maxwell/mwexc.jl:36-42), with the traces of
// excitation.jl:168-180: TangTraceMW (n x E) x n, which tested with a tangential f
// equals f . E, and CrossTraceMW n x E, n the cell's normal.
#include <cmath>

#include "bh_device.h"
#include "bh_internal.h"

namespace bh {

struct PlaneWave {
    double d[3], pr[3], pi[3], gr, gi, ar, ai;
};

__global__ void k_planewave(const double *__restrict__ geo, const double *__restrict__ qp, int Q,
                            const int *__restrict__ ptr, const int *__restrict__ cell, const int *__restrict__ ref,
                            const double *__restrict__ coef, int ndofs, PlaneWave pw, int trace, int accumulate,
                            double *__restrict__ b)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ndofs) return;
    double sr = 0.0, si = 0.0;
    for (int k = ptr[m]; k < ptr[m + 1]; ++k) {
        const int c = cell[k], r = ref[k];
        const double *g = geo + (size_t)c * kGeo;
        const double inv2A = 1.0 / (2.0 * g[kGeoA]);
        // the polarisation as seen by f: pol itself, or n x pol for the cross trace
        double pr[3], pi[3];
        if (trace == BH_TRACE_CROSS) {
            const double *nrm = g + kGeoN;
            cross3(pr, nrm, pw.pr);
            cross3(pi, nrm, pw.pi);
        } else {
            for (int k2 = 0; k2 < 3; ++k2) { pr[k2] = pw.pr[k2]; pi[k2] = pw.pi[k2]; }
        }
        double ar = 0.0, ai = 0.0;
        for (int p = 0; p < Q; ++p) {
            const double *q = qp + ((size_t)c * Q + p) * 4;
            const double x[3] = {q[0] + g[kGeoC], q[1] + g[kGeoC + 1], q[2] + g[kGeoC + 2]};
            const double f[3] = {(x[0] - g[3 * r]) * inv2A, (x[1] - g[3 * r + 1]) * inv2A, (x[2] - g[3 * r + 2]) * inv2A};
            const double t = pw.d[0] * x[0] + pw.d[1] * x[1] + pw.d[2] * x[2];
            double s, co;
            sincos(pw.gi * t, &s, &co);
            const double ea = pw.gr == 0.0 ? 1.0 : exp(-pw.gr * t);
            // phase = amplitude * exp(-gamma t)
            const double er = ea * co, ei = -ea * s;
            const double phr = pw.ar * er - pw.ai * ei, phi = pw.ar * ei + pw.ai * er;
            const double er2 = f[0] * pr[0] + f[1] * pr[1] + f[2] * pr[2];
            const double ei2 = f[0] * pi[0] + f[1] * pi[1] + f[2] * pi[2];
            const double w = q[3];
            ar += w * (er2 * phr - ei2 * phi);
            ai += w * (er2 * phi + ei2 * phr);
        }
        sr += coef[k] * ar;
        si += coef[k] * ai;
    }
    if (accumulate) { sr += b[2 * m]; si += b[2 * m + 1]; }
    b[2 * m] = sr;
    b[2 * m + 1] = si;
}

}  // namespace bh

using namespace bh;

extern "C" int bh_assemble_planewave(bh_ctx *ctx, const bh_space *X, const double *direction, const double *pol_re,
                                     const double *pol_im, const double *gamma, const double *amplitude, int rule,
                                     int accumulate, double *b)
{
    return bh_assemble_excitation(ctx, X, BH_TRACE_TANGENTIAL, direction, pol_re, pol_im, gamma, amplitude, rule,
                                  accumulate, b);
}

extern "C" int bh_assemble_excitation(bh_ctx *ctx, const bh_space *X, int trace, const double *direction,
                                      const double *pol_re, const double *pol_im, const double *gamma,
                                      const double *amplitude, int rule, int accumulate, double *b)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && X && direction && pol_re && pol_im && gamma && amplitude, "NULL argument");
    BH_REQUIRE(trace == BH_TRACE_TANGENTIAL || trace == BH_TRACE_CROSS, "trace must be BH_TRACE_TANGENTIAL or "
               "BH_TRACE_CROSS");
    const int64_t n = X->ndofs;
    if (n == 0) return BH_OK;
    BH_REQUIRE(b != nullptr, "b is NULL");
    bh_space *sp = const_cast<bh_space *>(X);
    Mesh &m = *sp->mesh;
    int Q = 0;
    const double *qp = (const double *)quad_points(m, rule, &Q, false, ctx->stream);
    PlaneWave pw{};
    for (int k = 0; k < 3; ++k) { pw.d[k] = direction[k]; pw.pr[k] = pol_re[k]; pw.pi[k] = pol_im[k]; }
    pw.gr = gamma[0]; pw.gi = gamma[1]; pw.ar = amplitude[0]; pw.ai = amplitude[1];
    DevBuf<int> ptr(n + 1), cell(sp->cell.size() + 1), ref(sp->ref.size() + 1);
    DevBuf<double> coef(sp->coeff.size() + 1), db(2 * n);
    ptr.upload(sp->ptr.data(), n + 1, ctx->stream);
    cell.upload(sp->cell.data(), sp->cell.size(), ctx->stream);
    ref.upload(sp->ref.data(), sp->ref.size(), ctx->stream);
    coef.upload(sp->coeff.data(), sp->coeff.size(), ctx->stream);
    if (accumulate) db.upload(b, 2 * n, ctx->stream);
    k_planewave<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(m.d_geo.p, qp, Q, ptr.p, cell.p, ref.p, coef.p,
                                                                    (int)n, pw, trace, accumulate, db.p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) BH_FAIL(BH_ERR_HIP, "launch of k_planewave failed: %s", hipGetErrorString(e));
    BH_CHECK_HIP(hipMemcpyAsync(b, db.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, ctx->stream));
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    BH_API_END
}
