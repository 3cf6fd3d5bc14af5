/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's Galerkin assembly of the
 * Maxwell EFIE (MWSingleLayer3D), MFIE-K (MWDoubleLayer3D) and n x K
 * (DoubleLayerRotatedMW3D) operators on
 * Raviart-Thomas spaces, and of the plane-wave right-hand side.  It is used
 * only by tests/, __graft_entry__.smoke() and bench.py's CPU baseline, as the
 * checker.  The product (the HIP library) never links or calls it.
 *
 * It deliberately follows the reference's arithmetic arrangement (per point
 * pair, all 9 local entries, as in the reference's integrands) rather than the
 * factorised form of the GPU kernels, so that the two are independent.
 *
 * Reference files followed (paths relative to the reference repository):
 *   classification      src/quadrature/doublenumwiltonsauterqstrat.jl:32-69,
 *                       src/quadrature/doublenumsauterqstrat.jl:82-97,226-240
 *   double quadrature   src/quadrature/doublenumints.jl:12-43
 *   kernels             src/maxwell/mwops.jl:100-174
 *   RT shapes           src/bases/local/rtlocal.jl:8-27
 *   Wilton extraction   src/quadrature/singularityextractionints.jl:4-22,
 *                       src/maxwell/wiltonints.jl:8-84,
 *                       analytic integrals as restated in test/test_wiltonints.jl:60-232
 *   Sauter-Schwab       src/quadrature/sauterschwabints.jl:24-46,111-182; the
 *                       region parameterisations are those of Sauter & Schwab,
 *                       "Boundary Element Methods" (2011), Sec. 5.2 (the
 *                       reference delegates them to SauterSchwabQuadrature.jl 2.4,
 *                       which is not available here)
 *   scatter             src/integralop.jl:146-184, src/operator.jl:168-184
 *   plane-wave RHS      src/excitation.jl:38-115,149-198, src/maxwell/mwexc.jl:36-42
 *   rotated double layer src/maxwell/nxdbllayer.jl:17-44
 *   NCross, Identity    src/identityop.jl:6-22, src/localop.jl:114-146,362-386
 *   far fields          src/postproc.jl:174-211,273-288, src/maxwell/farfield.jl:56-75,98-100
 *
 * Parity pins (see tests/test_oracle_*.py): Dunavant exactness, Wilton vs brute
 * force, DQ vs Wilton-SE on far pairs, SS vs Wilton-SE at the reference test
 * tolerances, SS convergence in n, symmetry of the EFIE matrix, solve residual.
 * Quadrature-rule tables and the SS vertex-reorder convention of the external
 * Julia packages are parity unpinned at the 1e-12 level (no Julia available).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* Quadrature rules                                                     */
/* ------------------------------------------------------------------ */

/* Dunavant (1985) symmetric rules, orbit form, weights summing to 1. */
typedef struct { int kind; double w; double c[3]; } orbit_t;

static const orbit_t D1[] = {{1, 1.0, {0}}};
static const orbit_t D2[] = {{3, 0.333333333333333, {0.666666666666667, 0.166666666666667}}};
static const orbit_t D3[] = {{1, -0.5625, {0}}, {3, 0.520833333333333, {0.6, 0.2}}};
static const orbit_t D4[] = {{3, 0.223381589678011, {0.108103018168070, 0.445948490915965}},
                             {3, 0.109951743655322, {0.816847572980459, 0.091576213509771}}};
static const orbit_t D5[] = {{1, 0.225, {0}},
                             {3, 0.132394152788506, {0.059715871789770, 0.470142064105115}},
                             {3, 0.125939180544827, {0.797426985353087, 0.101286507323456}}};
static const orbit_t D6[] = {{3, 0.116786275726379, {0.501426509658179, 0.249286745170910}},
                             {3, 0.050844906370207, {0.873821971016996, 0.063089014491502}},
                             {6, 0.082851075618374, {0.053145049844817, 0.310352451033784, 0.636502499121399}}};
static const orbit_t D7[] = {{1, -0.149570044467682, {0}},
                             {3, 0.175615257433208, {0.479308067841920, 0.260345966079040}},
                             {3, 0.053347235608838, {0.869739794195568, 0.065130102902216}},
                             {6, 0.077113760890257, {0.048690315425316, 0.312865496004874, 0.638444188569810}}};
static const orbit_t D8[] = {{1, 0.144315607677787, {0}},
                             {3, 0.095091634267285, {0.081414823414554, 0.459292588292723}},
                             {3, 0.103217370534718, {0.658861384496480, 0.170569307751760}},
                             {3, 0.032458497623198, {0.898905543365938, 0.050547228317031}},
                             {6, 0.027230314174435, {0.008394777409958, 0.263112829634638, 0.728492392955404}}};

static const orbit_t D9[] = {{1, 0.097135796282799, {0}},
                             {3, 0.031334700227139, {0.020634961602525, 0.489682519198738}},
                             {3, 0.077827541004774, {0.125820817014127, 0.437089591492937}},
                             {3, 0.079647738927210, {0.623592928761935, 0.188203535619033}},
                             {3, 0.025577675658698, {0.910540973211095, 0.044729513394453}},
                             {6, 0.043283539377289, {0.036838412054736, 0.221962989160766, 0.741198598784498}}};
static const orbit_t D10[] = {{1, 0.090817990382754, {0}},
                              {3, 0.036725957756467, {0.028844733232685, 0.485577633383657}},
                              {3, 0.045321059435528, {0.781036849029926, 0.109481575485037}},
                              {6, 0.072757916845420, {0.141707219414880, 0.307939838764121, 0.550352941820999}},
                              {6, 0.028327242531057, {0.025003534762686, 0.246672560639903, 0.728323904597411}},
                              {6, 0.009421666963733, {0.009540815400299, 0.066803251012200, 0.923655933587500}}};

static const orbit_t *DTAB[11] = {0, D1, D2, D3, D4, D5, D6, D7, D8, D9, D10};
static const int DLEN[11] = {0, 1, 1, 2, 2, 3, 3, 4, 5, 6, 6};

#define OR_MAXQ 25

/* Fills u, v, w (weights sum to 1/2); returns the number of points or -1. */
int or_trgauss(int degree, double *u, double *v, double *w)
{
    if (degree < 1 || degree > 10) return -1;
    int n = 0;
    for (int o = 0; o < DLEN[degree]; ++o) {
        const orbit_t *ob = &DTAB[degree][o];
        double bary[6][3];
        int m;
        if (ob->kind == 1) {
            m = 1; bary[0][0] = bary[0][1] = bary[0][2] = 1.0 / 3.0;
        } else if (ob->kind == 3) {
            double a = ob->c[0], b = ob->c[1];
            m = 3;
            double t[3][3] = {{a, b, b}, {b, a, b}, {b, b, a}};
            memcpy(bary, t, sizeof t);
        } else {
            double a = ob->c[0], b = ob->c[1], c = ob->c[2];
            m = 6;
            double t[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
            memcpy(bary, t, sizeof t);
        }
        for (int k = 0; k < m; ++k) {
            u[n] = bary[k][0]; v[n] = bary[k][1]; w[n] = 0.5 * ob->w; ++n;
        }
    }
    return n;
}

/* Gauss-Legendre on [0,1] by Newton iteration on P_n (mwops.jl:46-51 semantics):
 * nodes ascending, weights summing to 1. */
static double legendre_p(int n, double z, double *dp)
{
    double p0 = 1.0, p1 = z;
    if (n == 0) { *dp = 0.0; return 1.0; }
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
    }
    *dp = n * (z * p1 - p0) / (z * z - 1.0);
    return p1;
}

void or_legendre01(int n, double *x, double *w)
{
    for (int i = 0; i < n; ++i) {
        double z = cos(OR_PI * (i + 0.75) / (n + 0.5)), dp;
        for (int it = 0; it < 100; ++it) {
            double p = legendre_p(n, z, &dp);
            double dz = p / dp;
            z -= dz;
            if (fabs(dz) < 1e-15) break;
        }
        legendre_p(n, z, &dp);
        x[n - 1 - i] = 0.5 * (z + 1.0);
        w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* ------------------------------------------------------------------ */
/* Small vector helpers                                                 */
/* ------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void sub3(double *r, const double *a, const double *b) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
static inline void cross3(double *r, const double *a, const double *b)
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }

static double tri_area(const double *v /* 9 */)
{
    double t1[3], t2[3], n[3];
    sub3(t1, v + 0, v + 6);
    sub3(t2, v + 3, v + 6);
    cross3(n, t1, t2);
    return 0.5 * norm3(n);
}

/* ------------------------------------------------------------------ */
/* Operator description                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int kind;            /* 0 = MWSingleLayer3D, 1 = MWDoubleLayer3D, 3 = DoubleLayerRotatedMW3D */
    cplx gamma, alpha, beta;
    double nx[3];        /* unit normal of the test cell (kind 3) */
} op_t;

enum { OR_OP_SL = 0, OR_OP_DL = 1, OR_OP_DL_ROTATED = 3 };

/* unit normal (v1 - v3) x (v2 - v3) / |.| of a cell, the chart normal */
static void cell_normal(const double *tv, double *nrm);

/* ------------------------------------------------------------------ */
/* Classification (doublenumwiltonsauterqstrat.jl:32-69)                */
/* ------------------------------------------------------------------ */
enum { OR_FAR = 0, OR_NEAR = 1, OR_CV = 2, OR_CE = 3, OR_CF = 4 };

/* Returns the pair class; hits beyond 3 return -1 (reference: @assert).
 * ti/sj receive the matching vertex pairs in test-major order. */
int or_classify(const double *tv, const double *sv, double k2, int wilton, int *ti, int *sj)
{
    const double dtol = 1.0e3 * 2.220446049250313e-16;
    int hits = 0;
    double dmin2 = 1.79769313486231570815e+308;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double dx = tv[3 * i] - sv[3 * j], dy = tv[3 * i + 1] - sv[3 * j + 1], dz = tv[3 * i + 2] - sv[3 * j + 2];
            volatile double d2 = dx * dx;
            d2 = d2 + dy * dy;
            d2 = d2 + dz * dz;
            double d = sqrt(d2);
            if (d2 < dmin2) dmin2 = d2;
            if (d < dtol) {
                if (hits < 3) { ti[hits] = i; sj[hits] = j; }
                ++hits;
            }
        }
    if (hits > 3) return -1;
    if (hits == 3) return OR_CF;
    if (hits == 2) return OR_CE;
    if (hits == 1) return OR_CV;
    if (wilton) {
        double h2 = tri_area(sv);
        double xtol2 = 0.2 * 0.2;
        double a = dmin2 * k2, b = dmin2 / (16.0 * h2);
        double m = a > b ? a : b;
        if (m < xtol2) return OR_NEAR;
    }
    return OR_FAR;
}

/* ------------------------------------------------------------------ */
/* Kernels evaluated as in mwops.jl:101-174                             */
/* ------------------------------------------------------------------ */
static const double I4PI = 1.0 / (4.0 * OR_PI);

static cplx cexpm1_acc(cplx z)
{
    double a = creal(z), b = cimag(z);
    double ea = exp(a);
    double sb2 = sin(0.5 * b);
    double re = expm1(a) * cos(b) - 2.0 * sb2 * sb2;
    double im = ea * sin(b);
    return re + I * im;
}

/* Adds w * integrand(x, y) to z for all local 3x3 shape pairs.
 * f and g are the (scaled) test/trial shape values; df, dg divergences. */
static void integrand_add(const op_t *op, int regular, const double *x, const double *y,
                          double f[3][3], const double df[3], double g[3][3], const double dg[3],
                          double wgt, cplx z[3][3])
{
    double r[3];
    sub3(r, x, y);
    double R = norm3(r);
    double iR = 1.0 / R;
    cplx gam = op->gamma;
    if (op->kind == 0) {
        cplx green;
        if (!regular) {
            green = cexp(-gam * R) * (I4PI * iR);
        } else {
            cplx gR = gam * R;
            green = (cexpm1_acc(-gR) + gR - 0.5 * gR * gR) / (4.0 * OR_PI * R);
        }
        cplx aG = op->alpha * green, bG = op->beta * green;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                z[i][j] += wgt * (aG * dot3(f[i], g[j]) + bG * (df[i] * dg[j]));
    } else if (op->kind == OR_OP_DL_ROTATED) {
        /* nxdbllayer.jl:27-44: K = -(gamma + 1/R) G r/R, f . (nx x (K x g)); the
         * reference has no regular/singular split for this operator */
        cplx green = cexp(-gam * R) / (4.0 * OR_PI * R);
        cplx K[3];
        for (int k = 0; k < 3; ++k) K[k] = -(gam + iR) * green * (iR * r[k]);
        const double *nx = op->nx;
        for (int j = 0; j < 3; ++j) {
            cplx Kg[3] = {K[1] * g[j][2] - K[2] * g[j][1], K[2] * g[j][0] - K[0] * g[j][2],
                          K[0] * g[j][1] - K[1] * g[j][0]};
            cplx nxKg[3] = {nx[1] * Kg[2] - nx[2] * Kg[1], nx[2] * Kg[0] - nx[0] * Kg[2],
                            nx[0] * Kg[1] - nx[1] * Kg[0]};
            for (int i = 0; i < 3; ++i)
                z[i][j] += wgt * (f[i][0] * nxKg[0] + f[i][1] * nxKg[1] + f[i][2] * nxKg[2]);
        }
    } else {
        cplx gg;
        if (!regular) {
            cplx green = cexp(-gam * R) * (iR * I4PI);
            gg = -(gam + iR) * green * iR;
        } else {
            cplx gR = gam * R;
            cplx expo = cexp(-gR);
            gg = (-(gR + 1.0) * expo + (1.0 - 0.5 * gR * gR)) * (I4PI * iR * iR * iR);
        }
        for (int j = 0; j < 3; ++j) {
            /* G_j = gradgreen x g_j, gradgreen = gg * r */
            cplx Gx = gg * (r[1] * g[j][2] - r[2] * g[j][1]);
            cplx Gy = gg * (r[2] * g[j][0] - r[0] * g[j][2]);
            cplx Gz = gg * (r[0] * g[j][1] - r[1] * g[j][0]);
            for (int i = 0; i < 3; ++i)
                z[i][j] += wgt * (f[i][0] * Gx + f[i][1] * Gy + f[i][2] * Gz);
        }
    }
}

/* RT values (rtlocal.jl:8-27) at physical point x on triangle tv with area A. */
static void rt_values(const double *tv, double A, const double *x, double f[3][3], double d[3])
{
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) f[i][k] = (x[k] - tv[3 * i + k]) / (2.0 * A);
        d[i] = 1.0 / A;
    }
}

static int tri_points(const double *tv, int degree, double *pts, double *wts)
{
    double u[OR_MAXQ], v[OR_MAXQ], w[OR_MAXQ];
    int n = or_trgauss(degree, u, v, w);
    double A = tri_area(tv);
    for (int q = 0; q < n; ++q) {
        for (int k = 0; k < 3; ++k)
            pts[3 * q + k] = u[q] * tv[k] + v[q] * tv[3 + k] + (1.0 - u[q] - v[q]) * tv[6 + k];
        wts[q] = w[q] * 2.0 * A;
    }
    return n;
}

static void cell_normal(const double *tv, double *nrm)
{
    double t1[3], t2[3];
    sub3(t1, tv, tv + 6);
    sub3(t2, tv + 3, tv + 6);
    cross3(nrm, t1, t2);
    const double nn = norm3(nrm);
    for (int q = 0; q < 3; ++q) nrm[q] /= nn;
}

/* DoubleQuadRule (doublenumints.jl:12-43) */
static void block_dq(const op_t *op, int regular, const double *tv, const double *sv,
                     int deg_o, int deg_i, cplx z[3][3])
{
    double xp[3 * OR_MAXQ], wx[OR_MAXQ], yq[3 * OR_MAXQ], wy[OR_MAXQ];
    int no = tri_points(tv, deg_o, xp, wx), ni = tri_points(sv, deg_i, yq, wy);
    double At = tri_area(tv), As = tri_area(sv);
    for (int p = 0; p < no; ++p) {
        double f[3][3], df[3];
        rt_values(tv, At, xp + 3 * p, f, df);
        for (int q = 0; q < ni; ++q) {
            double g[3][3], dg[3];
            rt_values(sv, As, yq + 3 * q, g, dg);
            integrand_add(op, regular, xp + 3 * p, yq + 3 * q, f, df, g, dg, wx[p] * wy[q], z);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Analytic potential integrals over a triangle, as restated in         */
/* test/test_wiltonints.jl:60-232 (legend :234-248).                    */
/* s[0..5] = int d/R^3, 1/R, 1, R, R^2, R^3;                            */
/* v[0..4] = int (y-rho) R^k for k = -3,-1,0,1,2;                       */
/* g[0..3] = int grad_y R^k for k = -1, 1, 2, 3.                        */
/* ------------------------------------------------------------------ */
void or_wiltonints(const double *p1, const double *p2, const double *p3, const double *ctr,
                   double s[6], double v[5][3], double g[4][3])
{
    const double tol = 2.220446049250313e-16 * 1e3;
    double t1[3], t2[3], n[3];
    sub3(t1, p2, p1);
    sub3(t2, p3, p1);
    cross3(n, t1, t2);
    double nn = norm3(n);
    for (int k = 0; k < 3; ++k) n[k] /= nn;
    double tmp[3];
    sub3(tmp, ctr, p1);
    double d = dot3(tmp, n);
    double D = fabs(d);
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = ctr[k] - d * n[k];

    const double *seg[3][2] = {{p1, p2}, {p2, p3}, {p3, p1}};
    int c_inside = 1, nthrough = 0, through[3] = {0, 0, 0};
    memset(s, 0, 6 * sizeof(double));
    memset(v, 0, 15 * sizeof(double));

    for (int i = 0; i < 3; ++i) {
        const double *a = seg[i][0], *b = seg[i][1];
        double ab[3], t[3], m[3], ca[3], cb[3];
        sub3(ab, b, a);
        double l = norm3(ab);
        for (int k = 0; k < 3; ++k) t[k] = ab[k] / l;
        cross3(m, t, n);
        sub3(ca, a, c);
        sub3(cb, b, c);
        double xa = dot3(ca, t), xb = dot3(cb, t);
        double p = dot3(ca, m), P = fabs(p);
        if (p < 0) c_inside = 0;
        double w2 = p * p + d * d;
        double ra = sqrt(xa * xa + w2), rb = sqrt(xb * xb + w2), w = sqrt(w2);
        if (P > tol) {
            double lg = log((xb + rb) / (xa + ra));
            double at = atan(d * xb / (p * rb)) - atan(d * xa / (p * ra));
            s[0] += -at;
            s[1] += p * lg + d * at;
            s[2] += 0.5 * p * (xb - xa);
            s[3] += (1.0 / 6.0) * p * (xb * rb - xa * ra) + (0.5 * d * d + (1.0 / 6.0) * p * p) * p * lg + (1.0 / 3.0) * d * d * d * at;
            s[4] += 0.5 * xb * p * ((1.0 / 6.0) * xb * xb + 0.5 * p * p + d * d) - 0.5 * xa * p * ((1.0 / 6.0) * xa * xa + 0.5 * p * p + d * d);
            {
                double pb = (1.0 / 20.0) * p * xb * rb * (xb * xb + 0.5 * (9 * d * d + 5 * p * p));
                double pa = (1.0 / 20.0) * p * xa * ra * (xa * xa + 0.5 * (9 * d * d + 5 * p * p));
                double p2 = (1.0 / 40.0) * p * (15 * d * d * d * d + 10 * d * d * p * p + 3 * p * p * p * p) * lg + 0.2 * d * d * d * d * d * at;
                s[5] += (pb - pa) + p2;
            }
            for (int k = 0; k < 3; ++k) {
                v[0][k] += -lg * m[k];
                v[1][k] += 0.5 * ((xb * rb - xa * ra) + w * w * lg) * m[k];
                v[2][k] += (0.5 * xb * ((1.0 / 3.0) * xb * xb + p * p) - 0.5 * xa * ((1.0 / 3.0) * xa * xa + p * p)) * m[k];
                v[3][k] += ((1.0 / 12.0) * (xb * rb * rb * rb - xa * ra * ra * ra) + 0.125 * w * w * (xb * rb - xa * ra) + 0.125 * w * w * w * w * lg) * m[k];
                {
                    double qb = 0.05 * pow(xb, 5) + (1.0 / 6.0) * p * p * xb * xb * xb + 0.25 * pow(p, 4) * xb + 0.5 * d * d * ((1.0 / 3.0) * xb * xb * xb + p * p * xb);
                    double qa = 0.05 * pow(xa, 5) + (1.0 / 6.0) * p * p * xa * xa * xa + 0.25 * pow(p, 4) * xa + 0.5 * d * d * ((1.0 / 3.0) * xa * xa * xa + p * p * xa);
                    v[4][k] += (qb - qa) * m[k];
                }
            }
        } else {
            if (xa * xb < tol * tol) { through[nthrough++] = i; }
            double lg;
            if (D > tol) {
                lg = log((xb + rb) / (xa + ra));
            } else {
                lg = (xb > 0) ? log(xb / xa) : log(xa / xb);
            }
            for (int k = 0; k < 3; ++k) {
                v[0][k] += -lg * m[k];
                if (D > tol)
                    v[1][k] += (0.5 * (xb * rb - xa * ra) + 0.5 * w * w * lg) * m[k];
                else
                    v[1][k] += 0.5 * (xb * rb - xa * ra) * m[k];
                v[2][k] += ((1.0 / 6.0) * xb * xb * xb - (1.0 / 6.0) * xa * xa * xa) * m[k];
                v[3][k] += ((1.0 / 12.0) * (xb * rb * rb * rb - xa * ra * ra * ra) + 0.125 * w * w * (xb * rb - xa * ra) + 0.125 * w * w * w * w * lg) * m[k];
                {
                    double qb = 0.05 * pow(xb, 5) + (1.0 / 6.0) * p * p * xb * xb * xb + 0.25 * pow(p, 4) * xb + 0.5 * d * d * ((1.0 / 3.0) * xb * xb * xb + p * p * xb);
                    double qa = 0.05 * pow(xa, 5) + (1.0 / 6.0) * p * p * xa * xa * xa + 0.25 * pow(p, 4) * xa + 0.5 * d * d * ((1.0 / 3.0) * xa * xa * xa + p * p * xa);
                    v[4][k] += (qb - qa) * m[k];
                }
            }
        }
    }
    double sgn = D < tol ? 0.0 : (d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0));
    double alpha = 0.0;
    if (nthrough == 0 && c_inside) alpha = 2.0 * OR_PI;
    else if (nthrough == 1) alpha = OR_PI;
    else if (nthrough == 2) {
        const double *a1 = seg[through[0]][0], *b1 = seg[through[0]][1];
        const double *a2 = seg[through[1]][0], *b2 = seg[through[1]][1];
        double e1[3], e2[3];
        sub3(e1, b1, a1);
        sub3(e2, b2, a2);
        double ca = -dot3(e1, e2) / (norm3(e1) * norm3(e2));
        if (ca > 1) ca = 1;
        if (ca < -1) ca = -1;
        alpha = acos(ca);
    }
    s[0] += alpha * sgn;
    s[1] += -alpha * D;
    s[3] += -(1.0 / 3.0) * alpha * D * D * D;
    s[5] += -(1.0 / 5.0) * alpha * D * D * D * D * D;
    for (int k = 0; k < 3; ++k) {
        g[0][k] = -(v[0][k] - s[0] * n[k]);
        g[1][k] = (v[1][k] - d * s[1] * n[k]);
        g[2][k] = 2.0 * (v[2][k] - d * s[2] * n[k]);
        g[3][k] = 3.0 * (v[3][k] - d * s[3] * n[k]);
    }
}

/* WiltonSERule (singularityextractionints.jl:4-22, wiltonints.jl:8-84) */
static void block_wilton(const op_t *op, const double *tv, const double *sv, int deg_o, int deg_i, cplx z[3][3])
{
    block_dq(op, 1, tv, sv, deg_o, deg_i, z);
    double xp[3 * OR_MAXQ], wx[OR_MAXQ];
    int no = tri_points(tv, deg_o, xp, wx);
    double At = tri_area(tv), As = tri_area(sv);
    const double *s1 = sv, *s2 = sv + 3, *s3 = sv + 6;
    double nrm[3], e13[3], e23[3];
    sub3(e13, s1, s3);
    sub3(e23, s2, s3);
    cross3(nrm, e13, e23);
    double nn = norm3(nrm);
    for (int k = 0; k < 3; ++k) nrm[k] /= nn;
    cplx gam = op->gamma;
    double sc = 1.0 / At / As / 4.0;
    for (int p = 0; p < no; ++p) {
        const double *x = xp + 3 * p;
        double dx = wx[p];
        double s[6], v[5][3], g[4][3];
        or_wiltonints(s1, s2, s3, x, s, v, g);
        if (op->kind == 0) {
            double xs1[3];
            sub3(xs1, x, s1);
            double h = dot3(xs1, nrm);
            double rho[3] = {x[0] - h * nrm[0], x[1] - h * nrm[1], x[2] - h * nrm[2]};
            /* scal[2..4] = s[1..3]; vec[2..4] = v[1..3] */
            cplx IG = (s[1] - gam * s[2] + 0.5 * gam * gam * s[3]) / (4.0 * OR_PI);
            cplx IGy[3];
            for (int k = 0; k < 3; ++k)
                IGy[k] = (v[1][k] + s[1] * rho[k] - gam * (v[2][k] + s[2] * rho[k]) + 0.5 * gam * gam * (v[3][k] + s[3] * rho[k])) / (4.0 * OR_PI);
            for (int i = 0; i < 3; ++i) {
                const double *a = tv + 3 * i;
                double gx[3] = {x[0] - a[0], x[1] - a[1], x[2] - a[2]};
                for (int j = 0; j < 3; ++j) {
                    const double *b = sv + 3 * j;
                    cplx IGf0 = IGy[0] - IG * b[0], IGf1 = IGy[1] - IG * b[1], IGf2 = IGy[2] - IG * b[2];
                    cplx IGdf = 2.0 * IG;
                    cplx dgIGf = gx[0] * IGf0 + gx[1] * IGf1 + gx[2] * IGf2;
                    z[i][j] += (sc * op->alpha * dgIGf + sc * op->beta * 2.0 * IGdf) * dx;
                }
            }
        } else {
            /* grad[1] = g[0] (int grad' 1/R), grad[3] = g[1] (int grad' R) */
            cplx IdG[3];
            for (int k = 0; k < 3; ++k) IdG[k] = (-g[0][k] - 0.5 * gam * gam * g[1][k]) / (4.0 * OR_PI);
            for (int i = 0; i < 3; ++i) {
                const double *a = tv + 3 * i;
                double gx[3] = {x[0] - a[0], x[1] - a[1], x[2] - a[2]};
                for (int j = 0; j < 3; ++j) {
                    const double *b = sv + 3 * j;
                    double xb[3] = {x[0] - b[0], x[1] - b[1], x[2] - b[2]};
                    double cr[3];
                    cross3(cr, xb, gx);
                    z[i][j] += sc * (cr[0] * IdG[0] + cr[1] * IdG[1] + cr[2] * IdG[2]) * dx;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Sauter-Schwab (sauterschwabints.jl:111-182; Sauter & Schwab Sec 5.2)  */
/* Reference element {0 <= x2 <= x1 <= 1}; a point (x1, x2) maps to     */
/* P1 + x1 (P2 - P1) + x2 (P3 - P2) with P the reordered vertices        */
/* (shared vertices first, in matching order).                           */
/* ------------------------------------------------------------------ */
static void ss_map(const double *P, double x1, double x2, double *out)
{
    for (int k = 0; k < 3; ++k)
        out[k] = P[k] + x1 * (P[3 + k] - P[k]) + x2 * (P[6 + k] - P[3 + k]);
}

/* region r of the class: writes xhat, yhat and returns the Jacobian factor */
static double ss_region(int cls, int r, double xi, double e1, double e2, double e3, double *xh, double *yh)
{
    if (cls == OR_CF) {
        double J = xi * xi * xi * e1 * e1 * e2;
        switch (r) {
        case 0: xh[0] = xi; xh[1] = xi * (1 - e1 + e1 * e2); yh[0] = xi * (1 - e1 * e2 * e3); yh[1] = xi * (1 - e1); break;
        case 1: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * (1 - e1); yh[0] = xi; yh[1] = xi * (1 - e1 + e1 * e2); break;
        case 2: xh[0] = xi; xh[1] = xi * e1 * (1 - e2 + e2 * e3); yh[0] = xi * (1 - e1 * e2); yh[1] = xi * e1 * (1 - e2); break;
        case 3: xh[0] = xi * (1 - e1 * e2); xh[1] = xi * e1 * (1 - e2); yh[0] = xi; yh[1] = xi * e1 * (1 - e2 + e2 * e3); break;
        case 4: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * e1 * (1 - e2 * e3); yh[0] = xi; yh[1] = xi * e1 * (1 - e2); break;
        default: xh[0] = xi; xh[1] = xi * e1 * (1 - e2); yh[0] = xi * (1 - e1 * e2 * e3); yh[1] = xi * e1 * (1 - e2 * e3); break;
        }
        return J;
    } else if (cls == OR_CE) {
        switch (r) {
        case 0: xh[0] = xi; xh[1] = xi * e1 * e3; yh[0] = xi * (1 - e1 * e2); yh[1] = xi * e1 * (1 - e2);
            return xi * xi * xi * e1 * e1;
        case 1: xh[0] = xi; xh[1] = xi * e1; yh[0] = xi * (1 - e1 * e2 * e3); yh[1] = xi * e1 * e2 * (1 - e3); break;
        case 2: xh[0] = xi * (1 - e1 * e2); xh[1] = xi * e1 * (1 - e2); yh[0] = xi; yh[1] = xi * e1 * e2 * e3; break;
        case 3: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * e1 * e2 * (1 - e3); yh[0] = xi; yh[1] = xi * e1; break;
        default: xh[0] = xi * (1 - e1 * e2 * e3); xh[1] = xi * e1 * (1 - e2 * e3); yh[0] = xi; yh[1] = xi * e1 * e2; break;
        }
        return xi * xi * xi * e1 * e1 * e2;
    } else {
        double J = xi * xi * xi * e2;
        if (r == 0) { xh[0] = xi; xh[1] = xi * e1; yh[0] = xi * e2; yh[1] = xi * e2 * e3; }
        else { xh[0] = xi * e2; xh[1] = xi * e2 * e3; yh[0] = xi; yh[1] = xi * e1; }
        return J;
    }
}

static int ss_nregions(int cls) { return cls == OR_CF ? 6 : (cls == OR_CE ? 5 : 2); }

/* Builds the reordered vertex lists (SauterSchwabQuadrature.reorder semantics:
 * shared vertices first, matched; then the rest in ascending local order). */
static void ss_reorder(int cls, const int *ti, const int *sj, int *pi, int *pj)
{
    int nh = cls == OR_CF ? 3 : (cls == OR_CE ? 2 : 1);
    int usedI[3] = {0, 0, 0}, usedJ[3] = {0, 0, 0};
    for (int h = 0; h < nh; ++h) { pi[h] = ti[h]; pj[h] = sj[h]; usedI[ti[h]] = 1; usedJ[sj[h]] = 1; }
    int a = nh, b = nh;
    for (int k = 0; k < 3; ++k) {
        if (!usedI[k]) pi[a++] = k;
        if (!usedJ[k]) pj[b++] = k;
    }
}

static void block_ss(const op_t *op, int cls, const int *ti, const int *sj, const double *tv, const double *sv,
                     int n, cplx z[3][3])
{
    double gx[64], gw[64];
    or_legendre01(n, gx, gw);
    int pi[3], pj[3];
    ss_reorder(cls, ti, sj, pi, pj);
    double Pt[9], Ps[9];
    for (int h = 0; h < 3; ++h)
        for (int k = 0; k < 3; ++k) { Pt[3 * h + k] = tv[3 * pi[h] + k]; Ps[3 * h + k] = sv[3 * pj[h] + k]; }
    int nreg = ss_nregions(cls);
    /* shapes unnormalised: f_i = x - a_i, div = 2 (sauterschwabints.jl:24-46) */
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            for (int c = 0; c < n; ++c)
                for (int d = 0; d < n; ++d) {
                    double w4 = gw[a] * gw[b] * gw[c] * gw[d];
                    for (int r = 0; r < nreg; ++r) {
                        double xh[2], yh[2], x[3], y[3];
                        double Jac = ss_region(cls, r, gx[a], gx[b], gx[c], gx[d], xh, yh);
                        ss_map(Pt, xh[0], xh[1], x);
                        ss_map(Ps, yh[0], yh[1], y);
                        double f[3][3], g[3][3], df[3] = {2, 2, 2}, dg[3] = {2, 2, 2};
                        for (int i = 0; i < 3; ++i)
                            for (int k = 0; k < 3; ++k) { f[i][k] = x[k] - tv[3 * i + k]; g[i][k] = y[k] - sv[3 * i + k]; }
                        integrand_add(op, 0, x, y, f, df, g, dg, w4 * Jac, z);
                    }
                }
}

/* ------------------------------------------------------------------ */
/* Local block for one pair under a quadstrat                           */
/*   qs_kind 0: DoubleNumWiltonSauterQStrat(of, if, on, in, tetr, face, edge, vert)
 *   qs_kind 1: DoubleNumSauterQstrat(o, i, tetr, face, edge, vert)
 *   qs_kind 2: DoubleNumQStrat(o, i)                                    */
/* ------------------------------------------------------------------ */
int or_block(int op_kind, const double *g, const double *al, const double *be, int qs_kind, const int *qs,
             const double *tv, const double *sv, double *zout /* 18, column-major 3x3 complex */, int *cls_out)
{
    op_t op = {op_kind, g[0] + I * g[1], al[0] + I * al[1], be[0] + I * be[1], {0, 0, 0}};
    /* the rotated double layer has no singularity extraction in the reference
     * (no regularpart/singularpart), so DoubleNumWiltonSauterQStrat does not apply */
    if (op_kind == OR_OP_DL_ROTATED && qs_kind == 0) return -3;
    cell_normal(tv, op.nx);
    cplx z[3][3];
    memset(z, 0, sizeof z);
    int ti[3], sj[3];
    double k2 = creal(op.gamma) * creal(op.gamma) + cimag(op.gamma) * cimag(op.gamma);
    int cls;
    if (qs_kind == 2) cls = OR_FAR;
    else cls = or_classify(tv, sv, k2, qs_kind == 0, ti, sj);
    if (cls < 0) return -1;
    if (cls_out) *cls_out = cls;
    int of = qs[0], inf = qs[1];
    if (cls == OR_FAR) block_dq(&op, 0, tv, sv, of, inf, z);
    else if (cls == OR_NEAR) block_wilton(&op, tv, sv, qs[2], qs[3], z);
    else {
        int face = qs_kind == 0 ? qs[5] : qs[3], edge = qs_kind == 0 ? qs[6] : qs[4], vert = qs_kind == 0 ? qs[7] : qs[5];
        int n = cls == OR_CF ? face : (cls == OR_CE ? edge : vert);
        block_ss(&op, cls, ti, sj, tv, sv, n, z);
    }
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) { zout[2 * (i + 3 * j)] = creal(z[i][j]); zout[2 * (i + 3 * j) + 1] = cimag(z[i][j]); }
    return 0;
}

/* Block with explicit strategy (for the reference's cross-check tests):
 * strat 0 = DQ(deg_o, deg_i), 1 = WiltonSE(outer deg_o, regular DQ(deg_o, deg_i)),
 * 2 = SS with n points (class from classification). */
int or_block_strategy(int op_kind, const double *g, const double *al, const double *be, int strat,
                      int deg_o, int deg_i, int n, const double *tv, const double *sv, double *zout)
{
    op_t op = {op_kind, g[0] + I * g[1], al[0] + I * al[1], be[0] + I * be[1], {0, 0, 0}};
    cell_normal(tv, op.nx);
    cplx z[3][3];
    memset(z, 0, sizeof z);
    if (strat == 1 && op_kind == OR_OP_DL_ROTATED) return -3;
    if (strat == 0) block_dq(&op, 0, tv, sv, deg_o, deg_i, z);
    else if (strat == 1) block_wilton(&op, tv, sv, deg_o, deg_i, z);
    else {
        int ti[3], sj[3];
        int cls = or_classify(tv, sv, 0.0, 0, ti, sj);
        if (cls < OR_CV) return -2;
        block_ss(&op, cls, ti, sj, tv, sv, n, z);
    }
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) { zout[2 * (i + 3 * j)] = creal(z[i][j]); zout[2 * (i + 3 * j) + 1] = cimag(z[i][j]); }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Full assembly (integralop.jl:146-184): Z[m,n] += a * z[i,j] * b       */
/* Z column-major, complex interleaved, ld = ntest.                      */
/* Spaces in shape CSR form: ptr[ndof+1], cell, ref, coeff.              */
/* ------------------------------------------------------------------ */
typedef struct { int32_t dof; double c; } slot_t;

static slot_t *build_cellmap(int nf, int ndof, const int32_t *ptr, const int32_t *cell, const int32_t *ref,
                             const double *coeff, int *K)
{
    int *cnt = calloc((size_t)nf * 3, sizeof(int));
    for (int b = 0; b < ndof; ++b)
        for (int k = ptr[b]; k < ptr[b + 1]; ++k) cnt[3 * cell[k] + ref[k]]++;
    int kmax = 1;
    for (int i = 0; i < 3 * nf; ++i) if (cnt[i] > kmax) kmax = cnt[i];
    slot_t *map = malloc(sizeof(slot_t) * (size_t)nf * 3 * kmax);
    for (size_t i = 0; i < (size_t)nf * 3 * kmax; ++i) { map[i].dof = -1; map[i].c = 0; }
    memset(cnt, 0, sizeof(int) * (size_t)nf * 3);
    for (int b = 0; b < ndof; ++b)
        for (int k = ptr[b]; k < ptr[b + 1]; ++k) {
            int s = 3 * cell[k] + ref[k];
            map[(size_t)s * kmax + cnt[s]].dof = b;
            map[(size_t)s * kmax + cnt[s]].c = coeff[k];
            cnt[s]++;
        }
    free(cnt);
    *K = kmax;
    return map;
}

int or_assemble(int op_kind, const double *g, const double *al, const double *be, int qs_kind, const int *qs,
                const double *Vt, const int32_t *Ct, int nft,
                int ndt, const int32_t *tptr, const int32_t *tcell, const int32_t *tref, const double *tcoef,
                const double *Vs, const int32_t *Cs, int nfs,
                int nds, const int32_t *sptr, const int32_t *scell, const int32_t *sref, const double *scoef,
                double *Z, int8_t *cls_out)
{
    int Kt;
    slot_t *tm = build_cellmap(nft, ndt, tptr, tcell, tref, tcoef, &Kt);
    int err = 0;
    const int PC = 64;  /* test cells per chunk */
    double *blk = malloc(sizeof(double) * 18 * (size_t)PC * (size_t)(nfs > 0 ? nfs : 1));
    for (int p0 = 0; p0 < nft; p0 += PC) {
        int pc = nft - p0 < PC ? nft - p0 : PC;
        /* local blocks of the chunk, all trial cells */
        #pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for (int pp = 0; pp < pc; ++pp)
            for (int q = 0; q < nfs; ++q) {
                int p = p0 + pp;
                double tv[9], sv[9];
                for (int a = 0; a < 3; ++a)
                    for (int k = 0; k < 3; ++k) {
                        tv[3 * a + k] = Vt[3 * Ct[3 * p + a] + k];
                        sv[3 * a + k] = Vs[3 * Cs[3 * q + a] + k];
                    }
                int cls = 0;
                const int rc = or_block(op_kind, g, al, be, qs_kind, qs, tv, sv, blk + 18 * ((size_t)pp * nfs + q),
                                        &cls);
                if (rc != 0) {
                    #pragma omp atomic write
                    err = rc;
                }
                if (cls_out) cls_out[(size_t)p * nfs + q] = (int8_t)cls;
            }
        /* signed scatter store(a*(z*b), m, n); each thread owns whole columns n,
         * so the summation order is fixed: chunk, trial shape, test cell, test shape */
        #pragma omp parallel for schedule(static)
        for (int n = 0; n < nds; ++n)
            for (int k = sptr[n]; k < sptr[n + 1]; ++k) {
                int q = scell[k], j = sref[k];
                double b = scoef[k];
                if (b == 0) continue;
                for (int pp = 0; pp < pc; ++pp) {
                    int p = p0 + pp;
                    const double *zb = blk + 18 * ((size_t)pp * nfs + q);
                    for (int i = 0; i < 3; ++i)
                        for (int ka = 0; ka < Kt; ++ka) {
                            slot_t ta = tm[(size_t)(3 * p + i) * Kt + ka];
                            if (ta.dof < 0 || ta.c == 0) continue;
                            size_t idx = 2 * ((size_t)ta.dof + (size_t)ndt * n);
                            double zr = zb[2 * (i + 3 * j)] * b, zi = zb[2 * (i + 3 * j) + 1] * b;
                            Z[idx] += ta.c * zr;
                            Z[idx + 1] += ta.c * zi;
                        }
                }
            }
    }
    free(blk);
    free(tm);
    (void)nds;
    return err;
}

/* ------------------------------------------------------------------ */
/* Local operators on RT x RT over one mesh (the matched-mesh loop         */
/* localop.jl:114-146 with cellinteractions_matched!, localop.jl:362-386;  */
/* SingleNumQStrat(degree), identityop.jl:66-67):                          */
/*   kind 0 NCross   (identityop.jl:20-22): locmat[i,j] = sum_q w_q f_i . (n x f_j) */
/*   kind 1 Identity (identityop.jl:9-10):  locmat[i,j] = sum_q w_q f_i . f_j        */
/*   Z[m,n] += scale * a * locmat[i,j] * b.                                  */
/* Z column-major, complex interleaved, ld = ntest.                          */
/* ------------------------------------------------------------------ */
int or_localop(int kind, const double *V, const int32_t *C, int nf, int nd, const int32_t *tptr,
                const int32_t *tcell, const int32_t *tref, const double *tcoef, int ns, const int32_t *sptr,
                const int32_t *scell, const int32_t *sref, const double *scoef, int degree, const double *scale,
                double *Z)
{
    if (kind != 0 && kind != 1) return -2;
    int Ks, Kt;
    slot_t *sm = build_cellmap(nf, ns, sptr, scell, sref, scoef, &Ks);
    slot_t *tm = build_cellmap(nf, nd, tptr, tcell, tref, tcoef, &Kt);
    const cplx sc = scale[0] + I * scale[1];
    for (int p = 0; p < nf; ++p) {
        double tv[9];
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < 3; ++q) tv[3 * a + q] = V[3 * C[3 * p + a] + q];
        double xp[3 * OR_MAXQ], wx[OR_MAXQ];
        const int np = tri_points(tv, degree, xp, wx);
        if (np < 0) { free(sm); free(tm); return -1; }
        const double A = tri_area(tv);
        double t1[3], t2[3], nrm[3];
        sub3(t1, tv, tv + 6);
        sub3(t2, tv + 3, tv + 6);
        cross3(nrm, t1, t2);
        const double nn = norm3(nrm);
        for (int q = 0; q < 3; ++q) nrm[q] /= nn;
        double loc[3][3] = {{0}};
        for (int k = 0; k < np; ++k) {
            const double *x = xp + 3 * k;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double fi[3], fj[3], nf_[3];
                    for (int q = 0; q < 3; ++q) {
                        fi[q] = (x[q] - tv[3 * i + q]) / (2.0 * A);
                        fj[q] = (x[q] - tv[3 * j + q]) / (2.0 * A);
                    }
                    if (kind == 0) {
                        cross3(nf_, nrm, fj);
                        loc[i][j] += wx[k] * dot3(fi, nf_);
                    } else {
                        loc[i][j] += wx[k] * dot3(fi, fj);
                    }
                }
        }
        /* store(a * locmat[i,j] * b, m, n) over the test and trial slots of the cell */
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int ka = 0; ka < Kt; ++ka) {
                    const slot_t ta = tm[(size_t)(3 * p + i) * Kt + ka];
                    if (ta.dof < 0) continue;
                    for (int kb = 0; kb < Ks; ++kb) {
                        const slot_t sb = sm[(size_t)(3 * p + j) * Ks + kb];
                        if (sb.dof < 0) continue;
                        const cplx v = sc * (ta.c * loc[i][j] * sb.c);
                        Z[2 * ((size_t)ta.dof + (size_t)nd * sb.dof)] += creal(v);
                        Z[2 * ((size_t)ta.dof + (size_t)nd * sb.dof) + 1] += cimag(v);
                    }
                }
    }
    free(sm);
    free(tm);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Plane-wave right-hand side (excitation.jl:38-115, mwexc.jl:36-42),   */
/* E(x) = amp exp(-gamma d.x) pol:                                        */
/*   trace 0, TangTraceMW (n x E) x n (excitation.jl:175-180):           */
/*     b[m] += a * sum_p w_p f_i(x_p) . E(x_p) (f_i tangential);          */
/*   trace 1, CrossTraceMW n x E (excitation.jl:168-173):                */
/*     b[m] += a * sum_p w_p f_i(x_p) . (n x E(x_p)).                     */
/* ------------------------------------------------------------------ */
int or_planewave(int trace, const double *V, const int32_t *C, int nf, int nd, const int32_t *ptr,
                 const int32_t *cell, const int32_t *ref, const double *coef, const double *dir,
                 const double *pol_re, const double *pol_im, const double *gam, const double *amp, int degree,
                 double *b)
{
    if (trace != 0 && trace != 1) return -2;
    cplx gm = gam[0] + I * gam[1], am = amp[0] + I * amp[1];
    for (int m = 0; m < nd; ++m)
        for (int k = ptr[m]; k < ptr[m + 1]; ++k) {
            int c = cell[k], r = ref[k];
            double tv[9];
            for (int a = 0; a < 3; ++a)
                for (int q = 0; q < 3; ++q) tv[3 * a + q] = V[3 * C[3 * c + a] + q];
            double xp[3 * OR_MAXQ], wx[OR_MAXQ];
            int np = tri_points(tv, degree, xp, wx);
            if (np < 0) return -1;
            double A = tri_area(tv);
            double nrm[3];
            cell_normal(tv, nrm);
            cplx acc = 0;
            for (int p = 0; p < np; ++p) {
                const double *x = xp + 3 * p;
                cplx ph = am * cexp(-gm * dot3(dir, x));
                double f[3];
                for (int q = 0; q < 3; ++q) f[q] = (x[q] - tv[3 * r + q]) / (2.0 * A);
                cplx E[3], e = 0;
                for (int q = 0; q < 3; ++q) E[q] = ph * (pol_re[q] + I * pol_im[q]);
                if (trace == 1) {
                    cplx nE[3] = {nrm[1] * E[2] - nrm[2] * E[1], nrm[2] * E[0] - nrm[0] * E[2],
                                  nrm[0] * E[1] - nrm[1] * E[0]};
                    for (int q = 0; q < 3; ++q) e += f[q] * nE[q];
                } else {
                    for (int q = 0; q < 3; ++q) e += f[q] * E[q];
                }
                acc += wx[p] * e;
            }
            cplx v = coef[k] * acc;
            b[2 * m] += creal(v);
            b[2 * m + 1] += cimag(v);
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Far fields (postproc.jl:174-211 potential!, :273-288 farfieldlocal!, */
/* with maxwell/farfield.jl:56-75,98-100): for each point y, for each   */
/* cell, zlocal[r] = sum_x dx I(y, krn f_r(x)), krn = exp(gamma y.x)    */
/* (expm1 for the drop-constant operator); then ff[y] += zlocal[r] b    */
/* coeffs[n] for every (n, b) of the cell's shape r.  kind: 0 single    */
/* layer (y x v) x y, 1 double layer y x v, 2 rotated double layer      */
/* y x (v x n), 3 drop-constant (y x v) x y with expm1.                 */
/* ------------------------------------------------------------------ */
static cplx or_cexpm1(cplx z)
{
    /* e^(a+ib) - 1 = (e^a - 1) - 2 e^a sin^2(b/2) + i e^a sin b */
    double a = creal(z), b = cimag(z), em = expm1(a), h = sin(0.5 * b);
    return (em - 2.0 * (em + 1.0) * h * h) + I * ((em + 1.0) * sin(b));
}

int or_farfield(const double *V, const int32_t *C, int nf, int nd, const int32_t *ptr, const int32_t *cell,
                const int32_t *ref, const double *coef, int kind, const double *gam, const double *amp, int npts,
                const double *pts, const double *cre, const double *cim, int degree, double *ff)
{
    if (kind < 0 || kind > 3) return -2;
    double u[OR_MAXQ], v[OR_MAXQ], w[OR_MAXQ];
    if (or_trgauss(degree, u, v, w) < 0) return -1;
    cplx gm = gam[0] + I * gam[1], am = amp[0] + I * amp[1];
    /* element-to-dof lists: shapes of each (cell, r) in ascending dof order */
    int ns = ptr[nd];
    int *cnt = calloc((size_t)3 * nf + 1, sizeof(int)), *sh = malloc(sizeof(int) * (size_t)(ns + 1));
    int *fill = calloc((size_t)3 * nf + 1, sizeof(int)), *dofof = malloc(sizeof(int) * (size_t)(ns + 1));
    for (int k = 0; k < ns; ++k) cnt[3 * cell[k] + ref[k] + 1]++;
    for (int s = 0; s < 3 * nf; ++s) cnt[s + 1] += cnt[s];
    for (int m = 0; m < nd; ++m)
        for (int k = ptr[m]; k < ptr[m + 1]; ++k) {
            int s = 3 * cell[k] + ref[k];
            sh[cnt[s] + fill[s]] = k;
            dofof[cnt[s] + fill[s]] = m;
            fill[s]++;
        }
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < npts; ++p) {
        const double *y = pts + 3 * p;
        cplx acc[3] = {0, 0, 0};
        for (int c = 0; c < nf; ++c) {
            double tv[9];
            for (int a = 0; a < 3; ++a)
                for (int q = 0; q < 3; ++q) tv[3 * a + q] = V[3 * C[3 * c + a] + q];
            double xp[3 * OR_MAXQ], wx[OR_MAXQ];
            int np = tri_points(tv, degree, xp, wx);
            double A = tri_area(tv), t1[3], t2[3], nrm[3];
            sub3(t1, tv + 0, tv + 6);
            sub3(t2, tv + 3, tv + 6);
            cross3(nrm, t1, t2);
            double nn = norm3(nrm);
            for (int q = 0; q < 3; ++q) nrm[q] /= nn;
            cplx z[3][3] = {{0}};
            for (int k = 0; k < np; ++k) {
                const double *x = xp + 3 * k;
                cplx krn = kind == 3 ? or_cexpm1(gm * dot3(y, x)) : cexp(gm * dot3(y, x));
                for (int r = 0; r < 3; ++r) {
                    cplx f[3], a[3], b[3];
                    for (int q = 0; q < 3; ++q) f[q] = krn * ((x[q] - tv[3 * r + q]) / (2.0 * A));
                    if (kind == 2) { /* (krn f) x n */
                        cplx g0 = f[1] * nrm[2] - f[2] * nrm[1], g1 = f[2] * nrm[0] - f[0] * nrm[2],
                             g2 = f[0] * nrm[1] - f[1] * nrm[0];
                        f[0] = g0; f[1] = g1; f[2] = g2;
                    }
                    /* a = y x f */
                    a[0] = y[1] * f[2] - y[2] * f[1];
                    a[1] = y[2] * f[0] - y[0] * f[2];
                    a[2] = y[0] * f[1] - y[1] * f[0];
                    if (kind == 0 || kind == 3) { /* (y x f) x y */
                        b[0] = a[1] * y[2] - a[2] * y[1];
                        b[1] = a[2] * y[0] - a[0] * y[2];
                        b[2] = a[0] * y[1] - a[1] * y[0];
                    } else {
                        b[0] = a[0]; b[1] = a[1]; b[2] = a[2];
                    }
                    for (int q = 0; q < 3; ++q) z[r][q] += (am * b[q]) * wx[k];
                }
            }
            for (int r = 0; r < 3; ++r) {
                int s = 3 * c + r;
                for (int e = cnt[s]; e < cnt[s + 1]; ++e) {
                    int m = dofof[e];
                    cplx cf = coef[sh[e]] * (cre[m] + I * cim[m]);
                    for (int q = 0; q < 3; ++q) acc[q] += z[r][q] * cf;
                }
            }
        }
        for (int q = 0; q < 3; ++q) {
            ff[6 * p + 2 * q] = creal(acc[q]);
            ff[6 * p + 2 * q + 1] = cimag(acc[q]);
        }
    }
    free(cnt); free(sh); free(fill); free(dofof);
    return rc;
}
