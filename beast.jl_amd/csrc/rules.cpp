// Quadrature rule tables of the library (host side).
//
// Triangle rules: Dunavant (1985) symmetric rules indexed by polynomial degree,
// the rules the reference requests through trgauss/quadpoints
// (src/quadrature/doublenumwiltonsauterqstrat.jl:12-29, src/excitation.jl:6-9).
// Points are chart parameters (u, v) of r = u v1 + v v2 + (1-u-v) v3; weights sum to 1/2.
// Gauss-Legendre on [0,1]: _legendre(n, 0, 1) of src/maxwell/mwops.jl:46-51.
#include <cmath>
#include <cstring>

#include "bh_internal.h"

namespace bh {

namespace {
struct Orbit {
    int kind;
    double w;
    double c[3];
};

const Orbit kD1[] = {{1, 1.0, {0, 0, 0}}};
const Orbit kD2[] = {{3, 0.333333333333333, {0.666666666666667, 0.166666666666667, 0}}};
const Orbit kD3[] = {{1, -0.5625, {0, 0, 0}}, {3, 0.520833333333333, {0.6, 0.2, 0}}};
const Orbit kD4[] = {{3, 0.223381589678011, {0.108103018168070, 0.445948490915965, 0}},
                     {3, 0.109951743655322, {0.816847572980459, 0.091576213509771, 0}}};
const Orbit kD5[] = {{1, 0.225, {0, 0, 0}},
                     {3, 0.132394152788506, {0.059715871789770, 0.470142064105115, 0}},
                     {3, 0.125939180544827, {0.797426985353087, 0.101286507323456, 0}}};
const Orbit kD6[] = {{3, 0.116786275726379, {0.501426509658179, 0.249286745170910, 0}},
                     {3, 0.050844906370207, {0.873821971016996, 0.063089014491502, 0}},
                     {6, 0.082851075618374, {0.053145049844817, 0.310352451033784, 0.636502499121399}}};
const Orbit kD7[] = {{1, -0.149570044467682, {0, 0, 0}},
                     {3, 0.175615257433208, {0.479308067841920, 0.260345966079040, 0}},
                     {3, 0.053347235608838, {0.869739794195568, 0.065130102902216, 0}},
                     {6, 0.077113760890257, {0.048690315425316, 0.312865496004874, 0.638444188569810}}};
const Orbit kD8[] = {{1, 0.144315607677787, {0, 0, 0}},
                     {3, 0.095091634267285, {0.081414823414554, 0.459292588292723, 0}},
                     {3, 0.103217370534718, {0.658861384496480, 0.170569307751760, 0}},
                     {3, 0.032458497623198, {0.898905543365938, 0.050547228317031, 0}},
                     {6, 0.027230314174435, {0.008394777409958, 0.263112829634638, 0.728492392955404}}};

const Orbit kD9[] = {{1, 0.097135796282799, {0, 0, 0}},
                     {3, 0.031334700227139, {0.020634961602525, 0.489682519198738, 0}},
                     {3, 0.077827541004774, {0.125820817014127, 0.437089591492937, 0}},
                     {3, 0.079647738927210, {0.623592928761935, 0.188203535619033, 0}},
                     {3, 0.025577675658698, {0.910540973211095, 0.044729513394453, 0}},
                     {6, 0.043283539377289, {0.036838412054736, 0.221962989160766, 0.741198598784498}}};
const Orbit kD10[] = {{1, 0.090817990382754, {0, 0, 0}},
                      {3, 0.036725957756467, {0.028844733232685, 0.485577633383657, 0}},
                      {3, 0.045321059435528, {0.781036849029926, 0.109481575485037, 0}},
                      {6, 0.072757916845420, {0.141707219414880, 0.307939838764121, 0.550352941820999}},
                      {6, 0.028327242531057, {0.025003534762686, 0.246672560639903, 0.728323904597411}},
                      {6, 0.009421666963733, {0.009540815400299, 0.066803251012200, 0.923655933587500}}};

const Orbit *kTab[11] = {nullptr, kD1, kD2, kD3, kD4, kD5, kD6, kD7, kD8, kD9, kD10};
const int kLen[11] = {0, 1, 1, 2, 2, 3, 3, 4, 5, 6, 6};
}  // namespace

int triangle_rule(int degree, double *u, double *v, double *w)
{
    if (degree < 1 || degree > 10) return -1;
    int n = 0;
    for (int o = 0; o < kLen[degree]; ++o) {
        const Orbit &ob = kTab[degree][o];
        double bary[6][3];
        int m = 0;
        if (ob.kind == 1) {
            m = 1;
            bary[0][0] = bary[0][1] = bary[0][2] = 1.0 / 3.0;
        } else if (ob.kind == 3) {
            const double a = ob.c[0], b = ob.c[1];
            const double t[3][3] = {{a, b, b}, {b, a, b}, {b, b, a}};
            m = 3;
            std::memcpy(bary, t, sizeof t);
        } else {
            const double a = ob.c[0], b = ob.c[1], c = ob.c[2];
            const double t[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
            m = 6;
            std::memcpy(bary, t, sizeof t);
        }
        for (int k = 0; k < m; ++k) {
            u[n] = bary[k][0];
            v[n] = bary[k][1];
            w[n] = 0.5 * ob.w;
            ++n;
        }
    }
    return n;
}

static double legendre_p(int n, double z, double *dp)
{
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    *dp = n * (z * p1 - p0) / (z * z - 1.0);
    return p1;
}

void legendre01(int n, double *x, double *w)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            const double p = legendre_p(n, z, &dp);
            const double dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        legendre_p(n, z, &dp);
        x[n - 1 - i] = 0.5 * (z + 1.0);
        w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

}  // namespace bh
