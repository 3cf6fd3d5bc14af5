"""Quadrature rules used by the Maxwell/RT assembly path.

Triangle rules: symmetric Dunavant rules (D. A. Dunavant, "High degree efficient
symmetrical Gaussian quadrature rules for the triangle", IJNME 21, 1985), indexed
by polynomial degree as ``trgauss(n)`` is in the reference quadrature strategies
(``src/quadrature/doublenumwiltonsauterqstrat.jl:12-29`` asks for degrees 2, 3, 6, 7;
``src/excitation.jl:6`` asks for 8).  Points are returned in the chart parameters
``(u, v)`` of ``r(u,v) = u*v1 + v*v2 + (1-u-v)*v3`` and weights sum to 1/2 (the
area of the reference simplex), so ``weight * 2*area`` is the physical weight.

Gauss-Legendre rules on [0, 1] follow ``_legendre(n, 0, 1)`` in
``src/maxwell/mwops.jl:46-51``.

The same tables are compiled into the HIP library (``csrc/rules.h``); a test
checks that both copies agree bit for bit.
"""
import numpy as np

# Orbit form: list of (kind, weight, coords) with kind 1 = centroid,
# 3 = (a, b, b) and permutations, 6 = (a, b, c) and permutations.
# Weights are normalised to sum to 1 over the triangle.
_DUNAVANT = {
    1: [(1, 1.0, ())],
    2: [(3, 0.333333333333333, (0.666666666666667, 0.166666666666667))],
    3: [(1, -0.5625, ()),
        (3, 0.520833333333333, (0.6, 0.2))],
    4: [(3, 0.223381589678011, (0.108103018168070, 0.445948490915965)),
        (3, 0.109951743655322, (0.816847572980459, 0.091576213509771))],
    5: [(1, 0.225, ()),
        (3, 0.132394152788506, (0.059715871789770, 0.470142064105115)),
        (3, 0.125939180544827, (0.797426985353087, 0.101286507323456))],
    6: [(3, 0.116786275726379, (0.501426509658179, 0.249286745170910)),
        (3, 0.050844906370207, (0.873821971016996, 0.063089014491502)),
        (6, 0.082851075618374, (0.053145049844817, 0.310352451033784, 0.636502499121399))],
    7: [(1, -0.149570044467682, ()),
        (3, 0.175615257433208, (0.479308067841920, 0.260345966079040)),
        (3, 0.053347235608838, (0.869739794195568, 0.065130102902216)),
        (6, 0.077113760890257, (0.048690315425316, 0.312865496004874, 0.638444188569810))],
    8: [(1, 0.144315607677787, ()),
        (3, 0.095091634267285, (0.081414823414554, 0.459292588292723)),
        (3, 0.103217370534718, (0.658861384496480, 0.170569307751760)),
        (3, 0.032458497623198, (0.898905543365938, 0.050547228317031)),
        (6, 0.027230314174435, (0.008394777409958, 0.263112829634638, 0.728492392955404))],
    9: [(1, 0.097135796282799, ()),
        (3, 0.031334700227139, (0.020634961602525, 0.489682519198738)),
        (3, 0.077827541004774, (0.125820817014127, 0.437089591492937)),
        (3, 0.079647738927210, (0.623592928761935, 0.188203535619033)),
        (3, 0.025577675658698, (0.910540973211095, 0.044729513394453)),
        (6, 0.043283539377289, (0.036838412054736, 0.221962989160766, 0.741198598784498))],
    10: [(1, 0.090817990382754, ()),
         (3, 0.036725957756467, (0.028844733232685, 0.485577633383657)),
         (3, 0.045321059435528, (0.781036849029926, 0.109481575485037)),
         (6, 0.072757916845420, (0.141707219414880, 0.307939838764121, 0.550352941820999)),
         (6, 0.028327242531057, (0.025003534762686, 0.246672560639903, 0.728323904597411)),
         (6, 0.009421666963733, (0.009540815400299, 0.066803251012200, 0.923655933587500))],
}

MAX_TRIANGLE_DEGREE = max(_DUNAVANT)


def _expand(orbits):
    pts = []
    for kind, w, c in orbits:
        if kind == 1:
            bary = [(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0)]
        elif kind == 3:
            a, b = c
            bary = [(a, b, b), (b, a, b), (b, b, a)]
        else:
            a, b, cc = c
            bary = [(a, b, cc), (a, cc, b), (b, a, cc), (b, cc, a), (cc, a, b), (cc, b, a)]
        for l1, l2, _l3 in bary:
            pts.append((l1, l2, 0.5 * w))
    return np.array(pts, dtype=np.float64)


def trgauss(degree):
    """Return ``(uv[2, Q], w[Q])`` for the Dunavant rule of the given degree.

    Mirrors the ``trgauss(n)`` call made by ``quadpoints`` for the reference's
    triangle rules; weights sum to 1/2.
    """
    if degree not in _DUNAVANT:
        raise ValueError(f"no triangle rule of degree {degree} (have 1..{MAX_TRIANGLE_DEGREE})")
    t = _expand(_DUNAVANT[degree])
    return t[:, :2].T.copy(), t[:, 2].copy()


def legendre01(n):
    """Gauss-Legendre rule on [0, 1], as ``_legendre(n, 0, 1)`` (mwops.jl:46-51)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, w / 2.0
