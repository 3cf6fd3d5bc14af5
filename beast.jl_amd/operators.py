"""Operators, excitations and quadrature strategies of the Maxwell/RT path.

Mirrors the reference's constructors and parameter handling:
``Maxwell3D.singlelayer/doublelayer/planewave/farfield`` (src/maxwell/maxwell.jl:16-107),
``MWSingleLayer3D``/``MWDoubleLayer3D`` (src/maxwell/mwops.jl:11-64),
``DoubleLayerRotatedMW3D`` (src/maxwell/nxdbllayer.jl:17-44), the local operators
``Identity`` and ``NCross`` (src/identityop.jl:6-22),
``gamma_wavenumber_handler`` (mwops.jl:213-238), the quadrature strategies
(src/quadrature/doublenumwiltonsauterqstrat.jl:1-10, doublenumsauterqstrat.jl:1-8,
doublenumqstrat.jl:1-4), ``defaultquadstrat`` (mwops.jl:55), and the traces
``(n x E) x n`` and ``n x E`` (src/excitation.jl:145-180) of plane waves and their curl
(src/maxwell/mwexc.jl:1-54).
"""
import numbers

import numpy as np


# ----------------------------------------------------------------- quadstrats
class AbstractQuadStrat:
    """A quadstrat object is also a function (op, X, Y) -> itself (strategies/quadstrat.jl:3-9)."""

    def __call__(self, *args):
        return self


class DoubleNumWiltonSauterQStrat(AbstractQuadStrat):
    def __init__(self, outer_rule_far, inner_rule_far, outer_rule_near, inner_rule_near,
                 sauter_schwab_common_tetr, sauter_schwab_common_face, sauter_schwab_common_edge,
                 sauter_schwab_common_vert):
        self.params = (outer_rule_far, inner_rule_far, outer_rule_near, inner_rule_near,
                       sauter_schwab_common_tetr, sauter_schwab_common_face, sauter_schwab_common_edge,
                       sauter_schwab_common_vert)

    kind = 0

    def __repr__(self):
        return "DoubleNumWiltonSauterQStrat%s" % (self.params,)


class DoubleNumSauterQstrat(AbstractQuadStrat):
    def __init__(self, outer_rule, inner_rule, sauter_schwab_common_tetr, sauter_schwab_common_face,
                 sauter_schwab_common_edge, sauter_schwab_common_vert):
        self.params = (outer_rule, inner_rule, sauter_schwab_common_tetr, sauter_schwab_common_face,
                       sauter_schwab_common_edge, sauter_schwab_common_vert)

    kind = 1

    def __repr__(self):
        return "DoubleNumSauterQstrat%s" % (self.params,)


class DoubleNumQStrat(AbstractQuadStrat):
    def __init__(self, outer_rule, inner_rule):
        self.params = (outer_rule, inner_rule)

    kind = 2

    def __repr__(self):
        return "DoubleNumQStrat%s" % (self.params,)


class SingleNumQStrat(AbstractQuadStrat):
    kind = 3

    def __init__(self, quad_rule):
        self.quad_rule = quad_rule
        self.params = (quad_rule,)

    def __repr__(self):
        return "SingleNumQStrat(%d)" % self.quad_rule


def defaultquadstrat(op, X=None, Y=None):
    """defaultquadstrat for Maxwell operators on RT x RT (mwops.jl:55), for local
    operators on RT x RT (identityop.jl:66-67), for functionals (excitation.jl:6)
    and for far-field operators (postproc/farfield.jl:3)."""
    if isinstance(op, FarField):
        return SingleNumQStrat(10)
    if isinstance(op, Functional):
        return SingleNumQStrat(8)
    if isinstance(op, LocalOperator):
        return SingleNumQStrat(6)
    if isinstance(op, DoubleLayerRotatedMW3D):
        return DoubleNumQStrat(6, 7)   # nxdbllayer.jl:23
    return DoubleNumWiltonSauterQStrat(2, 3, 6, 7, 5, 5, 4, 3)


def resolve_quadstrat(quadstrat, op, X, Y):
    if isinstance(quadstrat, AbstractQuadStrat):
        return quadstrat
    if callable(quadstrat):
        return quadstrat(op, X, Y)
    raise TypeError("quadstrat must be a quadrature strategy or a function (op, X, Y) -> strategy")


# ----------------------------------------------------------------- operators
def gamma_wavenumber_handler(gamma, wavenumber):
    """mwops.jl:213-238; returns (gamma, wavenumber), (None, None) for the static case."""
    if gamma is not None and wavenumber is not None:
        raise ValueError("Supplying both gamma and wavenumber is not supported.")
    if gamma is None and wavenumber is None:
        return None, None
    if gamma is None:
        k = complex(wavenumber)
        gamma = -k.imag if k.real == 0 else 1j * k
        return gamma, wavenumber
    g = complex(gamma)
    wavenumber = g.imag if g.real == 0 else -1j * g
    return gamma, wavenumber


class IntegralOperator:
    def __mul__(self, a):
        return LinearCombinationOfOperators([(complex(a), self)])

    __rmul__ = __mul__

    def __add__(self, other):
        return LinearCombinationOfOperators([(1.0, self)]) + other

    def __neg__(self):
        return self * -1

    def __sub__(self, other):
        return self + (-1.0) * other


class MWSingleLayer3D(IntegralOperator):
    """EFIE kernel G (alpha f.g + beta div f div g), G = exp(-gamma R)/(4 pi R)."""

    def __init__(self, gamma, alpha=None, beta=None):
        g = 0.0 if gamma is None else complex(gamma)
        self.gamma = g
        self.alpha = -g if alpha is None else complex(alpha)
        if beta is None:
            if g == 0:
                raise ValueError("static single layer needs explicit alpha and beta")
            beta = -1.0 / g
        self.beta = complex(beta)

    kind = 0

    def __repr__(self):
        return f"MWSingleLayer3D(gamma={self.gamma}, alpha={self.alpha}, beta={self.beta})"


class MWDoubleLayer3D(IntegralOperator):
    """MFIE-K kernel f . (grad G x g); like the reference, alpha is not used by the kernel."""

    def __init__(self, alpha, gamma):
        self.alpha = complex(alpha)
        self.gamma = 0.0 if gamma is None else complex(gamma)

    kind = 1

    def __repr__(self):
        return f"MWDoubleLayer3D(alpha={self.alpha}, gamma={self.gamma})"


class DoubleLayerRotatedMW3D(IntegralOperator):
    """n x K (nxdbllayer.jl:17-44): integrand f . (n x (grad G x g)), n the test cell's
    normal; like the reference's integrand, alpha is not used by the kernel."""

    kind = 3

    def __init__(self, alpha, gamma):
        self.alpha = complex(alpha)
        self.gamma = complex(gamma)

    def __repr__(self):
        return f"DoubleLayerRotatedMW3D(alpha={self.alpha}, gamma={self.gamma})"


class LocalOperator(IntegralOperator):
    """Operators whose matrix couples only functions sharing a cell (localop.jl:4)."""


class NCross(LocalOperator):
    """NCross (identityop.jl:14-22): the identity with the trial function rotated,
    integrand g . (n x f) for test g and trial f."""

    kind = 2

    def __repr__(self):
        return "NCross()"


class Identity(LocalOperator):
    """Identity (identityop.jl:6-11): integrand f . g on the common cells."""

    kind = 4

    def __repr__(self):
        return "Identity()"


class LinearCombinationOfOperators(IntegralOperator):
    """Sum of scaled operators (operator.jl:294-306): assembled term by term into one store."""

    def __init__(self, terms):
        self.terms = list(terms)

    def __mul__(self, a):
        return LinearCombinationOfOperators([(c * complex(a), op) for c, op in self.terms])

    __rmul__ = __mul__

    def __sub__(self, other):
        return self + (-1.0) * other

    def __add__(self, other):
        if isinstance(other, LinearCombinationOfOperators):
            return LinearCombinationOfOperators(self.terms + other.terms)
        if isinstance(other, IntegralOperator):
            return LinearCombinationOfOperators(self.terms + [(1.0, other)])
        return NotImplemented


class _Maxwell3D:
    """Namespace mirroring ``Maxwell3D`` (src/maxwell/maxwell.jl)."""

    @staticmethod
    def singlelayer(gamma=None, wavenumber=None, alpha=None, beta=None):
        gamma, wavenumber = gamma_wavenumber_handler(gamma, wavenumber)
        if gamma is None:
            if alpha is None or beta is None:
                raise AssertionError("static single layer needs alpha and beta")
            return MWSingleLayer3D(0.0, alpha, beta)
        g = complex(gamma)
        return MWSingleLayer3D(g, -g if alpha is None else alpha, -1.0 / g if beta is None else beta)

    @staticmethod
    def weaklysingular(wavenumber):
        return _Maxwell3D.singlelayer(wavenumber=wavenumber, alpha=-1j * wavenumber, beta=0.0)

    @staticmethod
    def hypersingular(wavenumber):
        return _Maxwell3D.singlelayer(wavenumber=wavenumber, alpha=0.0, beta=-1.0 / (1j * wavenumber))

    @staticmethod
    def doublelayer(alpha=None, gamma=None, wavenumber=None):
        gamma, wavenumber = gamma_wavenumber_handler(gamma, wavenumber)
        if alpha is None:
            alpha = 1.0
        return MWDoubleLayer3D(alpha, 0.0 if gamma is None else gamma)

    @staticmethod
    def planewave(direction, polarization, wavenumber, amplitude=1.0):
        return PlaneWaveMW(direction, polarization, 1j * wavenumber, amplitude)

    @staticmethod
    def farfield(wavenumber):
        """Maxwell3D.farfield(; wavenumber) (maxwell.jl:105-107)."""
        from .postproc import MWFarField3D
        return MWFarField3D(wavenumber=wavenumber)


Maxwell3D = _Maxwell3D()


# ----------------------------------------------------------------- excitations
class PlaneWaveMW:
    """E(x) = amplitude * exp(-gamma d.x) * polarisation (mwexc.jl:36-42)."""

    def __init__(self, direction, polarisation, gamma, amplitude=1.0):
        self.direction = np.asarray(direction, dtype=np.float64)
        self.polarisation = np.asarray(polarisation, dtype=np.complex128)
        self.gamma = complex(gamma)
        self.amplitude = complex(amplitude)

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        return self.amplitude * np.exp(-self.gamma * (x @ self.direction))[..., None] * self.polarisation

    def __rmul__(self, a):
        """a * E (mwexc.jl:54)."""
        if not isinstance(a, numbers.Number):
            return NotImplemented
        return PlaneWaveMW(self.direction, self.polarisation, self.gamma, complex(a) * self.amplitude)


def curl(field):
    """curl of a plane wave: PlaneWaveMW(d, d x u, gamma, -a gamma) (mwexc.jl:44-52)."""
    if not isinstance(field, PlaneWaveMW):
        raise TypeError("curl is defined for plane waves")
    return PlaneWaveMW(field.direction, np.cross(field.direction, field.polarisation), field.gamma,
                       -field.amplitude * field.gamma)


class NormalVector:
    """The symbol ``n`` of the reference (cross(n, E) builds traces)."""


n = NormalVector()


class Functional:
    pass


class FarField:
    """Far-field post-processing operators (postproc/farfield.jl:1); the concrete
    Maxwell ones are in the postproc module."""


class CrossTraceMW(Functional):
    def __init__(self, field):
        self.field = field


class TangTraceMW(Functional):
    """(n x E) x n: tested with tangential RT functions it equals f . E."""

    def __init__(self, field):
        self.field = field


def cross(a, b):
    """``cross(n, E) -> CrossTraceMW``, ``cross(CrossTraceMW, n) -> TangTraceMW`` (excitation.jl:166-167)."""
    if isinstance(a, NormalVector) and isinstance(b, FarField):
        return b._ncross()
    if isinstance(a, NormalVector) and isinstance(b, MWDoubleLayer3D):
        return DoubleLayerRotatedMW3D(b.alpha, b.gamma)   # nxdbllayer.jl:25
    if isinstance(a, NormalVector) and not isinstance(b, NormalVector):
        return CrossTraceMW(b)
    if isinstance(a, CrossTraceMW) and isinstance(b, NormalVector):
        return TangTraceMW(a.field)
    if isinstance(a, numbers.Number) or isinstance(b, numbers.Number):
        raise TypeError("cross of scalars")
    return np.cross(a, b)


x_hat = np.array([1.0, 0.0, 0.0])
y_hat = np.array([0.0, 1.0, 0.0])
z_hat = np.array([0.0, 0.0, 1.0])
