"""Far-field post-processing of a surface current on a Raviart-Thomas space.

Mirrors the reference's far-field operators (src/maxwell/farfield.jl:1-120) and
``potential(op, points, coeffs, basis; quadstrat)`` (src/postproc.jl:138-150,174-211);
the sum over cells and rule points runs in the native library (``bh_farfield``).
"""
import cmath
import ctypes as C

import numpy as np

from . import _native as nat
from .operators import (FarField, MWDoubleLayer3D, MWSingleLayer3D, SingleNumQStrat, defaultquadstrat,
                        gamma_wavenumber_handler, resolve_quadstrat)
from .rt import RTBasis

BH_FF_SINGLELAYER = 0
BH_FF_DOUBLELAYER = 1
BH_FF_DOUBLELAYER_ROTATED = 2
BH_FF_DROPCONSTANT = 3


class MWFarField(FarField):
    """Maxwell far-field operator with fields ``gamma`` and ``amplitude`` (farfield.jl:1-19)."""

    kind = None

    def __init__(self, gamma, amplitude):
        self.gamma = complex(gamma)
        self.amplitude = complex(amplitude)

    def __repr__(self):
        return f"{type(self).__name__}(gamma={self.gamma}, amplitude={self.amplitude})"


def _keyword_form(cls, args, gamma, wavenumber, amplitude):
    """The keyword constructors (farfield.jl:22-33,42-53,82-92): gamma from
    gamma_wavenumber_handler, static problems rejected, amplitude 1 by default."""
    if args:
        raise TypeError(f"{cls.__name__} takes (gamma, amplitude), an operator, or keywords")
    g, _ = gamma_wavenumber_handler(gamma, wavenumber)
    if g is None:
        raise AssertionError(f"{cls.__name__} needs gamma or wavenumber (static problems have no far field)")
    return g, 1.0 if amplitude is None else amplitude


class MWFarField3D(MWFarField):
    """Single-layer far field: integrand amplitude (y x (K J)) x y, K = exp(gamma y.x)
    (farfield.jl:3-6,22-35,59-63).

    ``MWFarField3D(gamma=, wavenumber=, amplitude=)``; ``MWFarField3D(op)`` for a
    ``MWSingleLayer3D`` op takes its gamma and amplitude sqrt(alpha beta) (farfield.jl:35);
    ``MWFarField3D(gamma, amplitude)`` is the field constructor, and ``MWFarField3D(gamma)``
    takes amplitude 1."""

    kind = BH_FF_SINGLELAYER

    def __init__(self, *args, gamma=None, wavenumber=None, amplitude=None):
        if len(args) == 1 and isinstance(args[0], MWSingleLayer3D):
            op = args[0]
            super().__init__(op.gamma, cmath.sqrt(op.alpha * op.beta))
        elif args and gamma is None and wavenumber is None and amplitude is None:
            super().__init__(args[0], args[1] if len(args) > 1 else 1.0)
        else:
            super().__init__(*_keyword_form(MWFarField3D, args, gamma, wavenumber, amplitude))


class MWDoubleLayerFarField3D(MWFarField):
    """Double-layer far field: integrand amplitude y x (K J) (farfield.jl:7-10,42-55,65-67).
    ``MWDoubleLayerFarField3D(op)`` for a ``MWDoubleLayer3D`` op takes its gamma and amplitude 1."""

    kind = BH_FF_DOUBLELAYER

    def __init__(self, *args, gamma=None, wavenumber=None, amplitude=None):
        if len(args) == 1 and isinstance(args[0], MWDoubleLayer3D):
            super().__init__(args[0].gamma, 1.0)
        elif args and gamma is None and wavenumber is None and amplitude is None:
            super().__init__(args[0], args[1] if len(args) > 1 else 1.0)
        else:
            super().__init__(*_keyword_form(MWDoubleLayerFarField3D, args, gamma, wavenumber, amplitude))

    def _ncross(self):
        """n x MWDoubleLayerFarField3D -> MWDoubleLayerRotatedFarField3D (farfield.jl:102)."""
        return MWDoubleLayerRotatedFarField3D(self.gamma, self.amplitude)


class MWDoubleLayerRotatedFarField3D(MWFarField):
    """Rotated double-layer far field: integrand amplitude y x ((K J) x n), n the cell
    normal (farfield.jl:12-15,82-100)."""

    kind = BH_FF_DOUBLELAYER_ROTATED

    def __init__(self, *args, gamma=None, wavenumber=None, amplitude=None):
        if args and gamma is None and wavenumber is None and amplitude is None:
            super().__init__(args[0], args[1] if len(args) > 1 else 1.0)
        else:
            super().__init__(*_keyword_form(MWDoubleLayerRotatedFarField3D, args, gamma, wavenumber, amplitude))

    def _ncross(self):
        """n x MWDoubleLayerRotatedFarField3D -> MWDoubleLayerFarField3D with the amplitude
        negated (farfield.jl:103)."""
        return MWDoubleLayerFarField3D(self.gamma, -self.amplitude)


class MWFarField3DDropConstant(MWFarField):
    """Single-layer far field with the constant of the kernel dropped: K = expm1(gamma y.x),
    integrand coeff (y x (K J)) x y (farfield.jl:69-75).  Field constructor (gamma, coeff)."""

    kind = BH_FF_DROPCONSTANT

    def __init__(self, gamma, coeff):
        super().__init__(gamma, coeff)

    @property
    def coeff(self):
        return self.amplitude


def potential(op, points, coeffs, basis, quadstrat=defaultquadstrat, ctx=None):
    """``potential(op, points, coeffs, basis; quadstrat)`` (postproc.jl:138-150) for the
    far-field operators: for every point y (a direction; used as given),
    ff[y] = sum over the cells and rule points x of the operator's integrand
    times the rule weight, with J(x) = sum_n coeffs[n] f_n(x) (postproc.jl:174-211,273-288).

    Returns complex128 of shape ``points.shape[:-1] + (3,)`` (the reference returns an
    array of SVector{3,ComplexF64} of the points' shape).  The default rule is
    SingleNumQStrat(10) (postproc/farfield.jl:3); triangle rules of degree 1..10."""
    from .assembly import default_context, space_handle
    if not isinstance(op, MWFarField):
        raise nat.UnsupportedOnGPU(nat.BH_ERR_UNSUPPORTED, f"potential: {type(op).__name__} is not a far-field operator")
    if not isinstance(basis, RTBasis):
        raise nat.UnsupportedOnGPU(nat.BH_ERR_UNSUPPORTED, "potential: basis must be a raviartthomas space")
    qs = resolve_quadstrat(quadstrat, op, basis, None)
    if not isinstance(qs, SingleNumQStrat):
        raise ValueError("far fields take a SingleNumQStrat")
    pts = np.asarray(points, dtype=np.float64)
    if pts.shape[-1:] != (3,):
        raise ValueError("points must be 3-vectors")
    u = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128).reshape(-1))
    if u.size != basis.numfunctions():
        raise AssertionError("length(coeffs) == numfunctions(basis)")
    ctx = ctx or default_context()
    lib = nat.load()
    flat = np.ascontiguousarray(pts.reshape(-1, 3))
    ff = np.zeros((len(flat), 3), dtype=np.complex128)
    g = np.array([op.gamma.real, op.gamma.imag])
    a = np.array([op.amplitude.real, op.amplitude.imag])
    nat.check(lib.bh_farfield(ctx.handle, op.kind, nat.dptr(g), nat.dptr(a), space_handle(basis, ctx),
                              u.ctypes.data_as(nat._D), len(flat), nat.dptr(flat), int(qs.quad_rule),
                              ff.ctypes.data_as(nat._D)))
    return ff.reshape(pts.shape[:-1] + (3,))
