"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Wraps liboracle.so (oracle.c, the FP64 restatement of the reference's
assembly) and adds the dense solve through LAPACK (scipy getrf/getrs, the
routines the reference reaches through LinearAlgebra, src/solvers/lusolver.jl:34).
Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_I32 = C.POINTER(C.c_int32)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_trgauss.argtypes = [C.c_int, _D, _D, _D]
        L.or_trgauss.restype = C.c_int
        L.or_legendre01.argtypes = [C.c_int, _D, _D]
        L.or_classify.argtypes = [_D, _D, C.c_double, C.c_int, _I, _I]
        L.or_classify.restype = C.c_int
        L.or_wiltonints.argtypes = [_D, _D, _D, _D, _D, _D, _D]
        L.or_block.argtypes = [C.c_int, _D, _D, _D, C.c_int, _I, _D, _D, _D, _I]
        L.or_block.restype = C.c_int
        L.or_block_strategy.argtypes = [C.c_int, _D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int, _D, _D, _D]
        L.or_block_strategy.restype = C.c_int
        L.or_assemble.argtypes = [C.c_int, _D, _D, _D, C.c_int, _I,
                                  _D, _I32, C.c_int, C.c_int, _I32, _I32, _I32, _D,
                                  _D, _I32, C.c_int, C.c_int, _I32, _I32, _I32, _D,
                                  _D, C.POINTER(C.c_int8)]
        L.or_assemble.restype = C.c_int
        L.or_planewave.argtypes = [C.c_int, _D, _I32, C.c_int, C.c_int, _I32, _I32, _I32, _D, _D, _D, _D, _D, _D,
                                   C.c_int, _D]
        L.or_planewave.restype = C.c_int
        L.or_localop.argtypes = [C.c_int, _D, _I32, C.c_int, C.c_int, _I32, _I32, _I32, _D, C.c_int, _I32, _I32,
                                 _I32, _D, C.c_int, _D, _D]
        L.or_localop.restype = C.c_int
        L.or_farfield.argtypes = [_D, _I32, C.c_int, C.c_int, _I32, _I32, _I32, _D, C.c_int, _D, _D, C.c_int, _D,
                                  _D, _D, C.c_int, _D]
        L.or_farfield.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _c2(z):
    z = complex(z)
    return np.array([z.real, z.imag])


def trgauss(degree):
    u, v, w = np.zeros(32), np.zeros(32), np.zeros(32)
    n = lib().or_trgauss(degree, _d(u), _d(v), _d(w))
    return np.stack([u[:n], v[:n]]), w[:n]


def legendre01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().or_legendre01(n, _d(x), _d(w))
    return x, w


def classify(tv, sv, k2, wilton=True):
    tv, sv = np.ascontiguousarray(tv, float).ravel(), np.ascontiguousarray(sv, float).ravel()
    ti, sj = (C.c_int * 3)(), (C.c_int * 3)()
    return lib().or_classify(_d(tv), _d(sv), float(k2), int(wilton), ti, sj)


def wiltonints(p1, p2, p3, x):
    """Analytic integrals as restated in the reference's test_wiltonints.jl:60-232."""
    s, v, g = np.zeros(6), np.zeros((5, 3)), np.zeros((4, 3))
    a = [np.ascontiguousarray(t, float) for t in (p1, p2, p3, x)]
    lib().or_wiltonints(*[_d(t) for t in a], _d(s), _d(v), _d(g))
    return s, v, g


def _opargs(kind, gamma, alpha, beta):
    return int(kind), _c2(gamma), _c2(alpha), _c2(beta)


def block(kind, gamma, alpha, beta, qs_kind, qs, tv, sv):
    """3x3 local block z[i, j] for one pair under a quadstrat; returns (z, class)."""
    k, g, a, b = _opargs(kind, gamma, alpha, beta)
    q = (C.c_int * 8)(*(list(qs) + [0] * (8 - len(qs))))
    tv, sv = np.ascontiguousarray(tv, float).ravel(), np.ascontiguousarray(sv, float).ravel()
    out = np.zeros(18)
    cls = C.c_int()
    rc = lib().or_block(k, _d(g), _d(a), _d(b), int(qs_kind), q, _d(tv), _d(sv), _d(out), C.byref(cls))
    _check_block(rc)
    z = (out[0::2] + 1j * out[1::2]).reshape(3, 3).T
    return z, cls.value


def block_strategy(kind, gamma, alpha, beta, strat, tv, sv, deg_o=6, deg_i=7, n=5):
    """Block with an explicit rule: strat 'dq', 'se' (Wilton) or 'ss' (Sauter-Schwab)."""
    k, g, a, b = _opargs(kind, gamma, alpha, beta)
    tv, sv = np.ascontiguousarray(tv, float).ravel(), np.ascontiguousarray(sv, float).ravel()
    out = np.zeros(18)
    s = {"dq": 0, "se": 1, "ss": 2}[strat]
    rc = lib().or_block_strategy(k, _d(g), _d(a), _d(b), s, deg_o, deg_i, n, _d(tv), _d(sv), _d(out))
    if rc == -3:
        raise NotImplementedError("no singularity extraction for the rotated double layer")
    if rc != 0:
        raise ValueError("strategy not applicable to this pair")
    return (out[0::2] + 1j * out[1::2]).reshape(3, 3).T


def _space_arrays(X):
    return (X.numfunctions(), np.ascontiguousarray(X.shape_ptr, np.int32), np.ascontiguousarray(X.shape_cell, np.int32),
            np.ascontiguousarray(X.shape_ref, np.int32), np.ascontiguousarray(X.shape_coeff, np.float64))


def assemble(kind, gamma, alpha, beta, qs_kind, qs, X, Y, want_classes=False):
    """Dense Z (column-major complex128) of the operator on test X, trial Y."""
    k, g, a, b = _opargs(kind, gamma, alpha, beta)
    q = (C.c_int * 8)(*(list(qs) + [0] * (8 - len(qs))))
    mt, ms = X.geo, Y.geo
    Vt, Ct = np.ascontiguousarray(mt.vertices, float), np.ascontiguousarray(mt.cells, np.int32)
    Vs, Cs = np.ascontiguousarray(ms.vertices, float), np.ascontiguousarray(ms.cells, np.int32)
    nt, tp, tc, tr, tco = _space_arrays(X)
    ns, sp, sc, sr, sco = _space_arrays(Y)
    Z = np.zeros((nt, ns), dtype=np.complex128, order="F")
    cls = np.zeros((mt.numcells, ms.numcells), dtype=np.int8) if want_classes else None
    i32 = lambda t: t.ctypes.data_as(_I32)
    rc = lib().or_assemble(k, _d(g), _d(a), _d(b), int(qs_kind), q,
                           _d(Vt), i32(Ct), mt.numcells, nt, i32(tp), i32(tc), i32(tr), _d(tco),
                           _d(Vs), i32(Cs), ms.numcells, ns, i32(sp), i32(sc), i32(sr), _d(sco),
                           Z.ctypes.data_as(_D), None if cls is None else cls.ctypes.data_as(C.POINTER(C.c_int8)))
    _check_block(rc)
    return (Z, cls) if want_classes else Z


def _check_block(rc):
    if rc == -3:
        # the reference has no regularpart/singularpart for DoubleLayerRotatedMW3D
        raise NotImplementedError("DoubleNumWiltonSauterQStrat does not apply to the rotated double layer")
    if rc != 0:
        raise AssertionError("hits <= 3")


def ncross(X, Y, degree=6, scale=1.0):
    """Dense NCross matrix (test X, trial Y on the same mesh), column-major complex128."""
    return localop(0, X, Y, degree, scale)


def identity(X, Y, degree=6, scale=1.0):
    """Dense Identity (RT x RT Gram) matrix on one mesh, column-major complex128."""
    return localop(1, X, Y, degree, scale)


def localop(kind, X, Y, degree=6, scale=1.0):
    """Local operator on RT x RT over one mesh: kind 0 NCross, 1 Identity."""
    if X.geo is not Y.geo:
        raise ValueError("local operator oracle: test and trial spaces must share the mesh")
    m = X.geo
    V, Cc = np.ascontiguousarray(m.vertices, float), np.ascontiguousarray(m.cells, np.int32)
    nt, tp, tc, tr, tco = _space_arrays(X)
    ns, sp, sc, sr, sco = _space_arrays(Y)
    Z = np.zeros((nt, ns), dtype=np.complex128, order="F")
    i32 = lambda t: t.ctypes.data_as(_I32)
    rc = lib().or_localop(int(kind), _d(V), i32(Cc), m.numcells, nt, i32(tp), i32(tc), i32(tr), _d(tco),
                          ns, i32(sp), i32(sc), i32(sr), _d(sco), int(degree), _d(_c2(scale)), Z.ctypes.data_as(_D))
    if rc != 0:
        raise ValueError("bad rule degree")
    return Z


def planewave(X, direction, polarization, gamma, amplitude=1.0, degree=8, trace="tangential"):
    """Plane wave tested with X: trace "tangential" ((n x E) x n) or "cross" (n x E)."""
    m = X.geo
    V, Cc = np.ascontiguousarray(m.vertices, float), np.ascontiguousarray(m.cells, np.int32)
    n, p, c, r, co = _space_arrays(X)
    pol = np.asarray(polarization, dtype=np.complex128)
    b = np.zeros(2 * n)
    i32 = lambda t: t.ctypes.data_as(_I32)
    rc = lib().or_planewave({"tangential": 0, "cross": 1}[trace], _d(V), i32(Cc), m.numcells, n, i32(p), i32(c), i32(r), _d(co),
                            _d(np.ascontiguousarray(direction, float)), _d(np.ascontiguousarray(pol.real)),
                            _d(np.ascontiguousarray(pol.imag)), _d(_c2(gamma)), _d(_c2(amplitude)), int(degree), _d(b))
    if rc != 0:
        raise ValueError("bad rule degree")
    return b[0::2] + 1j * b[1::2]


def farfield(kind, gamma, amplitude, points, coeffs, X, degree=10):
    """potential(op, points, coeffs, X) for the far-field operators (postproc.jl:174-211,273-288,
    maxwell/farfield.jl:56-75,98-100); kind 0 MWFarField3D, 1 MWDoubleLayerFarField3D,
    2 MWDoubleLayerRotatedFarField3D, 3 MWFarField3DDropConstant.  Returns [npoints, 3] complex."""
    m = X.geo
    V, Cc = np.ascontiguousarray(m.vertices, float), np.ascontiguousarray(m.cells, np.int32)
    n, p, c, r, co = _space_arrays(X)
    pts = np.ascontiguousarray(np.asarray(points, float).reshape(-1, 3))
    u = np.asarray(coeffs, dtype=np.complex128)
    cre, cim = np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag)
    ff = np.zeros(6 * len(pts))
    i32 = lambda t: t.ctypes.data_as(_I32)
    rc = lib().or_farfield(_d(V), i32(Cc), m.numcells, n, i32(p), i32(c), i32(r), _d(co), int(kind),
                           _d(_c2(gamma)), _d(_c2(amplitude)), len(pts), _d(pts), _d(cre), _d(cim), int(degree),
                           _d(ff))
    if rc != 0:
        raise ValueError("bad rule degree or far-field kind")
    return (ff[0::2] + 1j * ff[1::2]).reshape(-1, 3)


def lu_solve(Z, B):
    """LAPACK getrf + getrs (the reference's Z \\ b)."""
    import scipy.linalg as sla
    lu, piv = sla.lu_factor(Z, check_finite=False)
    return sla.lu_solve((lu, piv), B, check_finite=False)
