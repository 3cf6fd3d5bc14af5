"""``assemble(op, X, Y; quadstrat)`` on the GPU, through the C ABI.

Host-side mirror of src/operator.jl:100-116 (``assemble``) and
src/excitation.jl:17-30 (``assemble(field, tfs)``): operators, spaces and
quadrature strategies are resolved here, and the native library does all the
arithmetic.  There is no CPU path: unsupported inputs raise.
"""
import atexit
import ctypes as C
import threading
import weakref

import numpy as np

from . import _native as nat
from .operators import (CrossTraceMW, DoubleLayerRotatedMW3D, DoubleNumQStrat, DoubleNumSauterQstrat,
                        DoubleNumWiltonSauterQStrat, Identity, LinearCombinationOfOperators, MWDoubleLayer3D,
                        MWSingleLayer3D, NCross, PlaneWaveMW, SingleNumQStrat, TangTraceMW, defaultquadstrat,
                        resolve_quadstrat)
from .rt import RTBasis


class Context:
    """One GPU and one stream (replaces the reference's Threading{:multi} driver).

    A context is not re-entrant: ``lock`` serialises the calls of several threads on it
    (the calls of this module take it), e.g. a solve on a matrix another thread factors."""

    def __init__(self, device=0):
        lib = nat.load()
        h = C.c_void_p()
        nat.check(lib.bh_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device
        self._lib = lib
        self.lock = threading.RLock()
        _live["contexts"].add(self)

    def synchronize(self):
        with self.lock:
            nat.check(self._lib.bh_ctx_synchronize(self.handle))

    def profile(self, enable=True, only=None):
        """Bracket library kernel launches with HIP events (all of them, or only kernel ``only``)."""
        nat.check(self._lib.bh_ctx_profile_only(self.handle, (only or "").encode()))
        nat.check(self._lib.bh_ctx_profile(self.handle, int(bool(enable))))

    def pool(self, limit_bytes=-1):
        """Set the most bytes of freed matrix buffers the context keeps for reuse (-1 leaves
        it); returns the bytes it holds."""
        held = C.c_int64()
        nat.check(self._lib.bh_ctx_pool(self.handle, int(limit_bytes), C.byref(held)))
        return held.value

    def kernel_stats(self, name):
        """(launches, device seconds, algorithmic work) of one kernel since profiling was enabled."""
        n = C.c_int64()
        sec = C.c_double()
        work = C.c_double()
        nat.check(self._lib.bh_ctx_kernel_stats(self.handle, name.encode(), C.byref(n), C.byref(sec), C.byref(work)))
        return n.value, sec.value, work.value

    def close(self):
        if self.handle:
            self._lib.bh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None

# Live native objects, released in dependency order (matrices, spaces, meshes, then
# contexts) by an exit handler that runs before the interpreter's teardown, so that no
# finaliser calls into the HIP runtime while or after the runtime shuts down.
_live = {"matrices": weakref.WeakSet(), "handles": weakref.WeakSet(), "contexts": weakref.WeakSet()}


@atexit.register
def _release_native_objects():
    for M in list(_live["matrices"]):
        M.__del__()
    for h in sorted(list(_live["handles"]), key=lambda h: h.order):
        h.__del__()
    for c in list(_live["contexts"]):
        c.close()


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class _Handle:
    """A mesh (order 1) or space (order 0) handle; spaces are released before their meshes."""

    def __init__(self, h, destroy, keep, order):
        self.h = h
        self._destroy = destroy
        self._keep = keep
        self.order = order
        _live["handles"].add(self)

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
                self.h = None
        except Exception:
            pass


def mesh_handle(mesh, ctx):
    cache = mesh.__dict__.setdefault("_bh", {})
    key = id(ctx)
    if key not in cache:
        lib = nat.load()
        V = nat.as_f64(mesh.vertices)
        Cc = nat.as_i32(mesh.cells)
        h = C.c_void_p()
        with ctx.lock:
            nat.check(lib.bh_mesh_create(ctx.handle, nat.dptr(V), V.shape[0], nat.i32ptr(Cc), Cc.shape[0],
                                         C.byref(h)))
        cache[key] = _Handle(h, lib.bh_mesh_destroy, ctx, 1)
    return cache[key].h


def space_handle(space, ctx):
    cache = space.__dict__.setdefault("_bh", {})
    key = id(ctx)
    if key not in cache:
        lib = nat.load()
        mh = mesh_handle(space.geo, ctx)
        h = C.c_void_p()
        with ctx.lock:
            nat.check(lib.bh_space_create(mh, space.numfunctions(), nat.i32ptr(space.shape_ptr),
                                          nat.i32ptr(space.shape_cell), nat.i32ptr(space.shape_ref),
                                          nat.dptr(space.shape_coeff), C.byref(h)))
        cache[key] = _Handle(h, lib.bh_space_destroy, (ctx, mh), 0)
    return cache[key].h


def op_struct(op):
    s = nat.bh_operator()
    if isinstance(op, MWSingleLayer3D):
        s.kind = nat.OP_SINGLELAYER
        s.alpha[:] = [op.alpha.real, op.alpha.imag]
        s.beta[:] = [op.beta.real, op.beta.imag]
    elif isinstance(op, MWDoubleLayer3D):
        s.kind = nat.OP_DOUBLELAYER
        s.alpha[:] = [op.alpha.real, op.alpha.imag]
        s.beta[:] = [0.0, 0.0]
    elif isinstance(op, DoubleLayerRotatedMW3D):
        s.kind = nat.OP_DOUBLELAYER_ROTATED
        s.alpha[:] = [op.alpha.real, op.alpha.imag]
        s.beta[:] = [0.0, 0.0]
    elif isinstance(op, NCross):
        s.kind = nat.OP_NCROSS
        return s
    elif isinstance(op, Identity):
        s.kind = nat.OP_IDENTITY
        return s
    else:
        raise nat.UnsupportedOnGPU(nat.BH_ERR_UNSUPPORTED, f"operator {type(op).__name__} has no GPU path")
    g = complex(op.gamma)
    s.gamma[:] = [g.real, g.imag]
    return s


def qs_struct(qs):
    s = nat.bh_quadstrat()
    if isinstance(qs, (DoubleNumWiltonSauterQStrat, DoubleNumSauterQstrat, DoubleNumQStrat, SingleNumQStrat)):
        s.kind = qs.kind
        for i, v in enumerate(qs.params):
            s.params[i] = int(v)
    else:
        raise nat.UnsupportedOnGPU(nat.BH_ERR_UNSUPPORTED, f"quadrature strategy {qs!r} has no GPU path")
    return s


def _terms(op):
    if isinstance(op, LinearCombinationOfOperators):
        return op.terms
    return [(1.0, op)]


def assemble(op, X, Y=None, quadstrat=defaultquadstrat, precision="fp64", ctx=None, threading="gpu", out=None,
             row_offset=0, col_offset=0, **kwargs):
    """Dense matrix of ``op`` tested with ``X`` and trial ``Y`` (column-major complex).

    With ``Y`` omitted and ``op`` a functional, assembles the right-hand side
    (``assemble(e, X)``).  ``out``/offsets give the ``assemble!`` store semantics:
    contributions are added into ``out[row_offset:, col_offset:]``.
    """
    ctx = ctx or default_context()
    if Y is None:
        return assemble_rhs(op, X, quadstrat=quadstrat, ctx=ctx)
    if not isinstance(X, RTBasis) or not isinstance(Y, RTBasis):
        raise nat.UnsupportedOnGPU(nat.BH_ERR_UNSUPPORTED, "GPU assembly supports raviartthomas spaces only")
    prec = {"fp64": nat.FP64, "fp32": nat.FP32}[precision]
    dt = np.complex128 if prec == nat.FP64 else np.complex64
    M, N = X.numfunctions(), Y.numfunctions()
    if out is None:
        out = np.zeros((M, N), dtype=dt, order="F")
        row_offset = col_offset = 0
    if out.dtype != dt or not out.flags.f_contiguous or out.ndim != 2:
        raise ValueError("out must be a 2-d Fortran-ordered array of the assembly precision")
    if M == 0 or N == 0:
        return out
    # the block must lie inside out (the reference's store raises BoundsError otherwise)
    row_offset, col_offset = int(row_offset), int(col_offset)
    if row_offset < 0 or col_offset < 0 or row_offset + M > out.shape[0] or col_offset + N > out.shape[1]:
        raise IndexError(f"a {M} x {N} block at offsets ({row_offset}, {col_offset}) does not fit out of shape "
                         f"{out.shape}")
    lib = nat.load()
    with ctx.lock:
        xh, yh = space_handle(X, ctx), space_handle(Y, ctx)
        for coef, term in _terms(op):
            qs = resolve_quadstrat(quadstrat, term, X, Y)
            sc = (C.c_double * 2)(complex(coef).real, complex(coef).imag)
            nat.check(lib.bh_assemble(ctx.handle, C.byref(op_struct(term)), C.byref(qs_struct(qs)), prec, xh, yh, sc,
                                      1, out.ctypes.data_as(C.c_void_p), out.shape[0], row_offset, col_offset))
    return out


def assemble_rhs(field, X, quadstrat=defaultquadstrat, ctx=None):
    """``assemble((n x E) x n, X)`` or ``assemble(n x E, X)`` for a plane wave E
    (excitation.jl:17-115,145-180)."""
    ctx = ctx or default_context()
    if not isinstance(field, (TangTraceMW, CrossTraceMW)) or not isinstance(field.field, PlaneWaveMW):
        raise nat.UnsupportedOnGPU(nat.BH_ERR_UNSUPPORTED,
                                   "GPU right-hand sides support (n x E) x n and n x E of a plane wave")
    trace = nat.TRACE_CROSS if isinstance(field, CrossTraceMW) else nat.TRACE_TANGENTIAL
    qs = resolve_quadstrat(quadstrat, field, X, None)
    if not isinstance(qs, SingleNumQStrat):
        raise ValueError("right-hand sides take a SingleNumQStrat")
    E = field.field
    lib = nat.load()
    b = np.zeros(X.numfunctions(), dtype=np.complex128)
    d = nat.as_f64(E.direction)
    pr = nat.as_f64(E.polarisation.real)
    pi = nat.as_f64(E.polarisation.imag)
    g = np.array([E.gamma.real, E.gamma.imag])
    a = np.array([E.amplitude.real, E.amplitude.imag])
    with ctx.lock:
        nat.check(lib.bh_assemble_excitation(ctx.handle, space_handle(X, ctx), trace, nat.dptr(d), nat.dptr(pr),
                                             nat.dptr(pi), nat.dptr(g), nat.dptr(a), int(qs.quad_rule), 1,
                                             b.ctypes.data_as(nat._D)))
    return b


def classify(op, X, Y, quadstrat=defaultquadstrat, ctx=None):
    """Pair classes [nf_test, nf_trial] (0 far, 1 near, 2 CV, 3 CE, 4 CF)."""
    ctx = ctx or default_context()
    qs = resolve_quadstrat(quadstrat, op, X, Y)
    lib = nat.load()
    mx, my = X.geo if isinstance(X, RTBasis) else X, Y.geo if isinstance(Y, RTBasis) else Y
    out = np.zeros((mx.numcells, my.numcells), dtype=np.int8)
    nat.check(lib.bh_classify(ctx.handle, C.byref(op_struct(op)), C.byref(qs_struct(qs)), mesh_handle(mx, ctx),
                              mesh_handle(my, ctx), out.ctypes.data_as(C.POINTER(C.c_int8))))
    return out


def local_blocks(op, mesh_t, mesh_s, pairs, quadstrat=defaultquadstrat, ctx=None):
    """3x3 local blocks z[i, j] for (test cell, trial cell) pairs, as the GPU computes them."""
    ctx = ctx or default_context()
    qs = resolve_quadstrat(quadstrat, op, mesh_t, mesh_s)
    pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    tc, sc = nat.as_i32(pairs[:, 0]), nat.as_i32(pairs[:, 1])
    out = np.zeros((len(pairs), 18))
    lib = nat.load()
    nat.check(lib.bh_local_blocks(ctx.handle, C.byref(op_struct(op)), C.byref(qs_struct(qs)),
                                  mesh_handle(mesh_t, ctx), mesh_handle(mesh_s, ctx), len(pairs), nat.i32ptr(tc),
                                  nat.i32ptr(sc), nat.dptr(out)))
    z = out[:, 0::2] + 1j * out[:, 1::2]
    return z.reshape(-1, 3, 3).transpose(0, 2, 1)   # column-major 3x3 -> [i, j]


def element_to_dof(X, ctx=None):
    """The element-to-dof map the library built from X (dof[nf, 3, kmax], coeff)."""
    ctx = ctx or default_context()
    lib = nat.load()
    h = space_handle(X, ctx)
    k = C.c_int32()
    nat.check(lib.bh_space_e2d(h, C.byref(k), None, None, 0))
    n = X.geo.numcells * 3 * k.value
    dof = np.zeros(n, dtype=np.int32)
    cf = np.zeros(n)
    nat.check(lib.bh_space_e2d(h, C.byref(k), nat.i32ptr(dof), nat.dptr(cf), n))
    return dof.reshape(X.geo.numcells, 3, k.value), cf.reshape(X.geo.numcells, 3, k.value)


class DeviceMatrix:
    """A square system matrix resident in HBM (FP64) with its LU factors.

    Replaces the host Matrix{ComplexF64} + LAPACK LU of the reference
    (src/solvers/lusolver.jl:26-34, BEAST.lu in src/solvers/lu.jl:29-47).
    """

    def __init__(self, handle, n, ctx):
        self.h = handle
        self.n = n
        self.ctx = ctx
        self.stats = None
        _live["matrices"].add(self)

    @classmethod
    def assemble(cls, op, X, quadstrat=defaultquadstrat, ctx=None):
        """Resident matrix of ``op`` on X x X; a linear combination is the first term
        scaled, then the others accumulated in place (operator.jl:294-306)."""
        ctx = ctx or default_context()
        lib = nat.load()
        xh = space_handle(X, ctx)
        h = C.c_void_p()
        terms = list(_terms(op))
        # a term with coefficient 1 first (assembled in place, the others accumulated)
        ones = [i for i, (c, _) in enumerate(terms) if complex(c) == 1.0]
        if ones:
            terms.insert(0, terms.pop(ones[0]))
        c0, t0 = terms[0]
        qs = resolve_quadstrat(quadstrat, t0, X, X)
        with ctx.lock:
            nat.check(lib.bh_matrix_assemble(ctx.handle, C.byref(op_struct(t0)), C.byref(qs_struct(qs)), xh, C.byref(h)))
        M = cls(h, X.numfunctions(), ctx)
        if complex(c0) != 1.0:
            # the first term assembled unscaled: rescale by accumulating (c0 - 1) times it
            d = complex(c0) - 1.0
            sc = (C.c_double * 2)(d.real, d.imag)
            with M.ctx.lock:
                nat.check(lib.bh_matrix_accumulate(M.h, C.byref(op_struct(t0)), C.byref(qs_struct(qs)), xh, sc))
        for coef, term in terms[1:]:
            qs = resolve_quadstrat(quadstrat, term, X, X)
            sc = (C.c_double * 2)(complex(coef).real, complex(coef).imag)
            with M.ctx.lock:
                nat.check(lib.bh_matrix_accumulate(M.h, C.byref(op_struct(term)), C.byref(qs_struct(qs)), xh, sc))
        return M

    @classmethod
    def assemble_batch(cls, requests, quadstrat=defaultquadstrat, ctx=None):
        """Resident matrices of several (op, X) requests with shared launches
        (bh_matrix_assemble_batch): the integral operator terms of every request are
        assembled together, then each request's further terms are accumulated in place as
        in ``assemble``.  Each matrix equals the one ``assemble`` gives its request."""
        ctx = ctx or default_context()
        lib = nat.load()
        firsts, rests = [], []
        for op, X in requests:
            terms = list(_terms(op))
            # the first integral term with coefficient 1 goes into the batch, the rest is accumulated
            pick = next((i for i, (c, t) in enumerate(terms)
                         if complex(c) == 1.0 and not isinstance(t, (NCross, Identity))), None)
            if pick is None:
                pick = next(i for i, (c, t) in enumerate(terms) if not isinstance(t, (NCross, Identity)))
            c0, t0 = terms.pop(pick)
            if complex(c0) != 1.0:
                terms.insert(0, (complex(c0) - 1.0, t0))
            firsts.append((t0, X))
            rests.append(terms)
        n = len(firsts)
        ops = (nat.bh_operator * n)(*[op_struct(t) for t, _ in firsts])
        qss = (nat.bh_quadstrat * n)(*[qs_struct(resolve_quadstrat(quadstrat, t, X, X)) for t, X in firsts])
        sps = (C.c_void_p * n)(*[space_handle(X, ctx).value for _, X in firsts])
        outs = (C.c_void_p * n)()
        with ctx.lock:
            nat.check(lib.bh_matrix_assemble_batch(ctx.handle, n, ops, qss, sps, outs))
        mats = [cls(C.c_void_p(outs[i]), firsts[i][1].numfunctions(), ctx) for i in range(n)]
        for M, (t0, X), terms in zip(mats, firsts, rests):
            xh = space_handle(X, ctx)
            for coef, term in terms:
                qs = resolve_quadstrat(quadstrat, term, X, X)
                sc = (C.c_double * 2)(complex(coef).real, complex(coef).imag)
                with M.ctx.lock:
                    nat.check(lib.bh_matrix_accumulate(M.h, C.byref(op_struct(term)), C.byref(qs_struct(qs)), xh, sc))
        return mats

    @classmethod
    def from_host(cls, Z, ctx=None):
        ctx = ctx or default_context()
        Z = np.asfortranarray(Z, dtype=np.complex128)
        n = Z.shape[0]
        if Z.shape != (n, n):
            raise ValueError("square matrix expected")
        h = C.c_void_p()
        with ctx.lock:
            nat.check(nat.load().bh_matrix_from_host(ctx.handle, n, Z.ctypes.data_as(nat._D), n, C.byref(h)))
        return cls(h, n, ctx)

    def download(self):
        Z = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        with self.ctx.lock:
            nat.check(nat.load().bh_matrix_download(self.h, Z.ctypes.data_as(nat._D), self.n))
        return Z

    def factor(self, precision="fp64"):
        with self.ctx.lock:
            nat.check(nat.load().bh_matrix_factor(self.h, {"fp64": nat.FP64, "fp32": nat.FP32,
                                                           "fp32split": nat.FP32_SPLIT}[precision]))
        return self

    @staticmethod
    def factor_batch(mats, precision="fp32split"):
        """Factor several matrices of one context in shared launches
        (bh_matrix_factor_batch); each gets the factors ``factor`` gives it alone."""
        mats = list(mats)
        if not mats:
            return mats
        ctx = mats[0].ctx
        if any(M.ctx is not ctx for M in mats):
            raise ValueError("factor_batch: the matrices belong to different contexts")
        hs = (C.c_void_p * len(mats))(*[M.h.value for M in mats])
        prec = {"fp64": nat.FP64, "fp32": nat.FP32, "fp32split": nat.FP32_SPLIT}[precision]
        with ctx.lock:
            nat.check(nat.load().bh_matrix_factor_batch(ctx.handle, len(mats), hs, prec))
        return mats

    def pivots(self):
        ip = np.zeros(self.n, dtype=np.int64)
        with self.ctx.lock:
            nat.check(nat.load().bh_matrix_pivots(self.h, nat.i64ptr(ip)))
        return ip

    def solve(self, B, tol=0.0, maxit=0):
        """X = Z \\ B (B of shape [n] or [n, nrhs]); FP32 factors are refined to ``tol`` by GMRES-IR."""
        B = np.asarray(B, dtype=np.complex128)
        vec = B.ndim == 1
        Bf = np.asfortranarray(B.reshape(self.n, -1))
        Xf = np.zeros_like(Bf, order="F")
        st = nat.bh_solve_stats()
        with self.ctx.lock:
            nat.check(nat.load().bh_matrix_solve(self.h, Bf.shape[1], Bf.ctypes.data_as(nat._D),
                                                 Xf.ctypes.data_as(nat._D), float(tol), int(maxit), C.byref(st)))
        self.stats = {"iterations": st.iterations, "inner": st.inner, "residual": st.residual,
                      "seconds_factor": st.seconds_factor, "seconds_solve": st.seconds_solve}
        return Xf[:, 0] if vec else Xf

    def __del__(self):
        try:
            if self.h:
                with self.ctx.lock:
                    nat.load().bh_matrix_destroy(self.h)
                self.h = None
        except Exception:
            pass


def lu(A, precision="fp64", ctx=None):
    """``BEAST.lu``-style factorisation object of a host or resident matrix."""
    M = A if isinstance(A, DeviceMatrix) else DeviceMatrix.from_host(A, ctx)
    return M.factor(precision)
