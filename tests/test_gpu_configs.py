"""The benchmark configurations end to end on the GPU (needs an MI355X).

C2: plain FP64, EFIE and MFIE-K, jittered sphere of 11,604 cells (17,406 dofs), k = 2 pi.
C3: mixed precision at its own wavenumber (k = 4) against the plain FP64 path.
C4: the seeded 64-request batch (N in [500, 8000]) against the oracle + LAPACK.
C5: a seeded Poisson stream (N up to 30,000) against the plain FP64 GPU path.
"""
import importlib
import sys

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import relfro

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
QS = (2, 3, 6, 7, 5, 5, 4, 3)   # DoubleNumWiltonSauterQStrat defaults (mwops.jl:55)


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _serve(bj):
    return importlib.import_module("beast_jl_amd.serve")


def _oracle_Z(orc, bj, op, X, Y):
    """The oracle's matrix of an operator or a linear combination (default quadstrats)."""
    if isinstance(op, bj.LinearCombinationOfOperators):
        return sum(c * _oracle_Z(orc, bj, t, X, Y) for c, t in op.terms)
    if isinstance(op, bj.LocalOperator):
        return orc.localop({2: 0, 4: 1}[op.kind], X, Y, bj.defaultquadstrat(op, X, Y).quad_rule)
    qs = bj.defaultquadstrat(op, X, Y)
    return orc.assemble(op.kind, op.gamma, op.alpha, getattr(op, "beta", 0.0), qs.kind, qs.params, X, Y)


@pytest.fixture(scope="module")
def c2(bj):
    m = bj.meshsphere(1.0, nv=5804, jitter=0.02, seed=2)
    X = bj.raviartthomas(m)
    assert (m.numcells, X.numfunctions()) == (11604, 17406)
    return m, X


@pytest.mark.parametrize("which", ["efie", "mfie_k"])
def test_c2_fp64_rows_match_oracle_and_solve(bj, orc, c2, which):
    """C2 in plain FP64: rows of the resident matrix equal the oracle's to 1e-12 (a row
    sample through the host entry point, which must equal the resident matrix's rows).
    The EFIE's FP64 LU solve leaves a relative residual below 100 eps (test_itsolver.jl:21-24).
    The double layer alone has no identity term and is ill conditioned tested with RT, so
    its relative residual is amplified by the condition number (2e-12 measured); its solve
    is held to the normwise backward error ||A x - b|| / (||A|| ||x|| + ||b||) < 100 eps
    (infinity norms), the bound a backward-stable LU meets whatever the conditioning."""
    m, X = c2
    k = 2 * np.pi
    op = bj.Maxwell3D.singlelayer(wavenumber=k) if which == "efie" else bj.Maxwell3D.doublelayer(wavenumber=k)
    N = X.numfunctions()
    rows = np.arange(0, N, 97)
    _log(f"C2 {which}: rows {len(rows)} x {N}")
    Zs = bj.assemble(op, X.subset(rows), X)
    assert relfro(Zs, _oracle_Z(orc, bj, op, X.subset(rows), X)) < 1e-12
    M = bj.DeviceMatrix.assemble(op, X)
    Z = M.download()
    assert relfro(Z[rows], Zs) < 1e-14
    M.factor("fp64")
    E = bj.Maxwell3D.planewave(direction=bj.x_hat, polarization=bj.z_hat, wavenumber=k)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    u = M.solve(b)
    r = Z @ u - b
    res = np.linalg.norm(r) / np.linalg.norm(b)
    bwd = np.abs(r).max() / (np.abs(Z).sum(axis=1).max() * np.abs(u).max() + np.abs(b).max())
    _log(f"C2 {which}: FP64 solve relative residual {res:.3e}, backward error {bwd:.3e}")
    assert bwd < 100 * EPS
    if which == "efie":
        assert res < 100 * EPS


def test_c3_mixed_vs_fp64_at_k4(bj):
    """C3 at its own wavenumber: the mixed path (FP32 LU with split-bf16 updates + GMRES-IR)
    agrees with the plain FP64 GPU path to 1e-8."""
    k = 4.0
    m = bj.meshsphere(1.0, nv=13334)
    X = bj.raviartthomas(m)
    assert X.numfunctions() == 39996
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    E = bj.Maxwell3D.planewave(direction=bj.x_hat, polarization=bj.z_hat, wavenumber=k)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    _log("C3: FP64 path")
    M = bj.DeviceMatrix.assemble(t, X).factor("fp64")
    x64 = M.solve(b)
    del M
    _log("C3: mixed path")
    M = bj.DeviceMatrix.assemble(t, X).factor("fp32split")
    x = M.solve(b, tol=1e-11, maxit=10)
    assert M.stats["residual"] <= 1e-11
    err = relfro(x, x64)
    _log(f"C3: mixed vs FP64 {err:.3e}")
    assert err < 1e-8


def _oracle_rhs(orc, X, kind, k, d, p):
    """EFIE: (n x E) x n; MFIE: n x H, H = -1/(i k) curl E = (d x p) exp(-i k d.x)."""
    if kind == "efie":
        return orc.planewave(X, d, p, 1j * k)
    return orc.planewave(X, d, np.cross(d, p), 1j * k, trace="cross")


def test_c4_batch_against_oracle(bj, orc):
    """C4: the seeded 64-request batch; every request converges to its tolerance, and four
    requests spanning the N range agree with the oracle matrix + LAPACK to 1e-8."""
    sv = _serve(bj)
    reqs = sv.batched_requests(seed=4, count=64)
    assert len(reqs) == 64
    srv = sv.Server(tol=1e-10)
    st = srv.run(reqs, realtime=False)
    _log(f"C4: {st}")
    assert st["requests"] == 64
    for r in reqs:
        assert r.info["residual"] <= 1e-10, r.rid
    order = sorted(reqs, key=lambda r: r.dofs)
    picks = [order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]]
    for r in picks:
        m, X = srv._space(r.shape, r.dofs)
        op = srv._operator(r.kind, r.wavenumber)
        _log(f"C4: oracle for request {r.rid} ({r.shape}, N={X.numfunctions()}, k={r.wavenumber:.3f})")
        Z = _oracle_Z(orc, bj, op, X, X)
        B = np.stack([_oracle_rhs(orc, X, r.kind, r.wavenumber, d, p) for d, p in r.rhs], axis=1)
        U = sla.lu_solve(sla.lu_factor(Z, check_finite=False), B, check_finite=False)
        err = relfro(r.result.reshape(U.shape), U)
        _log(f"C4: request {r.rid} vs oracle {err:.3e}")
        assert err < 1e-8, r.rid
    srv.close()


def test_c5_stream_against_fp64_and_alone(bj):
    """C5: Poisson arrivals (20/s, seed 5) with N up to 30,000, EFIE and MFIE; sampled
    results agree with the plain FP64 GPU solve to 1e-8, and the largest request gives the
    same bits alone as in the mix."""
    sv = _serve(bj)
    reqs = sv.poisson_requests(seed=5, rate=20.0, count=48, nmax=30000)
    srv = sv.Server(tol=1e-10)
    st = srv.run(reqs)
    _log(f"C5: {st}")
    assert st["requests"] == 48
    for r in reqs:
        assert r.info["residual"] <= 1e-10, r.rid
    big = max(reqs, key=lambda r: r.dofs)
    picks = {big.rid: big}
    for r in reqs:
        if r.kind != big.kind and len(picks) < 2:
            picks[r.rid] = r
    picks.setdefault(reqs[1].rid, reqs[1])
    ctx = bj.default_context()
    for r in picks.values():
        m, X = srv._space(r.shape, r.dofs)
        op = srv._operator(r.kind, r.wavenumber)
        _log(f"C5: FP64 path for request {r.rid} ({r.kind}, {r.shape}, N={X.numfunctions()})")
        M = bj.DeviceMatrix.assemble(op, X).factor("fp64")
        B = np.stack([srv._rhs(X, r.kind, r.wavenumber, d, p, ctx) for d, p in r.rhs], axis=1)
        U = M.solve(B)
        del M
        err = relfro(r.result.reshape(U.shape), U)
        _log(f"C5: request {r.rid} vs FP64 {err:.3e}")
        assert err < 1e-8, r.rid
    srv.close()
    alone = sv.Request(big.rid, 0.0, big.shape, big.dofs, big.kind, big.wavenumber, big.rhs)
    s1 = sv.Server(tol=1e-10)
    s1.run([alone], realtime=False)
    s1.close()
    assert np.array_equal(alone.result, big.result)
