"""Dense LU and solves on the GPU against LAPACK (needs an MI355X)."""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import relfro

pytestmark = pytest.mark.gpu


def _rand(n, seed, cond_boost=True):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    if cond_boost:
        A += 0.5 * np.sqrt(n) * np.eye(n)
    return np.asfortranarray(A)


@pytest.mark.parametrize("n", [1, 5, 31, 33, 64, 130, 257, 700])
def test_lu_fp64_matches_lapack(bj, n):
    A = _rand(n, n)
    M = bj.DeviceMatrix.from_host(A)
    assert np.array_equal(M.download(), A)
    M.factor("fp64")
    lu, piv = sla.lu_factor(A)
    # LAPACK's pivot rule (largest |Re|+|Im|, first on ties) -> identical interchanges
    assert np.array_equal(M.pivots(), piv + 1)
    b = _rand(n, n + 1)[:, :3]
    x = M.solve(b)
    xr = sla.lu_solve((lu, piv), b)
    assert relfro(x, xr) < 1e-12
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 100 * np.finfo(float).eps * max(1, np.sqrt(n))


@pytest.mark.parametrize("precision", ["fp32", "fp32split"])
@pytest.mark.parametrize("n", [40, 300, 1025, 2100])
def test_lu_fp32_gmres_ir(bj, n, precision):
    A = _rand(n, 7 * n)
    b = _rand(n, 7 * n + 1)[:, 0]
    M = bj.DeviceMatrix.from_host(A).factor(precision)
    x = M.solve(b, tol=1e-12, maxit=10)
    xr = np.linalg.solve(A, b)
    assert M.stats["residual"] <= 1e-12
    assert relfro(x, xr) < 1e-10
    assert M.stats["iterations"] >= 1


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32split"])
def test_singular_pivot(bj, precision):
    A = _rand(40, 3)
    A[:, 7] = 0.0
    M = bj.DeviceMatrix.from_host(A)
    with pytest.raises(bj.SingularException):
        M.factor(precision)


def _perm_from_ipiv(ipiv1):
    """(P A)[i] = A[perm[i]] from 1-based sequential interchanges (LAPACK getrf convention)."""
    perm = np.arange(len(ipiv1))
    for i, p in enumerate(ipiv1 - 1):
        perm[i], perm[p] = perm[p], perm[i]
    return perm


@pytest.mark.parametrize("precision", ["fp32", "fp32split"])
@pytest.mark.parametrize("n", [300, 2100, 9000])
def test_fp32_tournament_finds_permutation(bj, n, precision):
    """A row-permuted, strongly diagonally dominant matrix: every pivot search (per
    workgroup, at every merge level, and in the final ordering) must pick the
    dominant entry, so P A is the unpermuted matrix and IR converges at once."""
    rng = np.random.default_rng(n)
    B = 0.01 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    B += np.diag(1.0 + rng.uniform(0, 1, n)) * np.exp(2j * np.pi * rng.uniform(0, 1, n))
    sigma = rng.permutation(n)
    A = np.asfortranarray(B[sigma])
    M = bj.DeviceMatrix.from_host(A).factor(precision)
    ipiv = M.pivots()
    assert ipiv.shape == (n,) and np.all(ipiv >= np.arange(1, n + 1)) and np.all(ipiv <= n)
    perm = _perm_from_ipiv(ipiv)
    assert np.array_equal(sigma[perm], np.arange(n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    M.solve(b, tol=1e-12, maxit=10)
    assert M.stats["residual"] <= 1e-12
    assert M.stats["iterations"] <= 3


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32split"])
def test_efie_end_to_end(bj, orc, golden, precision):
    # test_itsolver.jl:9-24 on sphere2.in: EFIE k=1, plane wave z / x-polarised
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    k = 1.0
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=k)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    M = bj.DeviceMatrix.assemble(t, X)
    Zo = orc.assemble(0, t.gamma, t.alpha, t.beta, 0, (2, 3, 6, 7, 5, 5, 4, 3), X, X)
    Zg = M.download()
    assert relfro(Zg, Zo) < 1e-12
    M.factor(precision)
    u = M.solve(b, tol=1e-13 if precision != "fp64" else 0.0)
    uo = orc.lu_solve(Zo, orc.planewave(X, bj.z_hat, bj.x_hat, 1j * k))
    # the residual of the system that was solved, as test_itsolver.jl:21-24 (LAPACK gives the same order)
    res = np.linalg.norm(Zg @ u - b) / np.linalg.norm(b)
    res_lapack = np.linalg.norm(Zg @ orc.lu_solve(Zg, b) - b) / np.linalg.norm(b)
    print(f"{precision}: residual {res:.2e} (LAPACK on the same matrix {res_lapack:.2e})")
    assert res < 100 * np.finfo(float).eps * (10 if precision != "fp64" else 1)
    assert relfro(u, uo) < 1e-10


def test_farfield_golden_gpu(bj, golden):
    """GPU assembly + RHS + mixed solve reproduce the reference's far-field golden values
    (test_farfield.jl:34-40) at the oracle's discretisation-level agreement."""
    import brute
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=1.0)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    M = bj.DeviceMatrix.assemble(t, X).factor("fp32")
    u = M.solve(b, tol=1e-12)
    a = brute.farfield_magnitudes(bj, u, X, np.linspace(0, np.pi, 7))
    assert np.abs(a - np.array(brute.FARFIELD_GOLDEN)).max() < 1e-3


def test_fp32split_deep_blocks_match_fp64(bj):
    """N >= 32768 takes 1024-deep FP32 blocks, each factored as two 512 halves whose
    interchanges, block row and split-bf16 update join them: the mixed solution must
    still meet the north-star bound against the plain FP64 path."""
    mesh = bj.meshsphere(1.0, nv=11000)
    X = bj.raviartthomas(mesh)
    assert X.numfunctions() >= 32768
    k = 2.0
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    E = bj.Maxwell3D.planewave(direction=bj.x_hat, polarization=bj.z_hat, wavenumber=k)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    M = bj.DeviceMatrix.assemble(t, X).factor("fp32split")
    x32 = M.solve(b, tol=1e-11, maxit=10)
    assert M.stats["residual"] <= 1e-11
    del M
    M = bj.DeviceMatrix.assemble(t, X).factor("fp64")
    x64 = M.solve(b)
    assert relfro(x32, x64) < 1e-8


def test_refactor_sequences(bj, golden):
    """The FP32 copy written by the resident assembly is consumed by the first FP32
    factorisation; factoring again (FP32 twice, then FP64, then FP32) converts from the
    FP64 matrix each time and gives the same solution; the context's cached FP32
    workspace serves matrices of different sizes in turn."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=1.0)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    M = bj.DeviceMatrix.assemble(t, X)
    sols = []
    for prec in ("fp32split", "fp32split", "fp64", "fp32"):
        M.factor(prec)
        sols.append(M.solve(b, tol=1e-13 if prec != "fp64" else 0.0))
    for u in sols[1:]:
        assert relfro(u, sols[0]) < 1e-10
    # a larger and then a smaller system on the same context
    m2 = bj.meshsphere(1.0, nv=900)
    X2 = bj.raviartthomas(m2)
    M2 = bj.DeviceMatrix.assemble(t, X2).factor("fp32split")
    b2 = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X2)
    u2 = M2.solve(b2, tol=1e-12)
    assert M2.stats["residual"] <= 1e-12
    M.factor("fp32split")
    assert relfro(M.solve(b, tol=1e-13), sols[0]) < 1e-10
    del u2


def test_transfers_with_padded_leading_dimension(bj):
    """bh_matrix_from_host / bh_matrix_download with ld > n read and write the n x n block
    only: the padding rows of the caller's columns are neither read nor overwritten."""
    import ctypes as C
    nat = bj._native
    n, ld = 37, 41
    A = _rand(n, 5)
    H = np.full((ld, n), np.nan + 1j * np.nan, dtype=np.complex128, order="F")
    H[:n] = A
    ctx = bj.default_context()
    h = C.c_void_p()
    nat.check(nat.load().bh_matrix_from_host(ctx.handle, n, H.ctypes.data_as(nat._D), ld, C.byref(h)))
    M = bj.DeviceMatrix(h, n, ctx)
    assert np.array_equal(M.download(), A)
    out = np.full((ld, n), 7.0 - 3.0j, dtype=np.complex128, order="F")
    nat.check(nat.load().bh_matrix_download(M.h, out.ctypes.data_as(nat._D), ld))
    assert np.array_equal(out[:n], A)
    assert np.all(out[n:] == 7.0 - 3.0j)


def test_golden_efie_solution_gpu(bj, golden):
    """The reference's EFIE solution on sphere2.in (tests/golden/efie_solution.txt) against the
    GPU assembly + mixed solve: sorted coefficient magnitudes (invariant under the dof order and
    signs) agree at quadrature level."""
    import brute
    u0 = brute.read_julia_complex_vector(os.path.join(golden, "efie_solution.txt"))
    X = bj.raviartthomas(bj.readmesh(os.path.join(golden, "sphere2.in")))
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=1.0)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    u = bj.DeviceMatrix.assemble(t, X).factor("fp32split").solve(b, tol=1e-12)
    ref = np.sort(np.abs(u0))
    assert np.abs(np.sort(np.abs(u)) - ref).max() / ref.max() < 2e-3


@pytest.mark.parametrize("precision,tol", [("fp32split", 1e-12), ("fp32", 1e-12), ("fp64", 1e-13)])
def test_right_hand_sides_in_lockstep_equal_one_at_a_time(bj, precision, tol):
    """GMRES-IR refines all right-hand sides of a solve together (one Krylov space per
    column, shared launches): each column equals, bit for bit, the solve of that column
    alone, with columns converging after different numbers of steps; the iteration counts
    add up over the columns."""
    n = 2100
    A = _rand(n, 77)
    rng = np.random.default_rng(78)
    B = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    B[:, 2] *= 1e-6                                   # a column far smaller than the others
    B[:, 4] = A @ (rng.standard_normal(n) + 0j)       # and one with a different spectrum mix
    M = bj.DeviceMatrix.from_host(A).factor(precision)
    X = M.solve(B, tol=tol, maxit=10)
    inner_all = M.stats["inner"]
    inner_sum = 0
    for c in range(B.shape[1]):
        x = M.solve(B[:, c], tol=tol, maxit=10)
        inner_sum += M.stats["inner"]
        assert np.array_equal(X[:, c], x), c
        assert np.linalg.norm(A @ x - B[:, c]) / np.linalg.norm(B[:, c]) < 1e-11
    assert inner_all == inner_sum
