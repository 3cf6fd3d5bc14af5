"""bh_matrix_solve_noalloc against bh_matrix_solve through the C ABI (needs an MI355X): two
calls in a row on the shared scratch, the second on a smaller system, each equal to the
allocating solve's result."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A += 0.5 * np.sqrt(n) * np.eye(n)
    return np.asfortranarray(A)


def _solve_noalloc(bj, M, B, tol, maxit):
    nat = bj._native
    B = np.asfortranarray(B, dtype=np.complex128)
    X = np.zeros_like(B, order="F")
    st = nat.bh_solve_stats()
    nat.check(nat.load().bh_matrix_solve_noalloc(M.h, B.shape[1], B.ctypes.data_as(nat._D),
                                                 X.ctypes.data_as(nat._D), float(tol), int(maxit), C.byref(st)))
    return X, st


@pytest.mark.parametrize("precision,tol", [("fp32split", 1e-12), ("fp64", 0.0)])
def test_noalloc_solve_matches_solve(bj, precision, tol):
    cases = [(3000, 4, 11), (1000, 2, 12)]   # the second call smaller than the first
    for n, nrhs, seed in cases:
        print(f"noalloc solve: {precision} n={n} nrhs={nrhs}", flush=True)
        A = _system(n, seed)
        rng = np.random.default_rng(seed + 100)
        B = np.asfortranarray(rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs)))
        M = bj.DeviceMatrix.from_host(A).factor(precision)
        X1, st = _solve_noalloc(bj, M, B, tol, 10)
        X2 = M.solve(B, tol=tol, maxit=10)
        assert np.array_equal(X1, X2)
        assert (st.iterations, st.inner, st.residual) == (M.stats["iterations"], M.stats["inner"],
                                                          M.stats["residual"])
        assert np.linalg.norm(A @ X1 - B) / np.linalg.norm(B) < 1e-11
