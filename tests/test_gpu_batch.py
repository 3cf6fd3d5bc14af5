"""Batched assembly (bh_matrix_assemble_batch): several requests' classification, near blocks and
far tiles in shared launches give every request the bits it gets alone (bh_matrix_assemble, a
batch of one), whatever the other requests of the batch (needs an MI355X)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(bj):
    sphere = bj.meshsphere(1.0, nv=300)
    big = bj.meshsphere(1.0, nv=2100)           # more than 2^22 cell pairs: grid classification
    ell = bj.meshellipsoid(2.0, 1.0, 1.0, nv=400)
    cube = bj.meshcube(4)
    K = bj.Maxwell3D.doublelayer
    return [
        (bj.Maxwell3D.singlelayer(wavenumber=2.0), bj.raviartthomas(sphere)),
        (bj.Maxwell3D.singlelayer(wavenumber=5.0), bj.raviartthomas(cube)),
        (bj.cross(bj.n, K(wavenumber=1.5)) - 0.5 * bj.Identity(), bj.raviartthomas(ell)),
        (K(wavenumber=3.0), bj.raviartthomas(sphere)),
        (bj.Maxwell3D.singlelayer(wavenumber=1.0), bj.raviartthomas(big)),
        (bj.Maxwell3D.singlelayer(gamma=0.4 + 2.0j), bj.raviartthomas(ell)),
        (bj.cross(bj.n, K(wavenumber=6.0)) - 0.5 * bj.Identity(), bj.raviartthomas(cube)),
    ]


def test_batch_equals_single_requests(bj):
    reqs = _requests(bj)
    print(f"batch of {len(reqs)}: N = {[X.numfunctions() for _, X in reqs]}", file=sys.stderr, flush=True)
    batch = bj.DeviceMatrix.assemble_batch(reqs)
    for (op, X), Mb in zip(reqs, batch):
        Ms = bj.DeviceMatrix.assemble(op, X)
        assert np.array_equal(Mb.download(), Ms.download()), op
        # the FP32 copy the factorisation starts from is the same too
        rng = np.random.default_rng(X.numfunctions())
        b = rng.standard_normal(X.numfunctions()) + 1j * rng.standard_normal(X.numfunctions())
        xb = Mb.factor("fp32split").solve(b, tol=1e-10, maxit=10)
        xs = Ms.factor("fp32split").solve(b, tol=1e-10, maxit=10)
        assert np.array_equal(xb, xs), op


def test_batch_of_one_and_errors(bj):
    X = bj.raviartthomas(bj.meshsphere(1.0, nv=200))
    t = bj.Maxwell3D.singlelayer(wavenumber=2.0)
    (M,) = bj.DeviceMatrix.assemble_batch([(t, X)])
    assert np.array_equal(M.download(), bj.DeviceMatrix.assemble(t, X).download())
    nxk = bj.cross(bj.n, bj.Maxwell3D.doublelayer(wavenumber=1.0))
    with pytest.raises(bj.UnsupportedOnGPU):
        bj.DeviceMatrix.assemble_batch([(t, X), (nxk, X)],
                                       quadstrat=bj.DoubleNumWiltonSauterQStrat(2, 3, 6, 7, 5, 5, 4, 3))


def _rand_complex(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 0.5 * np.sqrt(n) * np.eye(n)


def test_factor_batch_equals_factor_alone(bj):
    """bh_matrix_factor_batch: matrices of one to eleven 512-deep block columns, factored in
    shared launches, get the pivots and the solutions (bitwise, with the same GMRES-IR
    iterations) that bh_matrix_factor gives each alone."""
    k = bj.Maxwell3D.singlelayer(wavenumber=3.0)
    spaces = [bj.raviartthomas(bj.meshsphere(1.0, nv=v)) for v in (100, 302, 900, 1900)]
    spaces.append(bj.raviartthomas(bj.meshcube(7)))    # 882 dofs
    hosts = [_rand_complex(n, n) for n in (64, 513, 5500)]

    def build():
        mats = [bj.DeviceMatrix.assemble(k, X) for X in spaces]
        return mats + [bj.DeviceMatrix.from_host(Z) for Z in hosts]

    batch, alone = build(), build()
    print(f"factor batch of {len(batch)}: N = {[M.n for M in batch]}", file=sys.stderr, flush=True)
    bj.DeviceMatrix.factor_batch(batch, "fp32split")
    for Mb, Ma in zip(batch, alone):
        Ma.factor("fp32split")
        assert np.array_equal(Mb.pivots(), Ma.pivots()), Mb.n
        rng = np.random.default_rng(Mb.n + 1)
        B = rng.standard_normal((Mb.n, 2)) + 1j * rng.standard_normal((Mb.n, 2))
        xb = Mb.solve(B, tol=1e-11, maxit=10)
        xa = Ma.solve(B, tol=1e-11, maxit=10)
        assert np.array_equal(xb, xa), Mb.n
        assert all(Mb.stats[f] == Ma.stats[f] for f in ("iterations", "inner", "residual"))
        assert Mb.stats["residual"] <= 1e-11


def test_factor_batch_singular_and_other_precisions(bj):
    """A singular matrix in a batch is reported after the others are factored; FP64 and
    FP32 batches factor one by one, as bh_matrix_factor does."""
    good = [_rand_complex(n, n) for n in (40, 700)]
    bad = _rand_complex(300, 9)
    bad[:, 7] = 0.0
    mats = [bj.DeviceMatrix.from_host(Z) for Z in (good[0], bad, good[1])]
    with pytest.raises(bj.SingularException):
        bj.DeviceMatrix.factor_batch(mats, "fp32split")
    for M, Z in ((mats[0], good[0]), (mats[2], good[1])):
        b = np.ones(M.n, complex)
        x = M.solve(b, tol=1e-12, maxit=10)
        assert np.linalg.norm(Z @ x - b) / np.linalg.norm(b) < 1e-11
    for prec in ("fp64", "fp32"):
        ms = [bj.DeviceMatrix.from_host(Z) for Z in good]
        bj.DeviceMatrix.factor_batch(ms, prec)
        for M, Z in zip(ms, good):
            x = M.solve(np.ones(M.n, complex), tol=1e-12, maxit=10)
            assert np.linalg.norm(Z @ x - 1.0) / np.sqrt(M.n) < 1e-11
