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
