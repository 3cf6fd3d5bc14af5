"""The reference's RT-tested MFIE on the GPU: n x K (DoubleLayerRotatedMW3D), Identity and the
n x H right-hand side against the oracle, and the MFIE n x K - Id/2 end to end against the Mie
series (needs an MI355X)."""
import os
import sys

import numpy as np
import pytest

import brute
from conftest import relfro

pytestmark = pytest.mark.gpu


def _nxk_oracle(orc, op, qs, X, Y):
    return orc.assemble(3, op.gamma, op.alpha, 0.0, qs.kind, qs.params, X, Y)


def _meshes(bj, golden):
    return {"sphere45": bj.readmesh(os.path.join(golden, "sphere45.in")),
            "sphere2": bj.readmesh(os.path.join(golden, "sphere2.in")),
            "cube": bj.meshcube(3)}


@pytest.mark.parametrize("name", ["sphere45", "sphere2", "cube"])
@pytest.mark.parametrize("qs_args", [None, ("Sauter", 2, 3, 5, 5, 4, 3), ("DoubleNum", 2, 3)])
def test_rotated_double_layer_matches_oracle(bj, orc, golden, name, qs_args):
    """n x K on RT x RT, host mode, FP64, for the default DoubleNumQStrat(6,7)
    (nxdbllayer.jl:23) and the other strategies that apply: within 1e-12 of the oracle."""
    m = _meshes(bj, golden)[name]
    X = bj.raviartthomas(m)
    for gamma in (1.0j, 6.0j, 0.5 + 3.0j):
        op = bj.DoubleLayerRotatedMW3D(1.0, gamma)
        if qs_args is None:
            qs = bj.defaultquadstrat(op, X, X)
            assert qs.params == (6, 7)
        elif qs_args[0] == "Sauter":
            qs = bj.DoubleNumSauterQstrat(*qs_args[1:])
        else:
            qs = bj.DoubleNumQStrat(*qs_args[1:])
        Zg = bj.assemble(op, X, X, quadstrat=qs)
        assert relfro(Zg, _nxk_oracle(orc, op, qs, X, X)) < 1e-12, gamma


def test_rotated_double_layer_blocks_every_class(bj, orc, golden):
    """n x K blocks of every pair class under DoubleNumSauterQstrat (Sauter-Schwab on touching
    pairs, double Gauss elsewhere) against the oracle, block by block."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    op = bj.DoubleLayerRotatedMW3D(1.0, 6.0j)
    qs = bj.DoubleNumSauterQstrat(2, 3, 5, 5, 4, 3)
    cls = bj.classify(op, X, X, quadstrat=qs)
    pairs = []
    for c in (0, 2, 3, 4):
        pairs += [tuple(p) for p in np.argwhere(cls == c)[:20]]
    z = bj.local_blocks(op, m, m, pairs, quadstrat=qs)
    ref = []
    for p, q in pairs:
        zo, c = orc.block(3, op.gamma, op.alpha, 0.0, qs.kind, qs.params, m.vertices[m.cells[p]],
                          m.vertices[m.cells[q]])
        assert c == cls[p, q]
        ref.append(zo)
    # common-face blocks vanish on flat cells: the tolerance is relative to the largest block
    scale = max(np.abs(zo).max() for zo in ref)
    for (p, q), zg, zo in zip(pairs, z, ref):
        assert np.abs(zg - zo).max() <= 1e-12 * np.abs(zo).max() + 1e-13 * scale, (cls[p, q], p, q)


def test_rotated_double_layer_refuses_wilton(bj, golden):
    """The reference defines no singularity extraction for n x K: DoubleNumWiltonSauterQStrat is
    reported as unsupported (the caller's CPU path raises as the reference would)."""
    m = bj.readmesh(os.path.join(golden, "sphere45.in"))
    X = bj.raviartthomas(m)
    op = bj.cross(bj.n, bj.Maxwell3D.doublelayer(wavenumber=1.0))
    with pytest.raises(bj.UnsupportedOnGPU):
        bj.assemble(op, X, X, quadstrat=bj.DoubleNumWiltonSauterQStrat(2, 3, 6, 7, 5, 5, 4, 3))


def test_rotated_double_layer_fp32(bj, orc, golden):
    """FP32 n x K within 1e-5 of the FP64 oracle (relative Frobenius).  Under the default
    DoubleNumQStrat(6,7) every pair, touching ones included, is a plain double quadrature of
    a 1/R^2 kernel, which FP32 resolves to about 2e-6 per common-edge block (numpy emulation
    of the kernel's arithmetic); a flat cell paired with itself contributes exactly zero."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    op = bj.DoubleLayerRotatedMW3D(1.0, 2.0j)
    Z32 = bj.assemble(op, X, X, precision="fp32")
    err = relfro(Z32.astype(np.complex128), _nxk_oracle(orc, op, bj.defaultquadstrat(op, X, X), X, X))
    print(f"n x K FP32 vs oracle: {err:.2e}", file=sys.stderr)
    assert err < 1e-5


@pytest.mark.parametrize("mesh", ["sphere2", "cube"])
def test_identity_matches_oracle(bj, orc, golden, mesh):
    """Identity (identityop.jl:6-11) on RT x RT, host mode: within 1e-13 of the oracle."""
    m = _meshes(bj, golden)[mesh]
    X = bj.raviartthomas(m)
    assert relfro(bj.assemble(bj.Identity(), X, X), orc.identity(X, X)) < 1e-13


def test_cross_trace_rhs(bj, orc, golden):
    """n x H tested with RT (CrossTraceMW, excitation.jl:168-173), H = -1/(i omega) curl E of a
    plane wave: within 1e-13 of the oracle; the tangential trace still equals it too."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    k = 2.0
    d, p = np.array([0.3, -0.4, 0.866]), np.array([1.0, 0.2, 0.0])
    d /= np.linalg.norm(d)
    p -= p.dot(d) * d
    E = bj.Maxwell3D.planewave(direction=d, polarization=p, wavenumber=k)
    H = (-1.0 / (1j * k)) * bj.curl(E)
    b = bj.assemble(bj.cross(bj.n, H), X)
    bo = orc.planewave(X, d, np.cross(d, p), 1j * k, amplitude=H.amplitude, trace="cross")
    assert relfro(b, bo) < 1e-13
    bt = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    assert relfro(bt, orc.planewave(X, d, p, 1j * k)) < 1e-13


def test_nxmfie_resident_solve_and_mie(bj, orc):
    """The MFIE (n x K - Id/2) j = n x H of examples/nxmfie.jl:9-22 on a 2396-cell sphere, k = 1:
    the resident matrix (n x K assembled, Id/2 accumulated in place) equals the oracle's; the
    mixed solve (FP32 LU with split-bf16 updates + GMRES-IR) agrees with the oracle's LAPACK
    solve to 1e-8, and its far field with the PEC Mie series to 5e-3 (the oracle's level)."""
    m = bj.meshsphere(1.0, nv=1200)
    X = bj.raviartthomas(m)
    k = 1.0
    op = bj.cross(bj.n, bj.Maxwell3D.doublelayer(wavenumber=k)) - 0.5 * bj.Identity()
    M = bj.DeviceMatrix.assemble(op, X)
    Zo = orc.assemble(3, 1j * k, 1.0, 0.0, 2, (6, 7), X, X) - 0.5 * orc.identity(X, X)
    assert relfro(M.download(), Zo) < 1e-12
    d, u = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    E = bj.Maxwell3D.planewave(direction=d, polarization=u, wavenumber=k)
    H = (-1.0 / (1j * k)) * bj.curl(E)
    b = bj.assemble(bj.cross(bj.n, H), X)
    M.factor("fp32split")
    x = M.solve(b, tol=1e-12, maxit=10)
    print(f"nxMFIE: {M.stats}", file=sys.stderr)
    assert M.stats["residual"] <= 1e-12 and M.stats["inner"] <= 10
    assert relfro(x, orc.lu_solve(Zo, b)) < 1e-8
    th = np.linspace(0.05, np.pi - 0.05, 9)
    S1, S2 = brute.mie_pec_sphere(1.0, th)
    Ed = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    Hd = np.stack([0 * th, np.sin(th), np.cos(th)], 1)
    ff = bj.MWFarField3D(wavenumber=k)
    fe = np.linalg.norm(bj.potential(ff, Ed, x, X), axis=1) / (4 * np.pi)
    fh = np.linalg.norm(bj.potential(ff, Hd, x, X), axis=1) / (4 * np.pi)
    err = max(np.abs(fe - np.abs(S2)).max(), np.abs(fh - np.abs(S1)).max()) / np.abs(S2).max()
    print(f"nxMFIE 2396 cells, k = 1: max Mie error {err:.2e}", file=sys.stderr)
    assert err < 5e-3
