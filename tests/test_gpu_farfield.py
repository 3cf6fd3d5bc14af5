"""Far fields on the GPU (bh_farfield) against the oracle's restatement of
potential(op, points, coeffs, X) (postproc.jl:174-211,273-288, maxwell/farfield.jl:56-100)."""
import os

import numpy as np
import pytest

import brute

pytestmark = pytest.mark.gpu


def _relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _ops(bj, gamma, amp):
    return [bj.MWFarField3D(gamma, amp), bj.MWDoubleLayerFarField3D(gamma, amp),
            bj.MWDoubleLayerRotatedFarField3D(gamma, amp), bj.MWFarField3DDropConstant(gamma, amp)]


@pytest.mark.parametrize("gamma", [1.0j, 0.4 + 3.0j])
@pytest.mark.parametrize("rule", [10, 3])
def test_farfield_matches_oracle(bj, orc, golden, gamma, rule):
    """Every far-field kind on sphere2.in with a random current at 300 directions
    (not unit length, one past a multiple of the 256-lane block plus a ragged tail):
    FP64 end to end, so 1e-12 relative to the largest component."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(X.numfunctions()) + 1j * rng.standard_normal(X.numfunctions())
    Y = rng.standard_normal((300, 3)) * 1.3
    for op in _ops(bj, gamma, 0.8 - 0.3j):
        print(f"kind {op.kind} gamma {gamma} rule {rule}", flush=True)
        got = bj.potential(op, Y, u, X, quadstrat=bj.SingleNumQStrat(rule))
        want = orc.farfield(op.kind, gamma, op.amplitude, Y, u, X, degree=rule)
        assert got.shape == (300, 3)
        assert _relmax(got, want) < 1e-12, (op, _relmax(got, want))


@pytest.mark.parametrize("k", [0.2, 1.0, 4.0, 25.0])
def test_farfield_cell_factored_paths(bj, orc, golden, k):
    """Lossless kernels factor the phase per cell and take Taylor polynomials of degree
    8/12/18 for the in-cell phase (bound |k| max|y| max|x - c|: 0.03, 0.16, 0.64 on
    sphere2.in with unit directions), or the direct sincos path beyond that (k = 25):
    every path against the oracle's complex exponential at 1e-12."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    rng = np.random.default_rng(int(k * 10))
    u = rng.standard_normal(X.numfunctions()) + 1j * rng.standard_normal(X.numfunctions())
    Y = rng.standard_normal((77, 3))
    Y /= np.linalg.norm(Y, axis=1)[:, None]
    for op in (bj.MWFarField3D(wavenumber=k), bj.MWDoubleLayerRotatedFarField3D(wavenumber=k)):
        print(f"k {k} {type(op).__name__}", flush=True)
        got = bj.potential(op, Y, u, X)
        want = orc.farfield(op.kind, op.gamma, op.amplitude, Y, u, X)
        assert _relmax(got, want) < 1e-12, (k, _relmax(got, want))


def test_farfield_shapes_and_edges(bj, orc, golden):
    """One direction, an empty point set, a grid of points keeps its shape, a zero current
    gives zeros, and two calls give the same bits (fixed-order reduction)."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(X.numfunctions()) + 0j
    op = bj.MWFarField3D(wavenumber=2.0)
    y1 = np.array([[0.0, 0.0, 1.0]])
    assert _relmax(bj.potential(op, y1, u, X), orc.farfield(0, 2.0j, 1.0, y1, u, X)) < 1e-12
    assert bj.potential(op, np.zeros((0, 3)), u, X).shape == (0, 3)
    th, ph = np.meshgrid(np.linspace(0, np.pi, 5), np.linspace(0, 2 * np.pi, 4), indexing="ij")
    G = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    g = bj.potential(op, G, u, X)
    assert g.shape == (5, 4, 3)
    assert _relmax(g.reshape(-1, 3), orc.farfield(0, 2.0j, 1.0, G.reshape(-1, 3), u, X)) < 1e-12
    assert np.all(bj.potential(op, G, np.zeros_like(u), X) == 0)
    a, b = bj.potential(op, G, u, X), bj.potential(op, G, u, X)
    assert np.array_equal(a, b)


def test_farfield_large_mesh_properties(bj, orc):
    """A 26664-triangle sphere (the EFIE solve workload's mesh) at 4096 directions: a subset
    of directions against the oracle, and at full size the properties that do not need it:
    linearity in the current, (single layer) = (double layer) x y, and the drop-constant
    field = single layer minus its gamma = 0 field."""
    m = bj.meshsphere(1.0, nv=13334)
    X = bj.raviartthomas(m)
    n = X.numfunctions()
    rng = np.random.default_rng(7)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Y = rng.standard_normal((4096, 3))
    Y /= np.linalg.norm(Y, axis=1)[:, None]
    g = 4.0j
    sl = bj.potential(bj.MWFarField3D(g, 1.0), Y, u, X)
    sub = Y[::512]
    assert _relmax(sl[::512], orc.farfield(0, g, 1.0, sub, u, X)) < 1e-12
    slv = bj.potential(bj.MWFarField3D(g, 1.0), Y, v, X)
    lin = bj.potential(bj.MWFarField3D(g, 1.0), Y, 2.0 * u - 0.5j * v, X)
    assert _relmax(lin, 2.0 * sl - 0.5j * slv) < 1e-12
    dl = bj.potential(bj.MWDoubleLayerFarField3D(g, 1.0), Y, u, X)
    assert _relmax(np.cross(dl, Y), sl) < 1e-12
    dc = bj.potential(bj.MWFarField3DDropConstant(g, 1.0), Y, u, X)
    st = bj.potential(bj.MWFarField3D(0.0, 1.0), Y, u, X)
    assert _relmax(dc, sl - st) < 1e-10


def test_farfield_golden_through_gpu(bj, golden):
    """test_farfield.jl:34-40 fully on the GPU: assembly, right-hand side, mixed solve and
    the far field (MWFarField3D(1.0im), default SingleNumQStrat(10)), |ff| i omega/(4 pi c)
    against the golden magnitudes at the discretisation-level agreement of our own solve."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=1.0)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    u = bj.DeviceMatrix.assemble(t, X).factor("fp32").solve(b, tol=1e-12)
    th = np.linspace(0, np.pi, 7)
    pts = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    ff = bj.potential(bj.MWFarField3D(1.0j), pts, u, X) * (1j / (4 * np.pi))
    a = np.linalg.norm(ff, axis=1)
    assert np.abs(a - np.array(brute.FARFIELD_GOLDEN)).max() < 1e-3


def test_farfield_mie_through_gpu(bj):
    """EFIE on the 26664-triangle unit sphere at k = 2 fully on the GPU (assembly,
    right-hand side, mixed solve, far field) against the PEC Mie series, r|E_s| =
    |S2|/k in the E plane and |S1|/k in the H plane (i omega / (4 pi c) scaling of
    test_farfield.jl:26 with omega = k, c = 1)."""
    k = 2.0
    m = bj.meshsphere(1.0, nv=13334)
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=k)
    b = bj.assemble(bj.cross(bj.cross(bj.n, E), bj.n), X)
    u = bj.DeviceMatrix.assemble(t, X).factor("fp32split").solve(b, tol=1e-11)
    th = np.linspace(0.05, np.pi - 0.05, 13)
    S1, S2 = brute.mie_pec_sphere(k, th)
    Ye = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    Yh = np.stack([0 * th, np.sin(th), np.cos(th)], 1)
    ff = bj.MWFarField3D(wavenumber=k)
    fe = np.linalg.norm(bj.potential(ff, Ye, u, X), axis=1) * k / (4 * np.pi)
    fh = np.linalg.norm(bj.potential(ff, Yh, u, X), axis=1) * k / (4 * np.pi)
    ref = np.abs(S2).max() / k
    err = max(np.abs(fe - np.abs(S2) / k).max(), np.abs(fh - np.abs(S1) / k).max()) / ref
    print(f"26664 cells, k = 2: max Mie error {err:.2e}")
    assert err < 1e-3
