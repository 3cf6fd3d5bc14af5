"""GPU assembly parity against the CPU oracle (needs an MI355X)."""
import os

import numpy as np
import pytest

from conftest import relfro

pytestmark = pytest.mark.gpu

DEFAULT = (2, 3, 6, 7, 5, 5, 4, 3)


def _mesh(bj, golden, name):
    return bj.readmesh(os.path.join(golden, name))


def _oracle_Z(orc, op, qs, X, Y):
    kind = 0 if op.__class__.__name__ == "MWSingleLayer3D" else 1
    beta = op.beta if kind == 0 else 0.0
    return orc.assemble(kind, op.gamma, op.alpha, beta, qs.kind, qs.params, X, Y)


def _ops(bj, k):
    return [bj.Maxwell3D.singlelayer(wavenumber=k), bj.Maxwell3D.doublelayer(wavenumber=k)]


@pytest.mark.parametrize("name,k", [("sphere45.in", 3.0), ("sphere2.in", 1.0), ("sphere5.in", 2 * np.pi / 200)])
def test_classification_bitexact(bj, orc, golden, name, k):
    m = _mesh(bj, golden, name)
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    gpu = bj.classify(t, X, X)
    ref = np.zeros_like(gpu)
    for p in range(m.numcells):
        tv = m.vertices[m.cells[p]]
        for q in range(m.numcells):
            ref[p, q] = orc.classify(tv, m.vertices[m.cells[q]], k * k)
    assert np.array_equal(gpu, ref)


@pytest.mark.parametrize("name,k", [("sphere45.in", 3.0), ("sphere2.in", 1.0)])
@pytest.mark.parametrize("which", [0, 1])
def test_matrix_parity_default_quadstrat(bj, orc, golden, name, k, which):
    m = _mesh(bj, golden, name)
    X = bj.raviartthomas(m)
    op = _ops(bj, k)[which]
    qs = bj.defaultquadstrat(op, X, X)
    Z = bj.assemble(op, X, X)
    Zo = _oracle_Z(orc, op, qs, X, X)
    err = relfro(Z, Zo)
    print(f"{name} k={k} op={type(op).__name__} rel Frobenius {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("qs_args", [("DoubleNumQStrat", (2, 3)), ("DoubleNumQStrat", (4, 5)),
                                     ("DoubleNumSauterQstrat", (2, 3, 5, 5, 4, 3)),
                                     ("DoubleNumWiltonSauterQStrat", (3, 4, 7, 8, 5, 6, 5, 4))])
@pytest.mark.parametrize("which", [0, 1])
def test_matrix_parity_quadstrats(bj, orc, golden, qs_args, which):
    m = _mesh(bj, golden, "sphere45.in")
    X = bj.raviartthomas(m)
    op = _ops(bj, 2.5)[which]
    qs = getattr(bj, qs_args[0])(*qs_args[1])
    Z = bj.assemble(op, X, X, quadstrat=qs)
    Zo = _oracle_Z(orc, op, qs, X, X)
    assert relfro(Z, Zo) < 1e-12


def test_quadstrat_as_function(bj, golden):
    # test_quadstrat_as_fn.jl: a function (op, X, Y) -> strategy gives the same matrix as the object
    m = _mesh(bj, golden, "sphere45.in")
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    qs = bj.DoubleNumSauterQstrat(2, 3, 5, 5, 4, 3)
    Z1 = bj.assemble(t, X, X, quadstrat=qs)
    Z2 = bj.assemble(t, X, X, quadstrat=lambda op, a, b: qs)
    assert np.array_equal(Z1, Z2)


def test_cube_and_distinct_meshes(bj, orc):
    # coplanar neighbours (flat faces) and test/trial on different meshes
    cube = bj.meshcube(4)
    X = bj.raviartthomas(cube)
    t = bj.Maxwell3D.singlelayer(wavenumber=2.0)
    qs = bj.defaultquadstrat(t, X, X)
    assert relfro(bj.assemble(t, X, X), _oracle_Z(orc, t, qs, X, X)) < 1e-12
    s1 = bj.meshsphere(nv=120)
    s2 = bj.Mesh(s1.vertices * 0.7 + np.array([2.2, 0.3, 0.0]), s1.cells)
    Y = bj.raviartthomas(s2)
    X1 = bj.raviartthomas(s1)
    for op in _ops(bj, 3.0):
        assert relfro(bj.assemble(op, X1, Y), _oracle_Z(orc, op, qs, X1, Y)) < 1e-12


def test_lossy_gamma(bj, orc, golden):
    # complex gamma (lossy medium) and the static-like real gamma of test_sauterschwabints.jl
    m = _mesh(bj, golden, "sphere45.in")
    X = bj.raviartthomas(m)
    for op in (bj.MWSingleLayer3D(0.5 + 2.0j), bj.MWSingleLayer3D(1.0, 250.0, 1.0), bj.MWDoubleLayer3D(1.0, 0.3 + 1.0j)):
        qs = bj.defaultquadstrat(op, X, X)
        assert relfro(bj.assemble(op, X, X), _oracle_Z(orc, op, qs, X, X)) < 1e-12


def test_local_blocks_every_class(bj, orc, golden):
    m = _mesh(bj, golden, "sphere2.in")
    X = bj.raviartthomas(m)
    k = 6.0
    for op in _ops(bj, k):
        cls = bj.classify(op, X, X)
        pairs = []
        for c in range(5):
            idx = np.argwhere(cls == c)[:20]
            pairs += [tuple(p) for p in idx]
        z = bj.local_blocks(op, m, m, pairs)
        kind = 0 if op.__class__.__name__ == "MWSingleLayer3D" else 1
        beta = op.beta if kind == 0 else 0.0
        ref = []
        for p, q in pairs:
            zo, c = orc.block(kind, op.gamma, op.alpha, beta, 0, DEFAULT, m.vertices[m.cells[p]], m.vertices[m.cells[q]])
            assert c == cls[p, q]
            ref.append(zo)
        # common-face MFIE-K blocks vanish on flat cells (both sides are rounding noise), so the
        # tolerance is relative to the largest block of the operator
        scale = max(np.abs(zo).max() for zo in ref)
        for (p, q), zg, zo in zip(pairs, z, ref):
            assert np.abs(zg - zo).max() <= 1e-12 * np.abs(zo).max() + 1e-13 * scale, (cls[p, q], p, q)


def test_fp32_far_field_error_profile(bj, orc, golden):
    m = _mesh(bj, golden, "sphere2.in")
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    Z32 = bj.assemble(t, X, X, precision="fp32")
    Z64 = bj.assemble(t, X, X)
    err = relfro(Z32.astype(np.complex128), Z64)
    print(f"FP32 far field rel Frobenius {err:.3e}")
    assert Z32.dtype == np.complex64 and err < 2e-6


def test_linear_combination_and_store(bj, golden):
    m = _mesh(bj, golden, "sphere45.in")
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    k = bj.Maxwell3D.doublelayer(wavenumber=1.0)
    Zt, Zk = bj.assemble(t, X, X), bj.assemble(k, X, X)
    Z = bj.assemble(2.0 * t + (0.5 - 1.0j) * k, X, X)
    assert relfro(Z, 2.0 * Zt + (0.5 - 1.0j) * Zk) < 1e-14
    # assemble! store semantics: add into a block of a bigger matrix (operator.jl:319-356 offsets)
    N = X.numfunctions()
    big = np.ones((N + 5, N + 7), dtype=np.complex128, order="F")
    bj.assemble(t, X, X, out=big, row_offset=3, col_offset=4)
    assert np.array_equal(big[3:3 + N, 4:4 + N], Zt + 1.0)
    assert np.all(big[:3] == 1.0) and np.all(big[:, :4] == 1.0)


def test_empty_space(bj, golden):
    m = _mesh(bj, golden, "sphere45.in")
    X = bj.raviartthomas(m)
    E = X.subset([])
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    Z = bj.assemble(t, E, X)
    assert Z.shape == (0, X.numfunctions())


def test_subset_rows(bj, golden):
    # dofsplitting assembles row chunks of subset(X, lo:hi); the rows agree with the full matrix
    m = _mesh(bj, golden, "sphere45.in")
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=3.0)
    Z = bj.assemble(t, X, X)
    rows = np.arange(37, 201)
    Zs = bj.assemble(t, X.subset(rows), X)
    assert relfro(Zs, Z[rows]) < 1e-15


def test_hits_assertion_gpu(bj):
    V = np.array([[0, 0, 0], [1e-14, 0, 0], [0, 1e-14, 0], [1.0, 1.0, 0], [1.0, 0.0, 0.0]], float)
    m = bj.Mesh(V, np.array([[0, 1, 2], [0, 4, 3]]))
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    X = _two_dof_space(bj, m)
    with pytest.raises(AssertionError):
        bj.assemble(t, X, X)


def _two_dof_space(bj, m):
    return bj.RTBasis(m, [0, 1], [0], [0], [1.0])


def test_planewave_rhs(bj, orc, golden):
    m = _mesh(bj, golden, "sphere2.in")
    X = bj.raviartthomas(m)
    k = 1.0
    E = bj.Maxwell3D.planewave(direction=bj.z_hat, polarization=bj.x_hat, wavenumber=k)
    e = bj.cross(bj.cross(bj.n, E), bj.n)
    b = bj.assemble(e, X)
    bo = orc.planewave(X, bj.z_hat, bj.x_hat, 1j * k)
    assert relfro(b, bo) < 1e-13


@pytest.mark.parametrize("mesh", ["sphere2.in", "cube"])
def test_ncross_matches_oracle(bj, orc, golden, mesh):
    """NCross (identityop.jl:20-22) on RT x RT, host mode, FP64: within 1e-12 of the oracle."""
    m = bj.readmesh(os.path.join(golden, mesh)) if mesh.endswith(".in") else bj.meshcube(3)
    X = bj.raviartthomas(m)
    Zg = bj.assemble(bj.NCross(), X, X)
    Zo = orc.ncross(X, X)
    assert relfro(Zg, Zo) < 1e-12


def test_mfie_resident_combination(bj, orc, golden):
    """Resident K + N/2 (a linear combination: K assembled, N/2 accumulated in place) equals the
    host assembly of K plus half the oracle's NCross; then an MFIE solve converges."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    k = 1.0
    K = bj.Maxwell3D.doublelayer(wavenumber=k)
    M = bj.DeviceMatrix.assemble(K + 0.5 * bj.NCross(), X)
    Zref = bj.assemble(K, X, X) + 0.5 * orc.ncross(X, X)
    assert relfro(M.download(), Zref) < 1e-13
    d, p = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    H = bj.Maxwell3D.planewave(direction=d, polarization=np.cross(d, p), wavenumber=k)
    b = bj.assemble(bj.cross(bj.cross(bj.n, H), bj.n), X)
    M.factor("fp32split")
    u = M.solve(b, tol=1e-12)
    assert np.linalg.norm(Zref @ u - b) / np.linalg.norm(b) < 1e-11


@pytest.mark.parametrize("k", [1.0, 6.0])
def test_grid_classification_large_meshes(bj, orc, k):
    """Meshes with more than 2^22 cell pairs classify through the trial-cell grid;
    the matrix must equal the oracle's (which classifies every pair), also for two
    distinct, nearly touching meshes."""
    s1 = bj.meshsphere(1.0, nv=1100)
    assert s1.numcells ** 2 >= (1 << 22)
    X = bj.raviartthomas(s1)
    t = bj.Maxwell3D.singlelayer(wavenumber=k)
    qs = bj.defaultquadstrat(t, X, X)
    rows = np.arange(0, X.numfunctions(), 11)
    Xs = X.subset(rows)
    assert relfro(bj.assemble(t, Xs, X), _oracle_Z(orc, t, qs, Xs, X)) < 1e-12
    s2 = bj.Mesh(s1.vertices * 0.985, s1.cells)   # a second sphere just inside the first
    Y = bj.raviartthomas(s2)
    Ys = Y.subset(np.arange(0, Y.numfunctions(), 13))
    for op in _ops(bj, k):
        assert relfro(bj.assemble(op, Xs, Ys), _oracle_Z(orc, op, qs, Xs, Ys)) < 1e-12
