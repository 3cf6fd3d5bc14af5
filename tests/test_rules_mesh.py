"""Quadrature tables, meshes and the RT dof map (CPU)."""
import math
import os

import numpy as np
import pytest


def _moment(a, b):
    # int_T u^a v^b over the reference triangle {u, v >= 0, u + v <= 1}
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("deg", range(1, 11))
def test_dunavant_exactness(bj, orc, deg):
    uv, w = bj.trgauss(deg)
    assert abs(w.sum() - 0.5) < 1e-14
    assert np.all(uv >= 0) and np.all(uv.sum(axis=0) <= 1 + 1e-15)
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            q = np.sum(w * uv[0] ** a * uv[1] ** b)
            assert abs(q - _moment(a, b)) < 2e-14, (deg, a, b)
    # one degree more is not integrated exactly (the degree label is right)
    errs = [abs(np.sum(w * uv[0] ** a * uv[1] ** (deg + 1 - a)) - _moment(a, deg + 1 - a)) for a in range(deg + 2)]
    assert max(errs) > 1e-10
    # the oracle's copy of the table is the same, bit for bit
    uo, wo = orc.trgauss(deg)
    assert np.array_equal(uo, uv) and np.array_equal(wo, w)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 12, 16])
def test_legendre(bj, orc, n):
    x, w = bj.legendre01(n)
    xo, wo = orc.legendre01(n)
    assert np.allclose(xo, x, rtol=0, atol=2e-15) and np.allclose(wo, w, rtol=0, atol=2e-15)
    for k in range(2 * n):
        assert abs(np.sum(wo * xo ** k) - 1.0 / (k + 1)) < 1e-14


def test_readmesh_fixtures(bj, golden):
    m = bj.readmesh(os.path.join(golden, "sphere45.in"))
    assert (m.numvertices, m.numcells) == (109, 212)
    m2 = bj.readmesh(os.path.join(golden, "sphere2.in"))
    assert (m2.numvertices, m2.numcells) == (415, 824)
    assert bj.raviartthomas(m).numfunctions() == 318
    assert bj.raviartthomas(m2).numfunctions() == 1236


def test_writemesh_roundtrip(bj, tmp_path):
    m = bj.meshsphere(nv=60)
    p = tmp_path / "s.in"
    bj.writemesh(str(p), m)
    m2 = bj.readmesh(str(p))
    assert np.array_equal(m.cells, m2.cells) and np.array_equal(m.vertices, m2.vertices)


@pytest.mark.parametrize("nv", [12, 236, 1000])
def test_fibonacci_sphere(bj, nv):
    m = bj.meshsphere(nv=nv)
    assert m.numcells == 2 * nv - 4
    c = m.vertices[m.cells]
    nrm = np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0])
    assert np.all(np.einsum("ij,ij->i", nrm, c.mean(axis=1)) > 0)   # outward


def _rt_invariants(bj, m):
    X = bj.raviartthomas(m)
    F = m.numcells
    assert X.numfunctions() == 3 * F // 2
    assert np.all(np.diff(X.shape_ptr) == 2)
    assert np.all(X.shape_coeff[0::2] == 1.0) and np.all(X.shape_coeff[1::2] == -1.0)
    for b in range(X.numfunctions()):
        (c1, e1, _), (c2, e2, _) = X.basisfunction(b)
        v1, v2 = m.cells[c1], m.cells[c2]
        shared = set(v1) & set(v2)
        assert len(shared) == 2                               # rtspace.jl:55
        assert v1[e1] not in shared and v2[e2] not in shared  # rtspace.jl:56
    dof, cf = bj.assemblydata(X)
    assert dof.shape == (F, 3, 1) and np.all(dof >= 0)         # closed: no holes
    return X


def test_rt_sphere(bj, golden):
    _rt_invariants(bj, bj.readmesh(os.path.join(golden, "sphere45.in")))
    _rt_invariants(bj, bj.meshsphere(nv=300))
    _rt_invariants(bj, bj.meshcube(3))


def test_rt_rectangle_counts(bj):
    # test_raviartthomas.jl:28-58: meshrectangle(1, 1, 0.5): 9 vertices, 16 edges, 8 RT dofs
    m = bj.meshrectangle(1.0, 1.0, 0.5)
    assert m.numvertices == 9
    assert abs(m.areas()[0] - 1 / 8) < 1e-15
    edges = bj.mesh_edges(m)[0]
    assert len(edges) == 16
    assert bj.raviartthomas(m).numfunctions() == 8


def test_cube_counts(bj):
    for n in (1, 2, 4):
        m = bj.meshcube(n)
        assert m.numcells == 12 * n * n
        assert bj.raviartthomas(m).numfunctions() == 18 * n * n
