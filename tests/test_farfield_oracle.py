"""Far-field post-processing: pins of the oracle's far field (postproc.jl:174-211,273-288,
maxwell/farfield.jl:56-75,98-100) and the host-side operator constructors
(farfield.jl:1-103, test_handlers.jl:72-74).  CPU only."""
import os

import numpy as np
import pytest

import brute


def _tri_space(bj, V, cells, refs, coeffs):
    """One triangle and one dof whose shapes are (cell 0, ref r, coeff c)."""
    m = bj.Mesh(np.asarray(V, float), np.array([[0, 1, 2]]))
    return bj.RTBasis(m, [0, len(refs)], [0] * len(refs), refs, coeffs)


def _divdiff_exp(s):
    """exp[s0, ..., sn] for distinct nodes: sum_i e^{s_i} / prod_{j != i} (s_i - s_j)."""
    s = list(s)
    out = 0j
    for i, si in enumerate(s):
        den = 1.0 + 0j
        for j, sj in enumerate(s):
            if j != i:
                den *= si - sj
        out += np.exp(si) / den
    return out


def _plane_wave_moments(V, k):
    """I0 = int_T exp(k.x) dA and I1 = int_T x exp(k.x) dA in closed form.

    Hermite-Genocchi: int over the unit simplex of f''(lambda . s) = f[s0, s1, s2], so
    I0 = 2A exp[s0, s1, s2] with s_i = k . v_i; and d/ds_m f[s0, s1, s2] = f[s0, s1, s2, s_m]
    gives I1 = grad_k I0 = 2A sum_m exp[s0, s1, s2, s_m] v_m, the repeated-node difference
    taken as (f[a, b, c] - f[a, a, b]) / (c - a), f[a, a, b] = (f[a, b] - e^a) / (b - a)."""
    V = np.asarray(V, float)
    A = 0.5 * np.linalg.norm(np.cross(V[0] - V[2], V[1] - V[2]))
    s = V @ k
    I0 = 2 * A * _divdiff_exp(s)
    I1 = np.zeros(3, complex)
    for m in range(3):
        a = s[m]
        b, c = [s[j] for j in range(3) if j != m]
        fab = (np.exp(b) - np.exp(a)) / (b - a)
        faab = (fab - np.exp(a)) / (b - a)
        fabc = _divdiff_exp([a, b, c])
        I1 += 2 * A * (fabc - faab) / (c - a) * V[m]
    return I0, I1, A


def _integrand(kind, y, S, nrm, amp):
    if kind in (0, 3):
        return amp * np.cross(np.cross(y, S), y)
    if kind == 1:
        return amp * np.cross(y, S)
    return amp * np.cross(y, np.cross(S, nrm))


TRI = [[0.10, -0.05, 0.20], [0.38, 0.02, 0.11], [0.15, 0.29, 0.26]]


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("gamma", [2.0j, 0.4 + 1.7j])
def test_oracle_single_shape_closed_form(bj, orc, kind, gamma):
    """One RT shape f_r = (x - v_r)/(2A) on one triangle: S = int exp(gamma y.x) f_r dA
    = (I1 - v_r I0)/(2A) at k = gamma y (expm1 kernel: minus the k = 0 moments), against
    the degree-10 rule of the oracle.  A dropped term, a wrong sign, index or transposed
    cross product fails at O(1)."""
    V = np.array(TRI)
    y = np.array([0.3, -0.5, 0.81])
    amp = 0.7 - 0.2j
    nrm = np.cross(V[0] - V[2], V[1] - V[2])
    nrm /= np.linalg.norm(nrm)
    for r in range(3):
        X = _tri_space(bj, V, [0], [r], [1.0])
        I0, I1, A = _plane_wave_moments(V, gamma * y)
        S = (I1 - V[r] * I0) / (2 * A)
        if kind == 3:
            S = S - (V.mean(0) - V[r]) / 2   # int f_r dA = (centroid - v_r) A / (2A)
        want = _integrand(kind, y, S, nrm, amp)
        got = orc.farfield(kind, gamma, amp, [y], [1.0], X)[0]
        assert np.abs(got - want).max() < 1e-9 * np.abs(want).max(), (r, got, want)


@pytest.mark.parametrize("gamma", [3.0j, 0.5 + 2.5j])
def test_oracle_constant_current_closed_form(bj, orc, gamma):
    """Shapes r=0 and r=1 with coefficients +1, -1 make the constant current (v1 - v0)/(2A):
    S = I0 (v1 - v0)/(2A), the scalar Hermite-Genocchi closed form alone."""
    V = np.array(TRI)
    X = _tri_space(bj, V, [0, 0], [0, 1], [1.0, -1.0])
    for y in ([0.0, 0.0, 1.0], [0.6, 0.8, 0.0], [-0.2, 0.4, 0.3]):
        y = np.array(y)
        I0, _, A = _plane_wave_moments(V, gamma * y)
        want = np.cross(np.cross(y, I0 * (V[1] - V[0]) / (2 * A)), y)
        got = orc.farfield(0, gamma, 1.0, [y], [1.0], X)[0]
        assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


def test_oracle_kinds_agree(bj, orc):
    """On a closed surface with a random current: (single layer) = (double layer) x y at the
    same amplitude; the drop-constant field differs from the single layer by the
    gamma = 0 field; and every kind agrees with the numpy restatement."""
    m = bj.meshsphere(1.0, h=0.45)
    X = bj.raviartthomas(m)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(X.numfunctions()) + 1j * rng.standard_normal(X.numfunctions())
    Y = rng.standard_normal((9, 3))
    g = 0.3 + 2.0j
    sl = orc.farfield(0, g, 1.5, Y, u, X, degree=6)
    dl = orc.farfield(1, g, 1.5, Y, u, X, degree=6)
    dc = orc.farfield(3, g, 1.5, Y, u, X, degree=6)
    st = orc.farfield(0, 0.0, 1.5, Y, u, X, degree=6)
    assert np.abs(sl - np.cross(dl, Y)).max() < 1e-12 * np.abs(sl).max()
    assert np.abs(dc - (sl - st)).max() < 1e-12 * np.abs(sl).max()
    # every kind against the numpy restatement (its own summation arrangement)
    for kind in (0, 1, 2, 3):
        want = brute.farfield(bj, u, X, Y, kind, g, 1.5, degree=6)
        got = orc.farfield(kind, g, 1.5, Y, u, X, degree=6)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), kind


def test_oracle_farfield_golden(bj, orc, golden):
    """test_farfield.jl:34-40 through the oracle's far field at the reference default rule
    (degree 10): EFIE on sphere824.in (= sphere2.in) at kappa = 1, plane wave z / x, our
    own solve (the golden solution predates the current dof order), |ff| i omega/(4 pi c);
    and the independent numpy restatement of the same sum agrees to rounding."""
    m = bj.readmesh(os.path.join(golden, "sphere2.in"))
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    Z = orc.assemble(0, t.gamma, t.alpha, t.beta, 0, (2, 3, 6, 7, 5, 5, 4, 3), X, X)
    b = orc.planewave(X, [0, 0, 1.0], [1.0, 0, 0], 1.0j)
    u = orc.lu_solve(Z, b)
    th = np.linspace(0, np.pi, 7)
    Y = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    ff = orc.farfield(0, 1.0j, 1.0, Y, u, X) * (1j / (4 * np.pi))
    a = np.linalg.norm(ff, axis=1)
    assert np.abs(a - np.array(brute.FARFIELD_GOLDEN)).max() < 1e-3
    nb = brute.farfield_magnitudes(bj, u, X, th, degree=10)
    assert np.abs(a - nb).max() < 1e-12


def test_mie_series_pins():
    """The Mie comparator itself: the Rayleigh limit of a small conducting sphere
    (backscatter |S|/k -> (3/2) k^2 a^3 ... with a_1 -> -2i x^3/3, b_1 -> i x^3/3), and
    the reference's golden magnitudes (test_farfield.jl:34-40, its own 824-triangle
    solution) within the 2% of that mesh."""
    x = 1e-3
    S1, S2 = brute.mie_pec_sphere(x, [np.pi])
    # theta = pi: pi_1 = 1, tau_1 = -1: S = (3/2)(a_1 - b_1) -> (3/2)(-i x^3) for both
    assert abs(abs(S1[0]) - 1.5 * x ** 3) < 1e-6 * x ** 3
    assert abs(abs(S2[0]) - 1.5 * x ** 3) < 1e-6 * x ** 3
    th = np.linspace(0, np.pi, 7)
    _, S2 = brute.mie_pec_sphere(1.0, th)
    assert np.abs(np.abs(S2) - np.array(brute.FARFIELD_GOLDEN)).max() < 0.02 * np.abs(S2).max()


@pytest.mark.parametrize("nv", [400, 1200])
def test_oracle_farfield_mie(bj, orc, nv):
    """EFIE on a unit sphere at k = 1 (oracle assembly, LAPACK solve, oracle far field)
    against the PEC Mie series in the E and H planes (r|E_s| = |S|/k, k = 1 here); the
    agreement improves with the mesh (the far field converges to the exact scattered
    field)."""
    m = bj.meshsphere(1.0, nv=nv)
    X = bj.raviartthomas(m)
    t = bj.Maxwell3D.singlelayer(wavenumber=1.0)
    Z = orc.assemble(0, t.gamma, t.alpha, t.beta, 0, (2, 3, 6, 7, 5, 5, 4, 3), X, X)
    b = orc.planewave(X, [0, 0, 1.0], [1.0, 0, 0], 1.0j)
    u = orc.lu_solve(Z, b)
    th = np.linspace(0.05, np.pi - 0.05, 9)
    S1, S2 = brute.mie_pec_sphere(1.0, th)
    E = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    H = np.stack([0 * th, np.sin(th), np.cos(th)], 1)
    fe = np.linalg.norm(orc.farfield(0, 1.0j, 1.0, E, u, X), axis=1) / (4 * np.pi)
    fh = np.linalg.norm(orc.farfield(0, 1.0j, 1.0, H, u, X), axis=1) / (4 * np.pi)
    err = max(np.abs(fe - np.abs(S2)).max(), np.abs(fh - np.abs(S1)).max()) / np.abs(S2).max()
    print(f"nv={nv} cells={m.numcells}: max Mie error {err:.2e}")
    assert err < (0.015 if nv == 400 else 0.005)


def test_farfield_constructors(bj):
    """farfield.jl:22-53,82-103 and test_handlers.jl:72-74."""
    for cls in (bj.MWFarField3D, bj.MWDoubleLayerFarField3D, bj.MWDoubleLayerRotatedFarField3D):
        with pytest.raises(AssertionError):
            cls()
        op = cls(wavenumber=2.0)
        assert op.gamma == 2.0j and op.amplitude == 1.0
        op = cls(gamma=0.5 + 1j, amplitude=3.0)
        assert op.gamma == 0.5 + 1j and op.amplitude == 3.0
        with pytest.raises(ValueError):
            cls(gamma=1j, wavenumber=1.0)
    t = bj.Maxwell3D.singlelayer(wavenumber=2.0, alpha=-1j * 2.0 * 3.0, beta=-1 / (1j * 2.0 * 5.0))
    f = bj.MWFarField3D(t)
    assert f.gamma == t.gamma and abs(f.amplitude - np.sqrt(3.0 / 5.0)) < 1e-15
    assert bj.MWFarField3D(bj.Maxwell3D.singlelayer(wavenumber=1.0)).amplitude == 1.0
    k = bj.Maxwell3D.doublelayer(wavenumber=2.0)
    assert bj.MWDoubleLayerFarField3D(k).gamma == 2.0j
    r = bj.cross(bj.n, bj.MWDoubleLayerFarField3D(1j, 2.0))
    assert isinstance(r, bj.MWDoubleLayerRotatedFarField3D) and r.amplitude == 2.0
    d = bj.cross(bj.n, r)
    assert isinstance(d, bj.MWDoubleLayerFarField3D) and d.amplitude == -2.0
    assert bj.Maxwell3D.farfield(wavenumber=3.0).gamma == 3.0j
    assert bj.defaultquadstrat(bj.MWFarField3D(1j)).quad_rule == 10
    assert bj.MWFarField3DDropConstant(1j, 0.5).coeff == 0.5


def test_potential_rejects_bad_input(bj):
    m = bj.meshsphere(1.0, h=0.6)
    X = bj.raviartthomas(m)
    with pytest.raises(AssertionError):
        bj.potential(bj.MWFarField3D(1j), np.zeros((2, 3)), np.zeros(3), X)
    with pytest.raises(ValueError):
        bj.potential(bj.MWFarField3D(1j), np.zeros((2, 2)), np.zeros(X.numfunctions()), X)
    with pytest.raises(bj.UnsupportedOnGPU):
        bj.potential(bj.Maxwell3D.singlelayer(wavenumber=1.0), np.zeros((2, 3)), np.zeros(X.numfunctions()), X)
