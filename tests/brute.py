"""Brute-force comparators for the oracle's pins (test utilities).

``se_composite`` evaluates a pair block by singularity extraction with composite
rules: the outer triangle and the regular part use Dunavant rules on 4^L
subtriangles, the singular part the analytic potential integrals.  As L grows
it converges to the exact integral for any pair, touching pairs included, so
it checks the Sauter-Schwab blocks independently of their region formulas.
"""
import numpy as np


def subtriangles(v, L):
    tris = [np.asarray(v, float)]
    for _ in range(L):
        new = []
        for a, b, c in tris:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            new += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        tris = new
    return tris


def composite_points(trgauss, v, degree, L):
    uv, w = trgauss(degree)
    P, W = [], []
    for t in subtriangles(v, L):
        A = 0.5 * np.linalg.norm(np.cross(t[0] - t[2], t[1] - t[2]))
        for q in range(len(w)):
            u, vv = uv[0, q], uv[1, q]
            P.append(u * t[0] + vv * t[1] + (1 - u - vv) * t[2])
            W.append(w[q] * 2 * A)
    return np.array(P), np.array(W)


def area(v):
    return 0.5 * np.linalg.norm(np.cross(v[0] - v[2], v[1] - v[2]))


def se_composite(orc, trgauss, kind, gam, al, be, tv, sv, L=3):
    tv, sv = np.asarray(tv, float), np.asarray(sv, float)
    At, As = area(tv), area(sv)
    X, WX = composite_points(trgauss, tv, 8, L)
    Y, WY = composite_points(trgauss, sv, 6, L)      # no shared centroid point with degree 8
    z = np.zeros((3, 3), complex)
    nrm = np.cross(sv[0] - sv[2], sv[1] - sv[2])
    nrm /= np.linalg.norm(nrm)
    gvals = (Y[:, None, :] - sv[None, :, :]) / (2 * As)
    for x, wx in zip(X, WX):
        r = x - Y
        R = np.linalg.norm(r, axis=1)
        gR = gam * R
        f = (x - tv) / (2 * At)
        s, v, gg = orc.wiltonints(sv[0], sv[1], sv[2], x)
        if kind == 0:
            Gr = (np.expm1(-gR) + gR - 0.5 * gR ** 2) / (4 * np.pi * R)
            ff = np.einsum("ik,qjk->qij", f, gvals)
            z += wx * np.einsum("q,q,qij->ij", WY, Gr, al * ff) + wx * np.sum(WY * Gr) * be / (At * As)
            h = np.dot(x - sv[0], nrm)
            rho = x - h * nrm
            IG = (s[1] - gam * s[2] + 0.5 * gam ** 2 * s[3]) / (4 * np.pi)
            IGy = (v[1] + s[1] * rho - gam * (v[2] + s[2] * rho) + 0.5 * gam ** 2 * (v[3] + s[3] * rho)) / (4 * np.pi)
            for i in range(3):
                for j in range(3):
                    z[i, j] += wx * (al * np.dot(x - tv[i], IGy - IG * sv[j]) + be * 4 * IG) / (4 * At * As)
        else:
            e = np.exp(-gR)
            hh = (-(gR + 1) * e + 1 - 0.5 * gR ** 2) / (4 * np.pi * R ** 3)
            for j in range(3):
                cr = np.cross(r, gvals[:, j, :])
                for i in range(3):
                    z[i, j] += wx * np.sum(WY * hh * (cr @ f[i]))
            IdG = (-gg[0] - 0.5 * gam ** 2 * gg[1]) / (4 * np.pi)
            for i in range(3):
                for j in range(3):
                    z[i, j] += wx * np.dot(np.cross(x - sv[j], x - tv[i]), IdG) / (4 * At * As)
    return z


def wilton_brute(trgauss, p1, p2, p3, x, L=4):
    """Quadrature comparator of the analytic integrals (s, v, g as in the oracle)."""
    tri = np.array([p1, p2, p3], float)
    Y, W = composite_points(trgauss, tri, 8, L)
    n = np.cross(tri[0] - tri[2], tri[1] - tri[2])
    n /= np.linalg.norm(n)
    d = np.dot(x - tri[0], n)
    rho = x - d * n
    R = np.linalg.norm(Y - x, axis=1)
    s = np.array([np.sum(W * d / R ** 3), np.sum(W / R), np.sum(W), np.sum(W * R), np.sum(W * R ** 2),
                  np.sum(W * R ** 3)])
    yr = Y - rho
    v = np.array([(W / R ** 3) @ yr, (W / R) @ yr, W @ yr, (W * R) @ yr, (W * R ** 2) @ yr])
    g = np.array([-(v[0] - s[0] * n), v[1] - d * s[1] * n, 2 * (v[2] - d * s[2] * n), 3 * (v[3] - d * s[3] * n)])
    return s, v, g


def farfield(bj, u, X, Y, kind=0, gamma=1.0j, amplitude=1.0, degree=10):
    """Far field of the current sum_b u_b f_b at directions Y, summed as S = sum w K J over all
    rule points (numpy arrangement, independent of the oracle's per-cell loop):
    kind 0 (y x S) x y, 1 y x S, 2 y x S_n with S_n = sum w K (J x n), 3 as 0 with K = expm1."""
    mesh = X.geo
    uv, w = bj.trgauss(degree)
    V = mesh.vertices[mesh.cells]
    A = mesh.areas()
    nrm = np.cross(V[:, 0] - V[:, 2], V[:, 1] - V[:, 2])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    P = (uv[0][None, :, None] * V[:, 0:1, :] + uv[1][None, :, None] * V[:, 1:2, :]
         + (1 - uv[0] - uv[1])[None, :, None] * V[:, 2:3, :])
    W = w[None, :] * 2 * A[:, None]
    J = np.zeros(P.shape, complex)
    for b in range(X.numfunctions()):
        for (cell, r, cf) in X.basisfunction(b):
            J[cell] += cf * u[b] * (P[cell] - V[cell, r][None, :]) / (2 * A[cell])
    if kind == 2:
        J = np.cross(J, nrm[:, None, :])
    out = []
    for y in np.asarray(Y, float):
        ph = gamma * (P @ y)
        K = np.expm1(ph) if kind == 3 else np.exp(ph)
        S = np.einsum("fq,fq,fqk->k", W, K, J)
        v = np.cross(y, S)
        out.append(amplitude * (np.cross(v, y) if kind in (0, 3) else v))
    return np.array(out)


def farfield_magnitudes(bj, u, X, thetas, gamma=1.0j, omega=1.0, c=1.0, degree=8):
    """|E_far| of the current sum_b u_b f_b at directions (sin t, 0, cos t), as test_farfield.jl:8-40:
    MWFarField3D(gamma): integrand (y x (exp(gamma y.x) J)) x y, scaled by i omega / (4 pi c)."""
    Y = np.stack([np.sin(thetas), 0 * np.asarray(thetas), np.cos(thetas)], 1)
    ff = farfield(bj, u, X, Y, 0, gamma, 1.0, degree) * (1j * omega / (4 * np.pi * c))
    return np.linalg.norm(ff, axis=1)


# test/test_farfield.jl:34-40 (sphere824.in = sphere2.in, EFIE solution at kappa = 1)
FARFIELD_GOLDEN = [0.6426645162880508, 0.5222644345134998, 0.28287430186444407, 0.3873439919314161,
                   0.6788129041674158, 0.8800967217249486, 0.9487895695366841]


def mie_pec_sphere(ka, thetas):
    """Scattering amplitudes S1, S2 of a perfectly conducting sphere (Bohren & Huffman,
    "Absorption and Scattering of Light by Small Particles", 1983, Sec. 4.4, with the
    conductor limit of the coefficients): a_n = psi_n'(x)/xi_n'(x), b_n = psi_n(x)/xi_n(x),
    psi_n = x j_n(x), xi_n = x h_n(x), x = ka, and the angular functions pi_n, tau_n by
    their upward recurrences.  For an x-polarised plane wave along z the far field
    r |E_s| is |S2(theta)|/k in the plane phi = 0 and |S1(theta)|/k in phi = 90 degrees."""
    from scipy.special import spherical_jn, spherical_yn
    x = float(ka)
    nmax = int(x + 4 * x ** (1 / 3) + 12)
    n = np.arange(1, nmax + 1)
    jn, jnp = spherical_jn(n, x), spherical_jn(n, x, derivative=True)
    yn, ynp = spherical_yn(n, x), spherical_yn(n, x, derivative=True)
    hn, hnp = jn + 1j * yn, jnp + 1j * ynp
    a = (jn + x * jnp) / (hn + x * hnp)
    b = (x * jn) / (x * hn)
    c = (2 * n + 1) / (n * (n + 1))
    S1, S2 = [], []
    for th in thetas:
        mu = np.cos(th)
        pi = np.zeros(nmax + 1)
        pi[1] = 1.0
        for k in range(2, nmax + 1):
            pi[k] = (2 * k - 1) / (k - 1) * mu * pi[k - 1] - k / (k - 1) * pi[k - 2]
        tau = n * mu * pi[1:] - (n + 1) * pi[:-1]
        S1.append(np.sum(c * (a * pi[1:] + b * tau)))
        S2.append(np.sum(c * (a * tau + b * pi[1:])))
    return np.array(S1), np.array(S2)


def dq_double_layer(trgauss, gamma, tv, sv, deg_o, deg_i, rotate_test=False):
    """3x3 block z[i, j] = sum_p sum_q w_p w_q t_i(x_p) . (grad G(x_p - y_q) x g_j(y_q)) of the
    double layer by plain double quadrature (numpy arrangement), grad G = -(gamma + 1/R) G r / R,
    G = exp(-gamma R) / (4 pi R), with RT shapes f = (x - v) / (2A) as test functions t_i = f_i,
    or their rotations t_i = n x f_i (n the test cell's normal) when rotate_test is set."""
    tv, sv = np.asarray(tv, float), np.asarray(sv, float)
    At, As = area(tv), area(sv)
    nrm = np.cross(tv[0] - tv[2], tv[1] - tv[2])
    nrm /= np.linalg.norm(nrm)

    def pts(v, deg):
        uv, w = trgauss(deg)
        P = uv[0][:, None] * v[0] + uv[1][:, None] * v[1] + (1 - uv[0] - uv[1])[:, None] * v[2]
        return P, w * 2 * area(v)

    X, wx = pts(tv, deg_o)
    Y, wy = pts(sv, deg_i)
    F = (X[None, :, :] - tv[:, None, :]) / (2 * At)            # [i, p, 3]
    if rotate_test:
        F = np.cross(nrm[None, None, :], F)
    Gv = (Y[None, :, :] - sv[:, None, :]) / (2 * As)           # [j, q, 3]
    r = X[:, None, :] - Y[None, :, :]                          # [p, q, 3]
    R = np.linalg.norm(r, axis=2)
    grad = (-(gamma + 1 / R) * np.exp(-gamma * R) / (4 * np.pi * R) / R)[:, :, None] * r
    W = wx[:, None] * wy[None, :]
    z = np.zeros((3, 3), complex)
    for j in range(3):
        cr = np.cross(grad, Gv[j][None, :, :])                 # [p, q, 3]
        for i in range(3):
            z[i, j] = np.sum(W * np.einsum("pk,pqk->pq", F[i], cr))
    return z


def read_julia_complex_vector(path):
    """A column of Julia complex literals ("a + bim"), one per line."""
    vals = []
    for line in open(path):
        line = line.strip()
        if line:
            vals.append(complex(line.replace("im", "j").replace(" ", "")))
    return np.array(vals)
