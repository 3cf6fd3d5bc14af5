"""Triangle surface meshes: the ``.in`` file format and seeded mesh generators.

The ``.in`` format is the one ``readmesh`` reads for the reference's test
fixtures (``test/assets/*.in``): line 1 ``1``, line 2 ``nv nf``, then ``nv``
coordinate lines and ``nf`` lines of 1-based vertex indices.

``meshsphere``/``meshellipsoid``/``meshcube`` are seeded, deterministic
stand-ins for the gmsh-based generators of CompScienceMeshes (gmsh is not
available): Fibonacci points on the sphere triangulated by their convex hull
(F = 2*nv - 4 exactly), an x-stretched sphere, and a structured cube surface
(F = 12 n^2).  All produce outward-oriented triangles.
"""
import numpy as np


class Mesh:
    """Vertices ``[nv, 3]`` float64 and 0-based cells ``[nf, 3]`` int32."""

    def __init__(self, vertices, cells):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        if self.vertices.ndim != 2 or self.vertices.shape[1] != 3:
            raise ValueError("vertices must be [nv, 3]")
        if self.cells.ndim != 2 or self.cells.shape[1] != 3:
            raise ValueError("cells must be [nf, 3]")
        if self.cells.size and (self.cells.min() < 0 or self.cells.max() >= len(self.vertices)):
            raise ValueError("cell index out of range")

    @property
    def numvertices(self):
        return self.vertices.shape[0]

    @property
    def numcells(self):
        return self.cells.shape[0]

    def areas(self):
        v = self.vertices[self.cells]
        return 0.5 * np.linalg.norm(np.cross(v[:, 0] - v[:, 2], v[:, 1] - v[:, 2]), axis=1)

    def mean_edge_length(self):
        v = self.vertices[self.cells]
        e = np.concatenate([v[:, 1] - v[:, 0], v[:, 2] - v[:, 1], v[:, 0] - v[:, 2]])
        return float(np.linalg.norm(e, axis=1).mean())

    def subset(self, cell_ids):
        return Mesh(self.vertices, self.cells[np.asarray(cell_ids, dtype=np.int64)])

    def __len__(self):
        return self.numcells


def readmesh(path):
    """Read a ``.in`` mesh (1-based cells in the file, 0-based in memory)."""
    with open(path) as fh:
        toks = fh.read().split()
    if not toks or toks[0] != "1":
        raise ValueError(f"{path}: not a .in mesh file")
    nv, nf = int(toks[1]), int(toks[2])
    vals = toks[3:]
    if len(vals) < 3 * nv + 3 * nf:
        raise ValueError(f"{path}: truncated")
    V = np.array(vals[:3 * nv], dtype=np.float64).reshape(nv, 3)
    C = np.array(vals[3 * nv:3 * nv + 3 * nf], dtype=np.int64).reshape(nf, 3) - 1
    return Mesh(V, C.astype(np.int32))


def writemesh(path, mesh):
    """Write a ``.in`` mesh readable by ``readmesh`` here and in the reference."""
    with open(path, "w") as fh:
        fh.write("1\n%d\t%d\n" % (mesh.numvertices, mesh.numcells))
        for x, y, z in mesh.vertices:
            fh.write("%.17g\t%.17g\t%.17g\n" % (x, y, z))
        for a, b, c in mesh.cells + 1:
            fh.write("%d\t%d\t%d\n" % (a, b, c))


def _orient_outward(V, C):
    """Flip cells whose normal points towards the origin (star-shaped surfaces)."""
    p = V[C]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    flip = np.einsum("ij,ij->i", n, p.mean(axis=1)) < 0
    C = C.copy()
    C[flip, 1], C[flip, 2] = C[flip, 2].copy(), C[flip, 1].copy()
    return C


def fibonacci_sphere_points(nv):
    i = np.arange(nv, dtype=np.float64) + 0.5
    z = 1.0 - 2.0 * i / nv
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = np.pi * (3.0 - np.sqrt(5.0)) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def meshsphere(radius=1.0, nv=None, h=None, jitter=0.0, seed=None):
    """Unit-sphere-type mesh with ``nv`` Fibonacci vertices (or target edge ``h``).

    ``jitter`` is the standard deviation of Gaussian vertex noise in units of the
    mean edge length; jittered vertices are re-projected onto the sphere and the
    connectivity of the un-jittered hull is kept.
    """
    from scipy.spatial import ConvexHull
    if nv is None:
        if h is None:
            raise ValueError("give nv or h")
        # equilateral estimate: F ~ 29.02 / h^2 on the unit sphere, nv = F/2 + 2
        nv = int(round(0.5 * 4.0 * np.pi / (np.sqrt(3.0) / 4.0 * h * h) + 2))
    P = fibonacci_sphere_points(int(nv))
    C = ConvexHull(P).simplices.astype(np.int32)
    C = _orient_outward(P, C)
    mesh = Mesh(P, C)
    if jitter:
        rng = np.random.default_rng(seed)
        hmean = mesh.mean_edge_length()
        Q = P + rng.normal(0.0, jitter * hmean, size=P.shape)
        Q /= np.linalg.norm(Q, axis=1, keepdims=True)
        mesh = Mesh(Q, C)
    mesh.vertices *= radius
    return mesh


def meshellipsoid(a=2.0, b=1.0, c=1.0, nv=None, h=None, seed=None, jitter=0.0):
    """Sphere mesh scaled by (a, b, c) per axis (the 2x1x1 ellipsoid for a=2)."""
    m = meshsphere(1.0, nv=nv, h=h, jitter=jitter, seed=seed)
    return Mesh(m.vertices * np.array([a, b, c]), m.cells)


def meshcube(n, size=1.0):
    """Structured surface of the cube [0,size]^3: 6 faces of n x n squares, 2 triangles each."""
    verts = {}
    V = []

    def vid(p):
        key = tuple(int(k) for k in p)
        if key not in verts:
            verts[key] = len(V)
            V.append(key)
        return verts[key]

    cells = []
    for axis in range(3):
        for side in (0, n):
            a1, a2 = [k for k in range(3) if k != axis]
            for i in range(n):
                for j in range(n):
                    quad = []
                    for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0, 0, 0]
                        p[axis] = side
                        p[a1] = i + di
                        p[a2] = j + dj
                        quad.append(vid(p))
                    cells.append((quad[0], quad[1], quad[2]))
                    cells.append((quad[0], quad[2], quad[3]))
    V = np.array(V, dtype=np.float64) * (size / n)
    C = np.array(cells, dtype=np.int32)
    # orient outward w.r.t. the cube centre
    p = V[C]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    ctr = np.full(3, 0.5 * size)
    flip = np.einsum("ij,ij->i", nrm, p.mean(axis=1) - ctr) < 0
    C[flip, 1], C[flip, 2] = C[flip, 2].copy(), C[flip, 1].copy()
    return Mesh(V, C)


def meshrectangle(width, height, h):
    """Flat rectangle [0,w]x[0,h] in z=0 split into squares of side ``h``, 2 triangles each."""
    nx, ny = int(round(width / h)), int(round(height / h))
    xs, ys = np.linspace(0, width, nx + 1), np.linspace(0, height, ny + 1)
    V = np.array([(x, y, 0.0) for y in ys for x in xs])
    cells = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 2, a + nx + 1
            cells.append((a, b, c))
            cells.append((a, c, d))
    return Mesh(V, np.array(cells, dtype=np.int32))
