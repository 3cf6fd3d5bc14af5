"""Raviart-Thomas (RWG) spaces and the element-to-dof map.

Follows ``raviartthomas(mesh; sort=:spacefillingcurve)`` and
``raviartthomas(mesh, cellpairs)`` in ``src/bases/rtspace.jl:35-105``: one dof per
interior edge, made of two shapes ``Shape(c1, e1, +1)`` and ``Shape(c2, e2, -1)``
where ``e`` is the local index of the cell vertex opposite the shared edge (the
coefficient is +-1, not the edge length).  ``assemblydata`` follows
``src/bases/basis.jl:235-294``: for every cell ``c`` and local shape ``r`` the list
of ``(dof, coeff)`` pairs using it.

Conventions owned by this implementation (the reference takes them from
CompScienceMeshes ``skeleton``/``cellpairs``): ``c1`` is the cell in which the edge
runs from its lower to its higher vertex index in the cell's cyclic order, and
dofs are ordered along a Morton (Z-order) curve of the edge midpoints.  When a
space comes from the reference itself, its own map is passed through unchanged.
Indices are 0-based here.
"""
import numpy as np

from .mesh import Mesh


class RTBasis:
    """A div-conforming space: ``fns`` in CSR form over ``mesh``.

    ``shape_ptr[ndofs+1]``, ``shape_cell``, ``shape_ref`` (0-based local vertex
    opposite the edge), ``shape_coeff``: the shapes of dof ``b`` are entries
    ``shape_ptr[b]:shape_ptr[b+1]`` (``RTBasis.fns`` of rtspace.jl:3-7).
    """

    def __init__(self, mesh, shape_ptr, shape_cell, shape_ref, shape_coeff, pos=None):
        self.geo = mesh
        self.shape_ptr = np.ascontiguousarray(shape_ptr, dtype=np.int32)
        self.shape_cell = np.ascontiguousarray(shape_cell, dtype=np.int32)
        self.shape_ref = np.ascontiguousarray(shape_ref, dtype=np.int32)
        self.shape_coeff = np.ascontiguousarray(shape_coeff, dtype=np.float64)
        self.pos = pos
        n = len(self.shape_ptr) - 1
        if n < 0 or self.shape_ptr[0] != 0 or np.any(np.diff(self.shape_ptr) < 0):
            raise ValueError("shape_ptr must be a non-decreasing CSR pointer starting at 0")
        m = int(self.shape_ptr[-1])
        if not (len(self.shape_cell) == len(self.shape_ref) == len(self.shape_coeff) == m):
            raise ValueError("shape arrays disagree with shape_ptr")
        if m and (self.shape_cell.min() < 0 or self.shape_cell.max() >= mesh.numcells):
            raise ValueError("shape cell out of range")
        if m and (self.shape_ref.min() < 0 or self.shape_ref.max() > 2):
            raise ValueError("shape refid out of range")

    def numfunctions(self):
        return len(self.shape_ptr) - 1

    def basisfunction(self, b):
        lo, hi = self.shape_ptr[b], self.shape_ptr[b + 1]
        return [(int(self.shape_cell[k]), int(self.shape_ref[k]), float(self.shape_coeff[k]))
                for k in range(lo, hi)]

    def subset(self, idx):
        """``subset(rt, I)`` (rtspace.jl:16): keep dofs ``idx`` in that order."""
        idx = np.asarray(idx, dtype=np.int64)
        counts = (self.shape_ptr[idx + 1] - self.shape_ptr[idx]).astype(np.int64)
        ptr = np.zeros(len(idx) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum(counts)
        sel = np.concatenate([np.arange(self.shape_ptr[b], self.shape_ptr[b + 1]) for b in idx]) \
            if len(idx) else np.zeros(0, dtype=np.int64)
        return RTBasis(self.geo, ptr, self.shape_cell[sel], self.shape_ref[sel],
                       self.shape_coeff[sel], None if self.pos is None else self.pos[idx])

    def __len__(self):
        return self.numfunctions()


def _morton3(p):
    """Interleave 21-bit quantised coordinates of points ``p[n, 3]``."""
    lo, hi = p.min(axis=0), p.max(axis=0)
    q = ((p - lo) / np.maximum(hi - lo, 1e-300) * (2 ** 21 - 1)).astype(np.uint64)
    code = np.zeros(len(p), dtype=np.uint64)
    for bit in range(21):
        for ax in range(3):
            code |= ((q[:, ax] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + ax)
    return code


def mesh_edges(mesh):
    """All edges with their adjacent (cell, local-opposite-vertex) incidences.

    Returns ``(edges[ne, 2] (lo, hi), inc_ptr, inc_cell, inc_ref, inc_fwd)``
    where ``inc_fwd`` says whether the cell traverses the edge lo -> hi.
    """
    C = mesh.cells.astype(np.int64)
    nf = len(C)
    a = np.concatenate([C[:, 1], C[:, 2], C[:, 0]])   # edge opposite local vertex 0, 1, 2
    b = np.concatenate([C[:, 2], C[:, 0], C[:, 1]])
    ref = np.repeat(np.arange(3), nf)
    cell = np.tile(np.arange(nf), 3)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    fwd = a < b
    key = lo * (mesh.numvertices + 1) + hi
    order = np.lexsort((cell, key))
    key, lo, hi, ref, cell, fwd = key[order], lo[order], hi[order], ref[order], cell[order], fwd[order]
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    ptr = np.r_[start, len(key)]
    edges = np.stack([lo[start], hi[start]], axis=1)
    return edges, ptr, cell, ref, fwd


def raviartthomas(mesh, sort="spacefillingcurve"):
    """RT space on the interior edges of ``mesh`` (boundary and junction edges dropped)."""
    edges, ptr, cell, ref, fwd = mesh_edges(mesh)
    cnt = np.diff(ptr)
    keep = np.flatnonzero(cnt == 2)
    i0 = ptr[keep]
    i1 = i0 + 1
    # plus cell: the one traversing the edge lo -> hi; ties broken by cell order
    swap = (~fwd[i0]) & fwd[i1]
    p_idx = np.where(swap, i1, i0)
    m_idx = np.where(swap, i0, i1)
    V = mesh.vertices
    ed = edges[keep]
    mid = 0.5 * (V[ed[:, 0]] + V[ed[:, 1]])
    if sort == "spacefillingcurve" and len(keep):
        order = np.lexsort((ed[:, 1], ed[:, 0], _morton3(mid)))
    else:
        order = np.arange(len(keep))
    p_idx, m_idx = p_idx[order], m_idx[order]
    n = len(keep)
    shape_ptr = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    shape_cell = np.empty(2 * n, dtype=np.int32)
    shape_ref = np.empty(2 * n, dtype=np.int32)
    shape_coeff = np.empty(2 * n, dtype=np.float64)
    shape_cell[0::2], shape_cell[1::2] = cell[p_idx], cell[m_idx]
    shape_ref[0::2], shape_ref[1::2] = ref[p_idx], ref[m_idx]
    shape_coeff[0::2], shape_coeff[1::2] = 1.0, -1.0
    ctr = V[mesh.cells].mean(axis=1)
    pos = 0.5 * (ctr[shape_cell[0::2]] + ctr[shape_cell[1::2]])
    return RTBasis(mesh, shape_ptr, shape_cell, shape_ref, shape_coeff, pos)


def assemblydata(space):
    """Element-to-dof map of ``assemblydata`` (basis.jl:235-294), all cells active.

    Returns ``(dof[nf, 3, K] int32, coeff[nf, 3, K] float64)`` with ``dof = -1``
    where a slot is unused, ``K = max_celltonum``.
    """
    nf = space.geo.numcells
    n = space.numfunctions()
    counts = np.zeros((nf, 3), dtype=np.int64)
    np.add.at(counts, (space.shape_cell, space.shape_ref), 1)
    K = max(1, int(counts.max()) if counts.size else 1)
    dof = np.full((nf, 3, K), -1, dtype=np.int32)
    coeff = np.zeros((nf, 3, K), dtype=np.float64)
    fill = np.zeros((nf, 3), dtype=np.int64)
    owner = np.repeat(np.arange(n), np.diff(space.shape_ptr))
    for k in range(len(space.shape_cell)):
        c, r = space.shape_cell[k], space.shape_ref[k]
        s = fill[c, r]
        dof[c, r, s] = owner[k]
        coeff[c, r, s] = space.shape_coeff[k]
        fill[c, r] = s + 1
    return dof, coeff
