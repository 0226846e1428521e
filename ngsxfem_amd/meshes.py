"""Structured simplex meshes (stand-ins for ngsolve.meshes / netgen meshes).

The reference's tests build their inputs with ``ngsolve.meshes.Make1DMesh``,
``MakeStructured2DMesh`` and ``MakeStructured3DMesh`` (e.g.
tests/pytests/test_straightcutrule.py:14,44,113,195).  NGSolve is not part of
this build, so these generators produce the same geometry with this
project's own numbering:

* 1D: vertices x_i = i/n, segments (i, i+1).
* 2D: each cell [i,i+1]x[j,j+1] is split along its (0,0)-(1,1) diagonal into
  (v00, v10, v11) and (v00, v11, v01).
* 3D: each cube is split into the 6 Kuhn tetrahedra along its main diagonal
  (v000 -> v111), one per permutation of the axes.

Boundary elements (segments in 2D, triangles in 3D, points in 1D) are
listed as well; they carry the same vertex-array layout as the volume
elements one dimension lower.
"""
import itertools

import numpy as np


# local edges / faces of the reference quadrilateral and hexahedron (vertex
# order (0,0),(1,0),(1,1),(0,1), for the hexahedron that square at z = 0 then
# at z = 1); faces in cyclic vertex order
QUAD_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0)]
HEX_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
HEX_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]


class Mesh:
    """Simplex mesh held as flat arrays.

    coords : (nv, D) float64 vertex coordinates
    el2v   : (ne, D+1) int32 element -> vertex table
    bnd2v  : (nse, D) int32 boundary element -> vertex table
    """

    def __init__(self, coords, el2v, bnd2v=None):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        if self.coords.ndim == 1:
            self.coords = self.coords[:, None]
        self.el2v = np.ascontiguousarray(el2v, dtype=np.int32)
        self.dim = self.coords.shape[1]
        # quadrilaterals / hexahedra: 2^D vertices per element (D >= 2)
        self.is_quad = self.dim >= 2 and self.el2v.shape[1] == 2 ** self.dim
        self._bnd2v = None if bnd2v is None else np.ascontiguousarray(bnd2v, dtype=np.int32)
        self._facets = None

    @property
    def bnd2v(self):
        """boundary elements; computed from the facets when not given"""
        if self._bnd2v is None:
            if self.dim == 1:
                self._bnd2v = np.zeros((0, 1), np.int32)
            else:
                f2v, _, f2el = self.facets()
                self._bnd2v = np.ascontiguousarray(f2v[f2el[:, 1] < 0], dtype=np.int32)
        return self._bnd2v

    @bnd2v.setter
    def bnd2v(self, v):
        self._bnd2v = np.ascontiguousarray(v, dtype=np.int32)

    @property
    def nv(self):
        return self.coords.shape[0]

    def dofs(self, p):
        """global dof numbers of the degree-p Lagrange space, per element in
        the element basis' dof order (vertices, then edges (0,1),(0,2),...,
        lower local vertex first): (element dof table ne x ndof, ndof).
        Vertex dofs come first, then one dof per mesh edge (p = 2)."""
        if p == 1:
            return self.el2v, self.nv
        if p != 2:
            raise NotImplementedError("global dof numbering for p = 1, 2")
        e2v, el2e = self.edges()
        return np.ascontiguousarray(np.hstack([self.el2v, self.nv + el2e]).astype(np.int32)), self.nv + len(e2v)

    def bnd_dofs(self, p):
        """global dof numbers of the degree-p Lagrange space on the boundary
        elements, in their basis' dof order (vertices, then the edges
        (0,1),(0,2),(1,2) of the boundary simplex): (nse x ndof table, ndof)"""
        if p == 1:
            return self.bnd2v, self.nv
        if p != 2 or self.is_quad:
            raise NotImplementedError("boundary dof numbering for p = 1, 2 on simplices")
        e2v, _ = self.edges()
        key = e2v[:, 0].astype(np.int64) * self.nv + e2v[:, 1]
        nb = self.bnd2v.shape[1]
        pairs = [(i, j) for i in range(nb) for j in range(i + 1, nb)]
        cols = []
        for i, j in pairs:
            ev = np.sort(self.bnd2v[:, [i, j]], axis=1).astype(np.int64)
            k = ev[:, 0] * self.nv + ev[:, 1]
            pos = np.searchsorted(key, k)
            assert np.all(key[np.minimum(pos, len(key) - 1)] == k), "boundary edge not in the mesh"
            cols.append(self.nv + pos)
        tab = np.hstack([self.bnd2v, np.stack(cols, axis=1)]).astype(np.int32)
        return np.ascontiguousarray(tab), self.nv + len(e2v)

    def _local_edges(self):
        D = self.dim
        if self.is_quad:
            return QUAD_EDGES if D == 2 else HEX_EDGES
        return [(i, j) for i in range(D + 1) for j in range(i + 1, D + 1)]

    def edges(self):
        """(edge2v (nedge, 2) sorted vertex pairs in ascending order, el2e
        (ne, nE)): the element's edges in the local order (0,1),(0,2),...,
        (1,2),... (quads / hexes: the edges of the reference element, see
        QUAD_EDGES / HEX_EDGES).  In 1D the edges are the elements."""
        if getattr(self, "_edges", None) is None:
            pairs = self._local_edges()
            ev = np.stack([np.sort(self.el2v[:, [i, j]], axis=1) for i, j in pairs], axis=1)  # ne x nE x 2
            key = ev[..., 0].astype(np.int64) * self.nv + ev[..., 1]
            uniq, inv = np.unique(key.ravel(), return_inverse=True)
            e2v = np.stack([uniq // self.nv, uniq % self.nv], axis=1).astype(np.int32)
            self._edges = (e2v, np.ascontiguousarray(inv.reshape(key.shape).astype(np.int32)))
        return self._edges

    @property
    def ne(self):
        return self.el2v.shape[0]

    @property
    def nse(self):
        return self.bnd2v.shape[0]

    def GetNE(self, vb="VOL"):
        return self.ne if str(vb).endswith("VOL") else self.nse

    def facets(self):
        """(facet2v (nf, nvf), el2f (ne, nF), f2el (nf, 2)).

        Simplices: facet k of an element is the one opposite local vertex k,
        its vertices sorted.  Quads / hexes: the facets of the reference
        element (QUAD_EDGES / HEX_FACES), vertices in the cyclic order of the
        first element that has the facet.  f2el holds the two neighbouring
        elements (-1 on the boundary)."""
        if self._facets is None:
            D = self.dim
            ne = self.ne
            if self.is_quad:
                locf = QUAD_EDGES if D == 2 else HEX_FACES
            else:
                locf = [tuple(c for c in range(D + 1) if c != k) for k in range(D + 1)]
            nF = len(locf)
            rows = np.concatenate([self.el2v[:, list(f)] for f in locf], axis=0)  # block k: local facet k
            allf = np.sort(rows, axis=1)
            uniq, first_idx, inv = np.unique(allf, axis=0, return_index=True, return_inverse=True)
            inv = inv.reshape(-1)
            el2f = inv.reshape(nF, ne).T.copy().astype(np.int32)
            nf = uniq.shape[0]
            f2el = -np.ones((nf, 2), np.int64)
            elid = np.tile(np.arange(ne), nF)
            order = np.argsort(inv, kind="stable")
            sinv = inv[order]
            sel = elid[order]
            first = np.ones(len(sinv), bool)
            first[1:] = sinv[1:] != sinv[:-1]
            f2el[sinv[first], 0] = sel[first]
            second = ~first
            f2el[sinv[second], 1] = sel[second]
            f2v = rows[first_idx] if self.is_quad else uniq
            self._facets = (np.ascontiguousarray(f2v.astype(np.int32)), el2f, f2el.astype(np.int32))
        return self._facets

    def GetNFacets(self):
        return self.facets()[0].shape[0]

    def facet_dofs(self, p):
        """global dof numbers of the degree-p Lagrange space on each facet
        (nf x ndof_facet): the dofs of a neighbouring element that do not
        involve the vertex opposite the facet."""
        dofs, _ = self.dofs(p)
        f2v, el2f, f2el = self.facets()
        D = self.dim
        e = f2el[:, 0]
        k = np.argmax(el2f[e] == np.arange(len(f2el))[:, None], axis=1)  # local facet = opposite vertex
        locs = []
        for kk in range(D + 1):
            loc = [j for j in range(D + 1) if j != kk]
            if p == 2:
                pairs = [(i, j) for i in range(D + 1) for j in range(i + 1, D + 1)]
                loc += [D + 1 + t for t, (i, j) in enumerate(pairs) if kk not in (i, j)]
            locs.append(loc)
        locs = np.array(locs)
        return np.ascontiguousarray(np.take_along_axis(dofs[e], locs[k], axis=1).astype(np.int32))


def Make1DMesh(n, mapping=None):
    x = np.arange(n + 1, dtype=np.float64) / n
    if mapping is not None:
        x = np.array([mapping(v) for v in x], dtype=np.float64)
    el = np.stack([np.arange(n), np.arange(1, n + 1)], axis=1)
    bnd = np.array([[0], [n]], np.int32)
    return Mesh(x[:, None], el, bnd)


def _structured_quads(nx, ny, mapping):
    """nx x ny unit-square quadrilaterals; vertex (a, b) = a * (ny+1) + b, each
    element's vertices in the reference order (0,0),(1,0),(1,1),(0,1)"""
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    coords = np.stack([i.ravel() / nx, j.ravel() / ny], axis=1).astype(np.float64)
    if mapping is not None:
        coords = np.array([mapping(*p) for p in coords], dtype=np.float64)
    ci, cj = [a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")]
    vid = lambda a, b: a * (ny + 1) + b  # noqa: E731
    el2v = np.stack([vid(ci, cj), vid(ci + 1, cj), vid(ci + 1, cj + 1), vid(ci, cj + 1)], axis=1)
    return Mesh(coords, el2v)


def _structured_hexes(nx, ny, nz, mapping):
    """unit-cube hexahedra; vertex (a, b, c) = (a (ny+1) + b) (nz+1) + c, each
    element's vertices in the reference order: the square (0,0),(1,0),(1,1),
    (0,1) at z = 0, then at z = 1"""
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coords = np.stack([i.ravel() / nx, j.ravel() / ny, k.ravel() / nz], axis=1).astype(np.float64)
    if mapping is not None:
        coords = np.array([mapping(*p) for p in coords], dtype=np.float64)
    ci, cj, ck = [a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")]
    vid = lambda a, b, c: (a * (ny + 1) + b) * (nz + 1) + c  # noqa: E731
    el2v = np.stack([vid(ci, cj, ck), vid(ci + 1, cj, ck), vid(ci + 1, cj + 1, ck), vid(ci, cj + 1, ck),
                     vid(ci, cj, ck + 1), vid(ci + 1, cj, ck + 1), vid(ci + 1, cj + 1, ck + 1),
                     vid(ci, cj + 1, ck + 1)], axis=1)
    return Mesh(coords, el2v)


def MakeStructured2DMesh(quads=False, nx=10, ny=10, mapping=None):
    if quads:
        return _structured_quads(nx, ny, mapping)
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    coords = np.stack([i.ravel() / nx, j.ravel() / ny], axis=1).astype(np.float64)
    if mapping is not None:
        coords = np.array([mapping(*p) for p in coords], dtype=np.float64)

    def vid(a, b):
        return a * (ny + 1) + b

    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ci = ci.ravel()
    cj = cj.ravel()
    v00, v10, v11, v01 = vid(ci, cj), vid(ci + 1, cj), vid(ci + 1, cj + 1), vid(ci, cj + 1)
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v11, v01], axis=1)
    el = np.stack([t1, t2], axis=1).reshape(-1, 3)
    bnd = []
    for a in range(nx):
        bnd.append((vid(a, 0), vid(a + 1, 0)))
        bnd.append((vid(a + 1, ny), vid(a, ny)))
    for b in range(ny):
        bnd.append((vid(nx, b), vid(nx, b + 1)))
        bnd.append((vid(0, b + 1), vid(0, b)))
    return Mesh(coords, el, np.array(bnd, np.int32))


_KUHN = [p for p in itertools.permutations(range(3))]


def MakeStructured3DMesh(hexes=False, nx=10, ny=None, nz=None, mapping=None):
    if hexes:
        return _structured_hexes(nx, nx if ny is None else ny, nx if nz is None else nz, mapping)
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coords = np.stack([i.ravel() / nx, j.ravel() / ny, k.ravel() / nz], axis=1).astype(np.float64)
    if mapping is not None:
        coords = np.array([mapping(*p) for p in coords], dtype=np.float64)

    def vid(a, b, c):
        return (a * (ny + 1) + b) * (nz + 1) + c

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.ravel()
    cj = cj.ravel()
    ck = ck.ravel()
    tets = []
    for perm in _KUHN:
        off = np.zeros(3, np.int64)
        verts = [vid(ci, cj, ck)]
        for ax in perm:
            off = off.copy()
            off[ax] += 1
            verts.append(vid(ci + off[0], cj + off[1], ck + off[2]))
        tets.append(np.stack(verts, axis=1))
    el = np.stack(tets, axis=1).reshape(-1, 4)
    return Mesh(coords, el)
