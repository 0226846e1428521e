"""Synthetic workloads shared by tests and bench.py (inputs only, no compute).

C1: [-1.5,1.5]^2, 60x60 cells x 2 triangles, phi = |x| - 1 (P1, perturbed).
C2: unit cube, n^3 cells x 6 Kuhn tets, phi = |x - c| - 0.35 + 0.01 sin(7x)
    sin(7y) sin(7z) (P1, perturbed); n = 64 is the full size (1,572,864 tets).
"""
import math

import numpy as np

from ngsxfem_amd.meshes import MakeStructured2DMesh, MakeStructured3DMesh


def p1_perturb(v, eps=1e-14):
    v = np.array(v, dtype=np.float64)
    v[np.abs(v) < eps] = eps
    return v


def config1():
    mesh = MakeStructured2DMesh(nx=60, ny=60, mapping=lambda x, y: (-1.5 + 3 * x, -1.5 + 3 * y))
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    lset = p1_perturb(np.sqrt(x * x + y * y) - 1.0)
    return mesh, lset


def sphere_lset(coords):
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    return p1_perturb(r - 0.35 + 0.01 * np.sin(7 * x) * np.sin(7 * y) * np.sin(7 * z))


def config2(n=64):
    mesh = MakeStructured3DMesh(nx=n, ny=n, nz=n)
    return mesh, sphere_lset(mesh.coords)


def random_quadric_p2(nel, seed=42):
    """C4 inputs: P2 nodal values (10 per tet, this project's dof order) of
    x^T A x + b^T x + c on the reference tet, A sym N(0,1), b N(0,1), c shifted
    by the median of the nodal values so the tet is cut."""
    from itertools import product  # noqa: F401
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nel, 3, 3))
    A = 0.5 * (A + A.transpose(0, 2, 1))
    b = rng.standard_normal((nel, 3))
    nodes = p2_nodes()
    vals = np.einsum("ki,nij,kj->nk", nodes, A, nodes) + b @ nodes.T
    c = -np.median(vals, axis=1)
    return vals + c[:, None]


def p2_nodes():
    """reference coordinates of the 10 P2 Lagrange nodes in dof order"""
    V = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]], float)
    nodes = [V[k] for k in range(4)]
    for i in range(4):
        for j in range(i + 1, 4):
            nodes.append(0.5 * (V[i] + V[j]))
    return np.array(nodes)


def config3(nmesh=256, seed0=1000, nmin=8, nmax=48):
    """Batched uneven configuration: nmesh independent unit-cube meshes of
    n^3 cells x 6 Kuhn tets, n = round(exp(U[ln nmin, ln nmax])), each with
    the P1 interpolant (perturbed) of its own ellipsoid
    ((x-1/2)/a)^2 + ((y-1/2)/b)^2 + ((z-1/2)/c)^2 - 1, semi-axes U[0.15, 0.45],
    and its own H1 degree p in {1, 2, 3} (rule order 2p).  Mesh m draws from
    MT19937(seed0 + m).  Returns the concatenation (coords, el2v with global
    vertex ids, lset) and per-mesh (elem_off, p, n, semi-axes)."""
    coords, el2v, lset, elem_off, ps, ns, axes = [], [], [], [0], [], [], []
    voff = 0
    for m in range(nmesh):
        rng = np.random.Generator(np.random.MT19937(seed0 + m))
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        ax = rng.uniform(0.15, 0.45, 3)
        p = int(rng.integers(1, 4))
        mesh = MakeStructured3DMesh(nx=n, ny=n, nz=n)
        X = mesh.coords
        phi = ((X[:, 0] - 0.5) / ax[0]) ** 2 + ((X[:, 1] - 0.5) / ax[1]) ** 2 + ((X[:, 2] - 0.5) / ax[2]) ** 2 - 1.0
        coords.append(X)
        el2v.append(mesh.el2v + voff)
        lset.append(p1_perturb(phi))
        voff += mesh.nv
        elem_off.append(elem_off[-1] + mesh.ne)
        ps.append(p)
        ns.append(n)
        axes.append(ax)
    return (np.concatenate(coords), np.ascontiguousarray(np.concatenate(el2v), dtype=np.int32),
            np.concatenate(lset), np.array(elem_off, np.int32), np.array(ps, np.int32), np.array(ns),
            np.array(axes))


def config5_spheres(seed=7, nsph=64):
    """Large mixed configuration: the level set is the union (min) of nsph
    spheres, radii U[0.05, 0.15], centres U[0.2, 0.8]^3, drawn from
    MT19937(seed) in the order (centres, radii)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    centres = rng.uniform(0.2, 0.8, (nsph, 3))
    radii = rng.uniform(0.05, 0.15, nsph)
    return centres, radii


def union_of_spheres(x, centres, radii):
    """min_s (|x - c_s| - r_s) for points x (n x 3), numpy"""
    out = np.full(x.shape[0], np.inf)
    for c, r in zip(centres, radii):
        out = np.minimum(out, np.sqrt(((x - c) ** 2).sum(axis=1)) - r)
    return out


def p2_nodal_values_torch(el2v, coords, centres, radii, chunk=1 << 20):
    """P2 nodal values (Lattice<3,2> dof order: 4 vertices, then the edge
    midpoints (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)) of the union-of-spheres
    level set for every element, computed with torch on the device of el2v;
    returns an ne x 10 FP64 tensor."""
    import torch
    dev = el2v.device
    ne = el2v.shape[0]
    C = torch.as_tensor(centres, dtype=torch.float64, device=dev)
    R = torch.as_tensor(radii, dtype=torch.float64, device=dev)
    out = torch.empty((ne, 10), dtype=torch.float64, device=dev)
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for a in range(0, ne, chunk):
        b = min(ne, a + chunk)
        V = coords[el2v[a:b].long()]  # (m, 4, 3)
        nodes = torch.cat([V, torch.stack([0.5 * (V[:, i] + V[:, j]) for i, j in pairs], dim=1)], dim=1)
        best = torch.full(nodes.shape[:2], float("inf"), dtype=torch.float64, device=dev)
        for s in range(C.shape[0]):
            d = torch.linalg.vector_norm(nodes - C[s], dim=2) - R[s]
            best = torch.minimum(best, d)
        out[a:b] = best
    return out
