"""Element-boundary forms (dCut(..., element_boundary=True),
SymbolicCutBilinearFormIntegrator::T_CalcElementMatrixEBAdd,
symboliccutbfi.cpp:488-640): the cut rules of every element's facets with
the volume basis, on the GPU.

Pinned by an independent geometric computation: for the plane level set
x - 0.3 (exact in P1) each facet triangle is clipped by the half space
(Sutherland-Hodgman in numpy), giving the NEG part's area and its moments of
y and y^2 (a degree-2 exact rule on the clipped polygon's fan triangles),
and the interface segment's length; with the partition of unity,
1^T M_T 1 = sum over T's facets of the NEG areas, c_y^T M_T 1 and
c_{y^2}^T M_T 1 the moments (P1 / P2 interpolants of y, y^2 are exact), and
the interface form gives the segment lengths."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def X():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import xfem
    return xfem


def clip(P, phi):
    """polygon P (k x 3) clipped to phi < 0 (phi linear, evaluated per vertex)"""
    out = []
    k = len(P)
    for i in range(k):
        a, b = P[i], P[(i + 1) % k]
        fa, fb = phi(a), phi(b)
        if fa < 0:
            out.append(a)
        if (fa < 0) != (fb < 0):
            out.append(a + fa / (fa - fb) * (b - a))
    return np.array(out)


def poly_moments(Q):
    """area and integrals of y, y^2 over the planar polygon Q (fan)"""
    A = my = my2 = 0.0
    for i in range(1, len(Q) - 1):
        a, b, c = Q[0], Q[i], Q[i + 1]
        ar = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        mids = [(a + b) / 2, (b + c) / 2, (c + a) / 2]  # degree-2 exact
        A += ar
        my += ar * (a[1] + b[1] + c[1]) / 3
        my2 += ar * sum(m[1] ** 2 for m in mids) / 3
    return A, my, my2


def facet_refs(mesh, e, phi):
    V = mesh.coords[mesh.el2v[e]]
    A = my = my2 = L = 0.0
    for k in range(4):
        T = V[[j for j in range(4) if j != k]]
        Q = clip(T, phi)
        if len(Q) >= 3:
            a, b, c = poly_moments(Q)
            A, my, my2 = A + a, my + b, my2 + c
        f = np.array([phi(t) for t in T])
        cuts = [T[i] + f[i] / (f[i] - f[j]) * (T[j] - T[i]) for i in range(3) for j in range(i + 1, 3)
                if (f[i] < 0) != (f[j] < 0)]
        if len(cuts) == 2:
            L += np.linalg.norm(cuts[0] - cuts[1])
    return A, my, my2, L


@pytest.mark.parametrize("p", [1, 2])
def test_element_boundary_forms(X, p):
    mesh = X.MakeStructured3DMesh(nx=4, ny=4, nz=4)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.x - 0.3, lset)
    phi = lambda q: q[0] - 0.3
    ci = X.CutInfo(mesh, lset)
    els = ci.GetElementsOfType(X.HASNEG)
    V = X.H1(mesh, order=p)
    u, v = V.TnT()
    dofs, _ = V.dofs()
    gy, gy2 = X.GridFunction(V), X.GridFunction(V)
    gy.Set(X.y)
    gy2.Set(X.y * X.y)
    neg = ((u * v) * X.dCut(lset, X.NEG, element_boundary=True, definedonelements=els)).parts[0]
    e, M = neg.element_matrices(mesh, p, 0)
    itf = ((u * v) * X.dCut(lset, X.IF, element_boundary=True, definedonelements=els)).parts[0]
    ei, MI = itf.element_matrices(mesh, p, 0)
    cf = ((1 + X.y) * u * v * X.dCut(lset, X.NEG, element_boundary=True, definedonelements=els)).parts[0]
    ec, MC = cf.element_matrices(mesh, p, 0)
    assert np.array_equal(e, ei) and np.array_equal(e, ec) and len(e) > 0
    for i, el in enumerate(e):
        A, my, my2, L = facet_refs(mesh, el, phi)
        one = np.ones(M.shape[1])
        cy, cy2 = gy.vec[dofs[el]], gy2.vec[dofs[el]]
        assert abs(one @ M[i] @ one - A) < 1e-14
        assert abs(cy @ M[i] @ one - my) < 1e-14
        if p == 2:
            assert abs(cy2 @ M[i] @ one - my2) < 1e-14
        assert abs(one @ MI[i] @ one - L) < 1e-14
        assert abs(one @ MC[i] @ one - (A + my)) < 1e-14
    # gradients: K_T 1 = 0 on the element boundary too
    k = ((X.grad(u) * X.grad(v)) * X.dCut(lset, X.NEG, element_boundary=True, definedonelements=els)).parts[0]
    _, K = k.element_matrices(mesh, p, 0)
    assert np.abs(K.sum(axis=2)).max() < 1e-12 * np.abs(K).max()
    # assembled over all elements: every interior facet twice
    a = X.BilinearForm(V)
    a += (u * v) * X.dCut(lset, X.NEG, element_boundary=True)
    a.Assemble(precision=0)
    tot = sum(facet_refs(mesh, el, phi)[0] for el in range(mesh.ne))
    one = np.ones(V.ndof)
    assert abs(one @ (a.mat @ one) - tot) < 1e-12
