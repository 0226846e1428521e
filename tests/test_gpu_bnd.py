"""Boundary integrals (dCut(..., vb=BND), Integrate(..., VOL_or_BND=BND)) on
the GPU: the level set's trace on the boundary simplices, their cut rules and
the trace space's element matrices through the (D-1)-dimensional kernels,
coefficients at the ambient coordinates.

Pinned by closed forms on the unit cube / square cut by the plane x = 0.3
(P1 exact): the NEG part of the cube's surface has area 1 + 4 * 0.3 = 2.2,
its boundary curve on the surface length 4; int y = 1.1, int z = 1.1,
int y^2 = 1/3 + 0.3 + 2 * 0.1, int sin z = 1.6 (1 - cos 1) + 0.3 sin 1; the
surface gradient of x has |grad_G x|^2 = 1 on the four side faces and 0 on
x = 0, so c_x^T K c_x = 1.2 and K 1 = 0; the NEG part of the square's
boundary has length 1.6 and meets the interface in 2 points."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def X():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import xfem
    return xfem


def cube(X, n=4):
    mesh = X.MakeStructured3DMesh(nx=n, ny=n, nz=n)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.x - 0.3, lset)
    return mesh, lset


def test_boundary_integrals_3d(X):
    mesh, lset = cube(X)
    I = lambda dt, cf, order=4: X.Integrate({"levelset": lset, "domain_type": dt}, cf=cf, mesh=mesh,
                                            VOL_or_BND=X.BND, order=order)
    assert abs(I(X.NEG, 1) - 2.2) < 1e-13
    assert abs(I(X.POS, 1) - 3.8) < 1e-13
    assert abs(I(X.IF, 1) - 4.0) < 1e-13
    assert abs(I(X.NEG, X.y) - 1.1) < 1e-13
    assert abs(I(X.NEG, X.y * X.y) - (1 / 3 + 0.3 + 0.2)) < 1e-13
    ref = 1.6 * (1 - math.cos(1.0)) + 0.3 * math.sin(1.0)
    assert abs(I(X.NEG, X.sin(X.z), order=12) - ref) < 1e-10
    # the same through cf * dCut(..., vb=BND)
    v = X.Integrate(X.y * X.dCut(lset, X.NEG, vb=X.BND, order=4), mesh)
    assert abs(v - 1.1) < 1e-13
    ew = X.Integrate({"levelset": lset, "domain_type": X.NEG}, cf=1, mesh=mesh, VOL_or_BND=X.BND, order=2,
                     element_wise=True)
    assert len(ew) == mesh.nse and abs(ew.sum() - 2.2) < 1e-13 and ew.min() >= 0.0


def test_boundary_integrals_2d(X):
    mesh = X.MakeStructured2DMesh(nx=4, ny=4)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.x - 0.3, lset)
    I = lambda dt, cf: X.Integrate({"levelset": lset, "domain_type": dt}, cf=cf, mesh=mesh, VOL_or_BND=X.BND,
                                   order=4)
    assert abs(I(X.NEG, 1) - 1.6) < 1e-14
    assert abs(I(X.IF, 1) - 2.0) < 1e-14
    assert abs(I(X.NEG, X.y) - (0.5 + 0.3)) < 1e-14


@pytest.mark.parametrize("p", [1, 2])
def test_boundary_forms(X, p):
    mesh, lset = cube(X)
    V = X.H1(mesh, order=p)
    u, v = V.TnT()
    ci = X.CutInfo(mesh, lset)
    bels = ci.GetElementsOfType(X.HASNEG, X.BND)
    m = X.BilinearForm(V)
    m += (u * v) * X.dCut(lset, X.NEG, vb=X.BND, definedonelements=bels)
    m.Assemble(precision=0)
    k = X.BilinearForm(V)
    k += (X.grad(u) * X.grad(v)) * X.dCut(lset, X.NEG, vb=X.BND, definedonelements=bels)
    k.Assemble(precision=0)
    g = X.GridFunction(V)
    one = np.ones(V.ndof)
    g.Set(X.y)
    assert abs(one @ (m.mat @ one) - 2.2) < 1e-13
    assert abs(one @ (m.mat @ g.vec) - 1.1) < 1e-13
    if p == 2:
        g.Set(X.y * X.y)
        assert abs(one @ (m.mat @ g.vec) - (1 / 3 + 0.5)) < 1e-13
    g.Set(X.x)
    assert abs(g.vec @ (k.mat @ g.vec) - 1.2) < 1e-12
    assert np.abs(k.mat @ one).max() < 1e-12
    # a Robin-type form: volume stiffness plus boundary mass
    a = X.BilinearForm(V)
    a += (X.grad(u) * X.grad(v)) * X.dCut(lset, X.NEG) + (u * v) * X.dCut(lset, X.NEG, vb=X.BND)
    a.Assemble(precision=0)
    assert abs(one @ (a.mat @ one) - 2.2) < 1e-12
    # coefficient function with a P1 grid function on the boundary: int (1 + y)
    h = X.GridFunction(X.H1(mesh, order=1))
    h.Set(1 + X.y)
    c = X.BilinearForm(V)
    c += (h * u * v) * X.dCut(lset, X.NEG, vb=X.BND)
    c.Assemble(precision=0)
    assert abs(one @ (c.mat @ one) - 3.3) < 1e-12
    # source terms on the boundary
    f = X.LinearForm(V)
    f += X.z * v * X.dCut(lset, X.NEG, vb=X.BND)
    f.Assemble()
    assert abs(f.vec.sum() - 1.1) < 1e-13
    f2 = X.LinearForm(V)
    f2 += X.sin(X.z) * v * X.dCut(lset, X.NEG, vb=X.BND, order=12)
    f2.Assemble()
    assert abs(f2.vec.sum() - (1.6 * (1 - math.cos(1.0)) + 0.3 * math.sin(1.0))) < 1e-10
