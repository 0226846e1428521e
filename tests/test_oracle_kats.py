"""Pins the FP64 oracle against the reference's closed-form known-answer tests.

Constants and tolerances are transcribed from the reference test-suite:
  tests/pytests/test_straightcutrule.py:11-31, 111-126, 134-154, 166-185,
      191-207, 218-235, 242-280, 37-68
  tests/pytests/test_localgeom.py:24-45
Meshes are this project's structured generators (ngsxfem_amd.meshes), so the
mesh-independent closed forms carry over unchanged.
"""
import math

import numpy as np
import pytest

from ngsxfem_amd.meshes import Make1DMesh, MakeStructured2DMesh, MakeStructured3DMesh, Mesh

NEG, POS, IF = 0, 1, 2


def p1_interp(mesh, f, eps=1e-14):
    """InterpolateToP1 with the +eps perturbation of utils/p1interpol.cpp:106-107."""
    v = np.array([f(*p) for p in mesh.coords], dtype=np.float64)
    v[np.abs(v) < eps] = eps
    return v


def tol(order):
    return 5e-15 * (order + 1) * (order + 1)


# ---------------------------------------------------------------- base rules
@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5, 6, 8, 12])
def test_base_rule_exactness(orc, D, order):
    """sum w xi^alpha == alpha! D! / (|alpha| + D)!  on the unit simplex."""
    xi, w = orc.rule(D, order)
    import itertools
    for alpha in itertools.product(range(order + 1), repeat=D):
        if sum(alpha) > order:
            continue
        exact = math.prod(math.factorial(a) for a in alpha) / math.factorial(sum(alpha) + D)
        approx = float(np.sum(w * np.prod(xi ** np.array(alpha), axis=1)))
        assert abs(approx - exact) < 1e-15 * max(1.0, 10 * order), (alpha, approx, exact)


def test_order0_weights_exact(orc):
    """the order-0 rule weight is the reference volume, bit for bit"""
    for D, vol in [(1, 1.0), (2, 0.5), (3, 1.0 / 6.0)]:
        xi, w = orc.rule(D, 0)
        assert len(w) == 1 and w[0] == vol


# ---------------------------------------------------------------- 1D KAT
@pytest.mark.parametrize("domain", [NEG, POS, IF])
@pytest.mark.parametrize("alpha", [2, 4, 8])
def test_polynomial_ET_Segm(orc, domain, alpha):
    order = alpha
    mesh = Make1DMesh(1)
    x_ast = 0.78522
    lset = np.array([x_ast - x for x in mesh.coords[:, 0]])  # lset_approx.Set(levelset): exact P1
    ref = {POS: x_ast ** (alpha + 1) / (alpha + 1), NEG: (1 - x_ast ** (alpha + 1)) / (alpha + 1),
           IF: x_ast ** alpha}
    val = orc.integrate_mesh(mesh, lset, domain, order, terms=[(1.0, (alpha, 0, 0))])
    assert abs(val - ref[domain]) < tol(order)


# ---------------------------------------------------------------- 2D KATs
@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("domain", [NEG, POS, IF])
@pytest.mark.parametrize("N", [1, 10, 30])
def test_straight_cutted_quad2D(orc, order, domain, N):
    mesh = MakeStructured2DMesh(nx=N, ny=N)
    lset = p1_interp(mesh, lambda x, y: 1 - 2 * x - 2 * y)
    ref = {NEG: 7 / 8, POS: 1 / 8, IF: 1 / math.sqrt(2)}
    val = orc.integrate_mesh(mesh, lset, domain, order)
    assert abs(val - ref[domain]) < tol(order)


@pytest.mark.parametrize("order", [2])
@pytest.mark.parametrize("domain", [IF, NEG, POS])
@pytest.mark.parametrize("dim", [0, 1])
@pytest.mark.parametrize("eps", [1e-1, 1e-2, 5e-3, 1e-3, 0])
def test_orth_cutted_quad2D_epsiloned(orc, order, domain, dim, eps):
    mesh = MakeStructured2DMesh(nx=1, ny=1)
    if dim == 0:
        f = lambda x, y: 1 - 2 * x + eps * (y - 0.5)
    else:
        f = lambda x, y: 1 - 2 * y + eps * (x - 0.5)
    lset = p1_interp(mesh, f)
    ref = {NEG: 0.5, POS: 0.5, IF: math.sqrt(1 + eps * eps / 4)}
    val = orc.integrate_mesh(mesh, lset, domain, order)
    assert abs(val - ref[domain]) < tol(order)


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("domain", [NEG, POS])
@pytest.mark.parametrize("alpha", [0, 1, 2])
@pytest.mark.parametrize("dim", [0, 1])
def test_straight_cutted_quad2D_polynomial(orc, order, domain, alpha, dim):
    mesh = MakeStructured2DMesh(nx=1, ny=1)
    lset = p1_interp(mesh, lambda x, y: 1 - 2 * x - 2 * y)
    val_pos = 2 ** (-alpha - 2) / (alpha * alpha + 3 * alpha + 2)
    ref = {POS: val_pos, NEG: 1.0 / (alpha + 1) - val_pos}
    ex = [0, 0, 0]
    ex[dim] = alpha
    val = orc.integrate_mesh(mesh, lset, domain, order, terms=[(1.0, tuple(ex))])
    assert abs(val - ref[domain]) < tol(order)


@pytest.mark.parametrize("domain", [NEG, POS, IF])
@pytest.mark.parametrize("dim", [0, 1])
def test_orth_cutted_quad2D(orc, domain, dim):
    order = 2
    mesh = MakeStructured2DMesh(nx=1, ny=1)
    lset = p1_interp(mesh, lambda x, y: 1 - 3 * (x if dim == 0 else y))
    ref = {NEG: 2 / 3, POS: 1 / 3, IF: 1.0}
    val = orc.integrate_mesh(mesh, lset, domain, order)
    assert abs(val - ref[domain]) < tol(order)


# ---------------------------------------------------------------- 3D KATs
@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("domain", [NEG, POS, IF])
def test_straight_cutted_quad3D(orc, order, domain):
    mesh = MakeStructured3DMesh(nx=1, ny=1, nz=1)
    lset = p1_interp(mesh, lambda x, y, z: 1 - 2 * x - 2 * y - 2 * z)
    ref = {POS: 1 / 48, NEG: 47 / 48, IF: math.sqrt(3) / 8}
    val = orc.integrate_mesh(mesh, lset, domain, order)
    assert abs(val - ref[domain]) < tol(order)


@pytest.mark.parametrize("domain", [NEG, POS])
@pytest.mark.parametrize("alpha", [0, 1, 2])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_straight_cutted_quad3D_polynomial(orc, domain, alpha, dim):
    order = 4
    mesh = MakeStructured3DMesh(nx=1, ny=1, nz=1)
    lset = p1_interp(mesh, lambda x, y, z: 1 - 2 * x - 2 * y - 2 * z)
    val_pos = 2 ** (-alpha - 3) / (alpha ** 3 + 6 * alpha * alpha + 11 * alpha + 6)
    ref = {POS: val_pos, NEG: 1.0 / (alpha + 1) - val_pos}
    ex = [0, 0, 0]
    ex[dim] = alpha
    val = orc.integrate_mesh(mesh, lset, domain, order, terms=[(1.0, tuple(ex))])
    assert abs(val - ref[domain]) < tol(order)


@pytest.mark.parametrize("domain", [NEG, POS])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_orth_cutted_quad3D(orc, domain, dim):
    order = 2
    mesh = MakeStructured3DMesh(nx=1, ny=1, nz=1)
    lset = p1_interp(mesh, lambda x, y, z: 1 - 2 * (x, y, z)[dim])
    val = orc.integrate_mesh(mesh, lset, domain, order)
    assert abs(val - 0.5) < tol(order)


# ---------------------------------------------------------------- localgeom KAT
def test_cut_triangle_localgeom(orc):
    """test_localgeom.py:24-45: triangle (0,0),(1,0),(0,1) meshed with 4
    elements (red refinement), phi = x + y - 0.25, orders 0..15."""
    coords = np.array([[0, 0], [1, 0], [0, 1], [0.5, 0], [0.5, 0.5], [0, 0.5]], float)
    el = np.array([[0, 3, 5], [3, 1, 4], [5, 4, 2], [3, 4, 5]], np.int32)
    mesh = Mesh(coords, el)
    lset = p1_interp(mesh, lambda x, y: x + y - 0.25)
    for order in range(16):
        neg = orc.integrate_mesh(mesh, lset, NEG, order)
        pos = orc.integrate_mesh(mesh, lset, POS, order)
        t = 5e-16 * (order + 1) ** 2
        assert abs(neg - 1 / 32) < t
        assert abs(pos - 0.5 + 1 / 32) < t
        assert abs(neg + pos - 0.5) < t


# ---------------------------------------------------------------- convergence KAT
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("domain", [NEG, POS, IF])
def test_circle_eoc(orc, order, domain):
    """test_straightcutrule.py:37-68: r = 0.6, EOC > 1.75 over 2^i meshes"""
    r = 0.6
    ref = {POS: 1 - math.pi * r * r / 4, NEG: math.pi * r * r / 4, IF: r * math.pi / 2}
    errs = []
    for i in range(7):
        mesh = MakeStructured2DMesh(nx=2 ** i, ny=2 ** i)
        lset = p1_interp(mesh, lambda x, y: math.sqrt(x * x + y * y) - r)
        val = orc.integrate_mesh(mesh, lset, domain, order)
        errs.append(abs(val - ref[domain]))
    eoc = [math.log(errs[i + 1] / errs[i]) / math.log(0.5) for i in range(len(errs) - 1)]
    assert sum(eoc[1:]) / len(eoc[1:]) > 1.75


def test_config1_weight_sums(orc):
    """2D config: [-1.5,1.5]^2, 60x60x2 triangles, phi = |x| - 1: NEG ~ pi,
    IF ~ 2 pi (O(h^2))."""
    mesh = MakeStructured2DMesh(nx=60, ny=60, mapping=lambda x, y: (-1.5 + 3 * x, -1.5 + 3 * y))
    lset = p1_interp(mesh, lambda x, y: math.sqrt(x * x + y * y) - 1)
    assert int(np.sum(np.abs([math.sqrt(x * x + y * y) - 1 for x, y in mesh.coords]) < 1e-14)) == 12
    neg = orc.integrate_mesh(mesh, lset, NEG, 2)
    ifv = orc.integrate_mesh(mesh, lset, IF, 2)
    assert abs(neg - math.pi) < 5e-3
    assert abs(ifv - 2 * math.pi) < 5e-3
