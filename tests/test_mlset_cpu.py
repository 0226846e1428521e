"""The multiple-level-set rules of the CPU oracle (or_mlset_rule, restating
cutint/mlsetintegration.cpp:8-158 and straightcutrule.cpp:679-800) pinned
by the known answers of the reference's tests/pytests/test_mlset_integration.py
(areas, lengths, point values of polygons / polyhedra bounded by straight
level sets) on generically distorted structured meshes of the same domains,
by the one-level-set case (equal to the pinned straight cut rule) and by
partition invariants."""
import math

import numpy as np
import pytest

from ngsxfem_amd.meshes import MakeStructured2DMesh, MakeStructured3DMesh

NEG, POS, IF = 0, 1, 2


def mesh2d(x0, x1, y0, y1, n):
    def f(x, y):
        return (x0 + (x1 - x0) * (x + 0.04 * math.sin(math.pi * x) * math.sin(2 * math.pi * y + 0.3)),
                y0 + (y1 - y0) * (y + 0.04 * math.sin(math.pi * y) * math.sin(2 * math.pi * x + 0.7)))
    return MakeStructured2DMesh(nx=n, ny=n, mapping=f)


def mesh3d(a, n):
    def f(x, y, z):
        s = math.sin
        return (-a + 2 * a * (x + 0.03 * s(math.pi * x) * s(2 * math.pi * y + 0.3) * s(math.pi * z + 0.1)),
                -a + 2 * a * (y + 0.03 * s(math.pi * y) * s(2 * math.pi * z + 0.5) * s(math.pi * x + 0.2)),
                -a + 2 * a * (z + 0.03 * s(math.pi * z) * s(2 * math.pi * x + 0.9) * s(math.pi * y + 0.4)))
    return MakeStructured3DMesh(nx=n, ny=n, nz=n, mapping=f)


def lsets_at(mesh, fs):
    c = mesh.coords
    return np.stack([np.array([f(*p) for p in c]) for f in fs])


def integrate(orc, mesh, lv, dtts, order, f=None):
    """sum over elements of the oracle's mlset rule, mapped to the element"""
    D = mesh.dim
    tot = 0.0
    dtts = np.atleast_2d(np.asarray(dtts, dtype=np.int32))
    for e in range(mesh.ne):
        vs = mesh.el2v[e]
        l = lv[:, vs]
        V = mesh.coords[vs]
        pts, w = orc.mlset_rule(D, l, dtts, order, V=V)
        if len(w) == 0:
            continue
        F, _, ad = orc.jacobian(V)
        x = V[D][None, :] + pts[:, :D] @ F.T
        fv = np.ones(len(w)) if f is None else np.array([f(*p) for p in x])
        tot += float(np.sum(fv * w)) * ad
    return tot


@pytest.fixture(scope="module")
def tri2d(orc):
    mesh = mesh2d(-1, 1, -0.5, 1.5, 9)
    lv = lsets_at(mesh, [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y])
    return mesh, lv


def test_2d_triangle_measures(orc, tri2d):
    """test_mlset_integration.py:37-142: area 1/2, complement 3.5, sides 1,
    sqrt(5/4), sqrt(5/4), corner values of x + y"""
    mesh, lv = tri2d
    assert abs(integrate(orc, mesh, lv, [(NEG, NEG, NEG)], 0) - 0.5) < 1e-12
    comp = [t for t in [(a, b, c) for a in (NEG, POS) for b in (NEG, POS) for c in (NEG, POS)] if t != (NEG, NEG, NEG)]
    assert abs(integrate(orc, mesh, lv, comp, 0) - 3.5) < 1e-12
    assert abs(integrate(orc, mesh, lv, [(IF, NEG, NEG)], 0) - 1) < 1e-12
    assert abs(integrate(orc, mesh, lv, [(NEG, IF, NEG)], 0) - math.sqrt(5 / 4)) < 1e-12
    assert abs(integrate(orc, mesh, lv, [(NEG, NEG, IF)], 0) - math.sqrt(5 / 4)) < 1e-12
    per = [(IF, NEG, NEG), (NEG, IF, NEG), (NEG, NEG, IF)]
    assert abs(integrate(orc, mesh, lv, per, 0) - 1 - math.sqrt(5)) < 1e-12
    xy = lambda x, y: x + y  # noqa: E731
    assert abs(integrate(orc, mesh, lv, [(IF, IF, NEG)], 0, xy) - 1.5) < 1e-12
    assert abs(integrate(orc, mesh, lv, [(IF, NEG, IF)], 0, xy) - 0.5) < 1e-12
    assert abs(integrate(orc, mesh, lv, [(NEG, IF, IF)], 0, xy)) < 1e-12
    cnr = [(IF, IF, NEG), (IF, NEG, IF), (NEG, IF, IF)]
    assert abs(integrate(orc, mesh, lv, cnr, 0, xy) - 2.0) < 1e-12


def test_2d_high_order(orc):
    """test_mlset_integration.py:145-169: int_[0,1]^2 x(1-x)y(1-y) = 1/36"""
    mesh = mesh2d(-0.2, 1.2, -0.2, 1.2, 7)
    lv = lsets_at(mesh, [lambda x, y: -y, lambda x, y: x - 1, lambda x, y: y - 1, lambda x, y: -x])
    r = integrate(orc, mesh, lv, [(NEG, NEG, NEG, NEG)], 4, lambda x, y: x * (1 - x) * y * (1 - y))
    assert abs(r - 1 / 36) < 1e-12


def test_3d_cube(orc):
    """test_mlset_integration.py:304-415: the cube [-1/2,1/2]^3 inside
    [-0.8,0.8]^3: volume 1, complement 3.096, faces 1, edges 1, corner
    values of x + y + z"""
    mesh = mesh3d(0.8, 5)
    fs = [lambda x, y, z: x - 0.5, lambda x, y, z: x + 0.5, lambda x, y, z: y - 0.5, lambda x, y, z: y + 0.5,
          lambda x, y, z: z - 0.5, lambda x, y, z: z + 0.5]
    lv = lsets_at(mesh, fs)
    cube = (NEG, POS, NEG, POS, NEG, POS)
    assert abs(integrate(orc, mesh, lv, [cube], 0) - 1) < 1e-12
    import itertools
    comp = [t for t in itertools.product((NEG, POS), repeat=6) if t != cube]
    assert abs(integrate(orc, mesh, lv, comp, 0) - 3.096) < 1e-12
    for i in range(6):
        side = list(cube)
        side[i] = IF
        assert abs(integrate(orc, mesh, lv, [tuple(side)], 0) - 1) < 1e-12
    edges = [(IF, POS, IF, POS, NEG, POS), (IF, POS, NEG, IF, NEG, POS), (NEG, IF, NEG, POS, IF, POS),
             (NEG, POS, NEG, IF, NEG, IF)]
    for ed in edges:
        assert abs(integrate(orc, mesh, lv, [ed], 0) - 1) < 1e-12
    pts = {(IF, POS, IF, POS, IF, POS): 1.5, (IF, POS, IF, POS, NEG, IF): 0.5, (NEG, IF, NEG, IF, NEG, IF): -1.5,
           (IF, POS, NEG, IF, NEG, IF): -0.5}
    for dom, val in pts.items():
        assert abs(integrate(orc, mesh, lv, [dom], 0, lambda x, y, z: x + y + z) - val) < 1e-12


def test_3d_codim2_line(orc):
    """test_mlset_integration.py:418-445: the segment x = y, z = 0 between
    x = -1/2 and x = 1/2 has length sqrt(2)"""
    mesh = mesh3d(0.8, 3)
    fs = [lambda x, y, z: x - 0.5, lambda x, y, z: x + 0.5, lambda x, y, z: x - y, lambda x, y, z: z]
    lv = lsets_at(mesh, fs)
    assert abs(integrate(orc, mesh, lv, [(NEG, POS, IF, IF)], 0) - math.sqrt(2)) < 1e-12


@pytest.mark.parametrize("D", [1, 2, 3])
def test_one_level_set_is_the_straight_cut_rule(orc, D):
    rng = np.random.default_rng(D)
    for trial in range(30):
        l = rng.standard_normal(D + 1)
        V = np.vstack([np.eye(D), np.zeros(D)]) + 0.2 * rng.standard_normal((D + 1, D))
        F, Fi, ad = orc.jacobian(V)
        for dt in (NEG, POS, IF):
            if D == 1 and dt == IF:
                continue
            p1, w1 = orc.mlset_rule(D, l[None, :], [(dt,)], 3, V=V)
            r = orc.straight_cut_rule(D, l, dt, 3, Finv=Fi)
            if r is None:
                assert len(w1) == 0
                continue
            p2, w2, _ = r
            assert np.array_equal(p1, p2) and np.array_equal(w1, w2)


@pytest.mark.parametrize("D", [2, 3])
def test_partition_of_the_element(orc, D):
    """the codim-0 parts of all 2^M sign tuples tile the element, and the
    interface of level set 0 splits into its (IF, NEG/POS, ...) parts"""
    import itertools
    rng = np.random.default_rng(10 + D)
    M = 3
    for trial in range(20):
        l = rng.standard_normal((M, D + 1))
        allt = list(itertools.product((NEG, POS), repeat=M))
        _, w = orc.mlset_rule(D, l, allt, 2)
        assert abs(w.sum() - 1 / math.factorial(D)) < 1e-14
        ift = [(IF,) + t for t in itertools.product((NEG, POS), repeat=M - 1)]
        _, wi = orc.mlset_rule(D, l, ift, 2)
        _, w0 = orc.mlset_rule(D, l[:1], [(IF,)], 2)
        assert abs(wi.sum() - w0.sum()) < 1e-13 * max(1.0, w0.sum())


def test_mark_of_type_and_contribution(orc, tri2d):
    """of_type: the element's tuple of CheckIfStraightCut values matches;
    contribution: positive order-0 weight (test_mlset_integration.py:62-77:
    ~of_type(NEG,NEG,NEG) == contribution(complement) on this mesh, and the
    boundary's contribution == hasneg & ~neg)"""
    mesh, lv = tri2d
    ot, co = orc.mlset_mark(2, mesh.el2v, mesh.coords, lv, [(NEG, NEG, NEG)])
    comp = [(a, b, c) for a in (NEG, POS) for b in (NEG, POS) for c in (NEG, POS) if (a, b, c) != (NEG, NEG, NEG)]
    _, co_c = orc.mlset_mark(2, mesh.el2v, mesh.coords, lv, comp)
    assert np.array_equal(~ot, co_c)
    _, co_b = orc.mlset_mark(2, mesh.el2v, mesh.coords, lv, [(IF, NEG, NEG), (NEG, IF, NEG), (NEG, NEG, IF)])
    assert np.array_equal(co_b, co & ~ot)
    assert ot.sum() > 0 and co_b.sum() > 0
