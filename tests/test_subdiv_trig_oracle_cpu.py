"""Order-0 subdivision rule on triangles (or_subdiv_trig_parts: the boundary
elements of CutInformation::Update with a P2 level set, xfem/cutinfo.cpp:
80-124 with cutint/xintegration.cpp:620-838,1019-1061,1391-1545), pinned by
the straight cut rule for linear level sets at every level, O(h^2)
convergence to the area of a disc, the partition of the triangle, and the
one-sided lattice-cut elements whose leaves all fall to one side."""
import math

import numpy as np
import pytest

NEG, POS, IF = 0, 1, 2
V = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
EDGES = [(0, 1), (0, 2), (1, 2)]


def nodal(f):
    return np.array([f(*v) for v in V] + [f(*(0.5 * (V[i] + V[j]))) for i, j in EDGES])


@pytest.mark.parametrize("L", [0, 1, 3, 5])
def test_linear_level_set_equals_straight_cut(orc, L):
    rng = np.random.default_rng(40 + L)
    for _ in range(40):
        a, b, c = rng.normal(size=3)
        c = c * 0.3
        f = lambda x, y: a * x + b * y + c  # noqa: E731
        l = np.array([f(*v) for v in V])
        if np.all(l > 0) or np.all(l < 0):
            continue
        dom, an, ap = orc.subdiv_trig_parts(nodal(f), 2, L)
        vn = np.sum(orc.straight_cut_rule(2, l, NEG, 0)[1])
        vp = np.sum(orc.straight_cut_rule(2, l, POS, 0)[1])
        assert dom == IF
        assert abs(an - vn) <= 1e-14 and abs(ap - vp) <= 1e-14, (L, an, vn, ap, vp)
        # the same level set given as P1 values (midpoints lifted by the oracle)
        dom1, an1, ap1 = orc.subdiv_trig_parts(l, 1, L)
        assert abs(an1 - vn) <= 1e-14 and abs(ap1 - vp) <= 1e-14


def test_disc_area_converges_quadratically(orc):
    r2 = 0.04
    g = lambda x, y: (x - 0.3) ** 2 + (y - 0.3) ** 2 - r2  # noqa: E731
    err = []
    for L in range(2, 6):
        dom, an, ap = orc.subdiv_trig_parts(nodal(g), 2, L)
        assert dom == IF and abs(an + ap - 0.5) < 1e-15
        err.append(abs(an - math.pi * r2))
    for e0, e1 in zip(err, err[1:]):
        assert e1 < e0 / 3.0, err


def test_partition_random_quadratics(orc):
    rng = np.random.default_rng(7)
    n_if = 0
    for _ in range(200):
        c = rng.normal(size=6) + rng.uniform(-1, 1)
        dom, an, ap = orc.subdiv_trig_parts(c, 2, 3)
        assert abs(an + ap - 0.5) <= 1e-15
        if dom != IF:
            assert (an, ap) == ((0.5, 0.0) if dom == NEG else (0.0, 0.5))
        n_if += dom == IF
    assert n_if > 20


@pytest.mark.parametrize("L", [1, 2, 4])
def test_lattice_cut_but_one_sided(orc, L):
    # an exact zero at an edge midpoint of a negative level set: the lattice
    # sees v >= 0, every leaf is NEG (Simplex::CheckIfCut, relative 1e-14)
    dom, an, ap = orc.subdiv_trig_parts(np.r_[-np.ones(5), 0.0], 2, L)
    assert dom == IF and ap == 0.0 and abs(an - 0.5) < 1e-15
    dom, an, ap = orc.subdiv_trig_parts(np.r_[np.ones(5), -1e-17], 2, L)
    assert dom == IF and an == 0.0 and abs(ap - 0.5) < 1e-15
    # at level 0 the lattice holds only the vertices
    assert orc.subdiv_trig_parts(np.r_[-np.ones(5), 0.0], 2, 0)[0] == NEG


@pytest.mark.parametrize("L", [1, 2, 3])
def test_tet_lattice_cut_but_one_sided(orc, L):
    """the same on tetrahedra (or_subdiv_rule3): the lattice domain is IF,
    but one order-0 rule is empty and the other holds the whole volume 1/6,
    so CutInformation::Update (part volumes) marks the element uncut"""
    dom, parts = orc.subdiv_rule3(np.r_[-np.ones(9), 0.0], 2, L, 0)
    assert dom == IF and len(parts[POS][1]) == 0 and abs(parts[NEG][1].sum() - 1 / 6) < 1e-15
    dom, parts = orc.subdiv_rule3(np.r_[np.ones(9), -1e-17], 2, L, 0)
    assert dom == IF and len(parts[NEG][1]) == 0 and abs(parts[POS][1].sum() - 1 / 6) < 1e-15
