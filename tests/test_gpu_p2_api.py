"""P2 level sets through the reference-shaped Python surface (CutInfo with a
P2 grid function / subdivlvl, Integrate / IntegrateX with subdivlvl, dCut
bilinear forms with subdivlvl) vs the FP64 oracle's subdivision rule
(or_subdiv_rule3, pinned in the oracle tests): element domains bit for bit,
cut ratios, polynomial integrals (degree <= 4, exact on the fixed-node
contraction) and element matrices to rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NEG, POS, IF = 0, 1, 2


@pytest.fixture(scope="module")
def X():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import xfem
    return xfem


def sphere_p2(X, n=5, r=0.31, c=(0.52, 0.47, 0.49)):
    mesh = X.MakeStructured3DMesh(nx=n, ny=n, nz=n)
    g = X.GridFunction(X.H1(mesh, order=2))
    g.Set(lambda x, y, z: (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 - r * r)
    return mesh, g


def nodal(mesh, g):
    dofs, _ = mesh.dofs(2)
    return np.ascontiguousarray(g.vec[dofs])


@pytest.mark.parametrize("L", [0, 1, 3])
def test_cutinfo_p2_domains_and_ratios(X, orc, L):
    mesh, g = sphere_p2(X)
    ci = X.CutInfo(mesh, g, subdivlvl=L)
    nod = nodal(mesh, g)
    dom = np.array([part_volume_domain(orc, v, L) for v in nod])
    for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
        assert np.array_equal(ci.GetElementsOfType(cdt).numpy(), dom == dt)
    assert (dom == IF).sum() > 20 and (dom == NEG).sum() > 0
    cr = ci.GetCutRatios()
    for e in np.nonzero(dom == IF)[0]:
        _, parts = orc.subdiv_rule3(nod[e], 2, L, 0)
        vn, vp = parts[0][1].sum(), parts[1][1].sum()
        assert abs(cr[e] - vn / (vn + vp)) < 1e-12
    assert np.all(cr[dom == NEG] == 1.0) and np.all(cr[dom == POS] == 0.0)
    # the sphere stays inside the cube: every boundary triangle is POS
    assert ci.GetElementsOfType(X.POS, X.BND).numpy().all()


@pytest.mark.parametrize("dt", [NEG, POS, IF])
def test_integrate_p2_polynomials(X, orc, dt):
    mesh, g = sphere_p2(X)
    nod = nodal(mesh, g)
    cf = X.x * X.y + 2 * X.z * X.z * X.x - 0.5 * X.y ** 4 + 1
    f = lambda p: p[:, 0] * p[:, 1] + 2 * p[:, 2] ** 2 * p[:, 0] - 0.5 * p[:, 1] ** 4 + 1  # noqa: E731
    L, order = 2, 4
    got = X.Integrate({"levelset": g, "domain_type": [X.NEG, X.POS, X.IF][dt], "subdivlvl": L}, cf=cf, mesh=mesh,
                      order=order, element_wise=True)
    for e in range(mesh.ne):
        V = mesh.coords[mesh.el2v[e]]
        F, Fi, ad = orc.jacobian(V)
        _, parts = orc.subdiv_rule3(nod[e], 2, L, order, Finv=Fi)
        pts, w = parts[dt]
        ref = 0.0
        if len(w):
            x = V[3][None, :] + pts @ F.T
            ref = float(np.sum(f(x) * w)) * ad
        assert abs(got[e] - ref) <= 1e-12 * max(1.0, abs(ref)), (e, got[e], ref)
    assert np.abs(got).max() > 0


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("dt", [NEG, IF])
def test_dcut_p2_element_matrices(X, orc, p, dt):
    mesh, g = sphere_p2(X, n=4)
    nod = nodal(mesh, g)
    ci = X.CutInfo(mesh, g, subdivlvl=2)
    els = ci.GetElementsOfType(X.IF)
    V_ = X.H1(mesh, order=p)
    u, v = V_.TnT()
    a = X.BilinearForm(V_)
    a += (X.grad(u) * X.grad(v) + u * v) * X.dCut(g, [X.NEG, X.POS, X.IF][dt], subdivlvl=2, order=2 * p,
                                                   definedonelements=els)
    a.Assemble(precision=X.PREC_FP64)
    e_ids, mats = a.elmats[0]
    assert len(e_ids) == els.NumSet() > 0
    for k, e in enumerate(e_ids):
        V = mesh.coords[mesh.el2v[e]]
        F, Fi, ad = orc.jacobian(V)
        _, parts = orc.subdiv_rule3(nod[e], 2, 2, 2 * p, Finv=Fi)
        pts, w = parts[dt]
        ref = orc.elmat(3, p, V, np.ascontiguousarray(pts), np.ascontiguousarray(w), mask=3)
        assert np.abs(mats[k] - ref).max() <= 1e-12 * np.abs(ref).max(), e


def part_volume_domain(orc, nod, L):
    """cutinfo.cpp:104-124 on the oracle's order-0 NEG and POS rules"""
    _, parts = orc.subdiv_rule3(nod, 2, L, 0)
    vn, vp = parts[NEG][1].sum(), parts[POS][1].sum()
    return (IF if vp > 0.0 else NEG) if vn > 0.0 else POS


@pytest.mark.parametrize("L", [1, 2, 3])
def test_cutinfo_p2_lattice_cut_but_one_sided(X, orc, L):
    """Elements the lattice check calls cut, whose leaves all fall to one side
    (exact zeros on a negative level set; tiny negative values on a positive
    one, below the leaves' relative 1e-14 threshold): the reference marks
    them by their part volumes, so they are uncut"""
    mesh = X.MakeStructured3DMesh(nx=3, ny=3, nz=3)
    dofs, _ = mesh.dofs(2)
    for sign, odd in ((-1.0, 0.0), (1.0, -1e-17)):
        g = X.GridFunction(X.H1(mesh, order=2))
        g.vec[:] = sign
        g.vec[dofs[5, 4:7]] = odd
        g.vec[dofs[40, 9]] = odd
        nod = np.ascontiguousarray(g.vec[dofs])
        lattice = np.array([orc.subdiv_rule3(v, 2, L, 0)[0] for v in nod])
        want = np.array([part_volume_domain(orc, v, L) for v in nod])
        assert (lattice == IF).sum() > 0 and not np.any(want == IF)
        ci = X.CutInfo(mesh, g, subdivlvl=L)
        for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
            assert np.array_equal(ci.GetElementsOfType(cdt).numpy(), want == dt), (sign, dt)
        assert np.array_equal(ci.GetElementsOfType(X.HASNEG).numpy(), want != POS)
        cr = ci.GetCutRatios()
        assert np.all(cr[want == NEG] == 1.0) and np.all(cr[want == POS] == 0.0)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("L", [0, 2, 3])
def test_cutinfo_p2_boundary_elements(X, orc, L, order):
    """Boundary triangles (cutinfo.cpp:80-124 with vb = BND) of a mesh cut by
    a sphere centred on a face: domains from the part areas of the oracle's
    triangle subdivision rule on each boundary triangle's restriction, bit
    for bit, cut ratios to 1e-12 (P2 level set, or P1 with subdivlvl)"""
    mesh = X.MakeStructured3DMesh(nx=4, ny=4, nz=4)
    g = X.GridFunction(X.H1(mesh, order=order))
    g.Set(lambda x, y, z: (x - 0.45) ** 2 + (y - 0.52) ** 2 + z * z - 0.4 ** 2)
    ci = X.CutInfo(mesh, g, subdivlvl=L)
    if order == 2:
        tab, _ = mesh.bnd_dofs(2)
        nod = g.vec[tab]
    else:
        b = mesh.bnd2v
        v = g.vec[b]
        nod = np.hstack([v, 0.5 * (v[:, [0, 0, 1]] + v[:, [1, 2, 2]])])
    want, wcr = [], []
    for c in nod:
        _, an, ap = orc.subdiv_trig_parts(np.ascontiguousarray(c), 2, L)
        want.append((IF if ap > 0 else NEG) if an > 0 else POS)
        wcr.append(an / (an + ap))
    want, wcr = np.array(want), np.array(wcr)
    assert (want == IF).sum() > 5 and (want == NEG).sum() > 0
    for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
        assert np.array_equal(ci.GetElementsOfType(cdt, X.BND).numpy(), want == dt), dt
    assert np.abs(np.asarray(ci.GetCutRatios(X.BND)) - wcr).max() <= 1e-12


@pytest.mark.parametrize("L", [2, 3])
def test_cutinfo_p2_zero_planes_count_as_pos(X, orc, L):
    """A P2 level set that vanishes on whole lattice planes, x * (y - h) with
    h = 2^-L scaled to a cell of the mesh: subtrees whose lattice values are
    all exactly zero belong to POS (xintegration.cpp:1159-1164, the lattice
    check's v >= 0), so NEG + POS part volumes make up every element and the
    cut ratios and domains are the oracle's (NEG / POS rules of the tree)"""
    mesh = X.MakeStructured3DMesh(nx=2, ny=2, nz=2)
    h = 0.5 * 2.0 ** -L
    g = X.GridFunction(X.H1(mesh, order=2))
    g.Set(lambda x, y, z: (x - 0.5) * (y - h))
    dofs, _ = mesh.dofs(2)
    nod = np.ascontiguousarray(g.vec[dofs])
    ci = X.CutInfo(mesh, g, subdivlvl=L)
    cr = np.asarray(ci.GetCutRatios())
    for e in range(mesh.ne):
        _, parts = orc.subdiv_rule3(nod[e], 2, L, 0)
        vn, vp = parts[NEG][1].sum(), parts[POS][1].sum()
        assert abs(vn + vp - 1.0 / 6.0) < 1e-14
        want = (IF if vp > 0.0 else NEG) if vn > 0.0 else POS
        for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
            assert ci.GetElementsOfType(cdt).numpy()[e] == (want == dt), (e, dt)
        assert abs(cr[e] - vn / (vn + vp)) <= 1e-12, e
    assert (np.asarray(ci.GetElementsOfType(X.IF).numpy())).sum() > 0


@pytest.mark.parametrize("p", [1, 2, 3])
def test_dcut_p2_pos_keeps_zero_planes(X, orc, p):
    """dCut(lset_p2, POS, subdivlvl=L) on the same zero planes: the POS
    element matrices are the oracle's POS rule's, and NEG + POS = the uncut
    matrix of every element"""
    L = 2
    mesh = X.MakeStructured3DMesh(nx=2, ny=2, nz=2)
    g = X.GridFunction(X.H1(mesh, order=2))
    g.Set(lambda x, y, z: (x - 0.5) * (y - 0.125))
    dofs, _ = mesh.dofs(2)
    nod = np.ascontiguousarray(g.vec[dofs])
    V_ = X.H1(mesh, order=p)
    u, v = V_.TnT()
    mats = {}
    for dt in (X.NEG, X.POS):
        a = X.BilinearForm(V_)
        a += (X.grad(u) * X.grad(v) + u * v) * X.dCut(g, dt, subdivlvl=L, order=2 * p)
        a.Assemble(precision=X.PREC_FP64)
        e_ids, m = a.elmats[0]
        mats[int(dt)] = np.asarray(m)
    for e in range(mesh.ne):
        Vx = mesh.coords[mesh.el2v[e]]
        F, Fi, ad = orc.jacobian(Vx)
        _, parts = orc.subdiv_rule3(nod[e], 2, L, 2 * p, Finv=Fi)
        pts, w = parts[POS]
        ref = orc.elmat(3, p, Vx, np.ascontiguousarray(pts), np.ascontiguousarray(w), mask=3)
        assert np.abs(mats[POS][e] - ref).max() <= 1e-12 * np.abs(ref).max(), e
        full = orc.elmat(3, p, Vx, orc.rule(3, 2 * p)[0], orc.rule(3, 2 * p)[1], mask=3)
        assert np.abs(mats[NEG][e] + mats[POS][e] - full).max() <= 1e-12 * np.abs(full).max(), e
