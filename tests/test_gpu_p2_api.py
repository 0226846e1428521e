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
    dom = np.array([orc.subdiv_rule3(v, 2, L, 0)[0] for v in nod])
    for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
        assert np.array_equal(ci.GetElementsOfType(cdt).numpy(), dom == dt)
    assert (dom == IF).sum() > 20 and (dom == NEG).sum() > 0
    cr = ci.GetCutRatios()
    for e in np.nonzero(dom == IF)[0]:
        _, parts = orc.subdiv_rule3(nod[e], 2, L, 0)
        vn, vp = parts[0][1].sum(), parts[1][1].sum()
        assert abs(cr[e] - vn / (vn + vp)) < 1e-12
    assert np.all(cr[dom == NEG] == 1.0) and np.all(cr[dom == POS] == 0.0)
    with pytest.raises(Exception):
        ci.GetElementsOfType(X.IF, X.BND)


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
