"""Node and dof marking on the GPU (csrc/xg_nodes.hip) vs the CPU oracle
(oracle/nodes.py, which restates xfem/cutinfo.cpp:135-224, 540-647): bit
for bit, through the C ABI and through the reference-shaped Python surface
(CutInfo node arrays, GetDofsOfElements, GetDofsOfFacets,
GetElementsWithNeighborFacets, GetElementsWithSharedVertex), as used by
ngsxfem's demos (unf_interf_prob.py:138, fictdom_mlset.py:108)."""
import numpy as np
import pytest

from ngsxfem_amd import workloads as W
from oracle import nodes as N

pytestmark = pytest.mark.gpu

NEG, POS, IF = 0, 1, 2


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from ngsxfem_amd import _abi
    ctx = _abi.Context(0)
    yield torch, _abi, ctx
    ctx.close()


@pytest.fixture(scope="module")
def X():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import xfem
    return xfem


def up(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def bits_of(torch, b):
    return up(torch, N.bool_to_bits(b).view(np.int32))


def got_bits(t, n):
    return N.bits_to_bool(t.cpu().numpy().view(np.uint32), n)


def test_mark_from_items_c2(dev, orc):
    """GetDofsOfElements on the 64^3 Kuhn mesh (1.57M tets), P1 and P2, HASNEG
    and IF marks, plus a random mark and the empty mark"""
    torch, _abi, ctx = dev
    mesh, lset = W.config2(64)
    dt, _, _ = orc.mark(3, mesh.el2v, lset)
    rng = np.random.default_rng(5)
    for p in (1, 2):
        dofs, ndof = mesh.dofs(p)
        d_dofs = up(torch, dofs)
        for mark in (dt != POS, dt == IF, rng.random(mesh.ne) < 0.01, np.zeros(mesh.ne, bool)):
            out = torch.full(((ndof + 31) // 32,), -1, dtype=torch.int32, device="cuda")
            ctx.mark_from_items(d_dofs, bits_of(torch, mark), ndof, out)
            ctx.sync()
            assert np.array_equal(got_bits(out, ndof), N.dofs_of_items(dofs, mark, ndof))


def test_mark_from_items_rejects_out_of_range(dev):
    torch, _abi, ctx = dev
    tab = up(torch, np.array([[0, 1], [2, 9]], np.int32))
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.mark_from_items(tab, None, 5, out)
    with pytest.raises(_abi.XgError, match="out of range"):
        ctx.sync()
    # the in-range bits are still set, the error word is cleared again
    assert np.array_equal(got_bits(out, 5), [True, True, True, False, False])
    ctx.mark_from_items(tab[:1].contiguous(), None, 5, out)
    ctx.sync()


def test_neighbourhoods_c2(dev, orc):
    torch, _abi, ctx = dev
    mesh, lset = W.config2(32)
    dt, _, _ = orc.mark(3, mesh.el2v, lset)
    f2v, el2f, f2el = mesh.facets()
    rng = np.random.default_rng(2)
    fm = rng.random(len(f2el)) < 0.02
    out = torch.zeros((mesh.ne + 31) // 32, dtype=torch.int32, device="cuda")
    ctx.elements_with_neighbor_facets(up(torch, f2el), bits_of(torch, fm), mesh.ne, out)
    ctx.sync()
    assert np.array_equal(got_bits(out, mesh.ne), N.elements_with_neighbor_facets(f2el, fm, mesh.ne))
    for em in (dt == IF, rng.random(mesh.ne) < 0.001):
        out.fill_(-1)
        ctx.elements_with_shared_vertex(3, up(torch, mesh.el2v), bits_of(torch, em), mesh.nv, out)
        ctx.sync()
        ref = N.elements_with_shared_vertex(mesh.el2v, em, mesh.nv)
        assert np.array_equal(got_bits(out, mesh.ne), ref)
        assert ref.sum() > em.sum()


@pytest.mark.parametrize("D", [1, 2, 3])
def test_node_domains_abi(dev, orc, D):
    torch, _abi, ctx = dev
    from ngsxfem_amd.meshes import Make1DMesh
    if D == 1:
        mesh = Make1DMesh(200)
        lset = np.abs(mesh.coords[:, 0] - 0.5) - 0.23
    elif D == 2:
        mesh, lset = W.config1()
    else:
        mesh, lset = W.config2(16)
    dt, cr, _ = orc.mark(D, mesh.el2v, lset)
    d_dt, d_cr = up(torch, dt), up(torch, cr)
    tables = [(mesh.el2v, mesh.nv)]
    if D >= 2:
        e2v, el2e = mesh.edges()
        tables.append((el2e, len(e2v)))
    if D == 3:
        tables.append((mesh.facets()[1], mesh.GetNFacets()))
    for tab, nn in tables:
        prev = np.zeros(nn, bool)
        prev[::97] = True  # flags of an earlier update are kept
        cutnb = bits_of(torch, prev)
        dom = torch.full((nn,), 7, dtype=torch.uint8, device="cuda")
        ctx.node_domains(up(torch, tab), d_dt, None, nn, cutnb, dom)
        ctx.sync()
        rc, rd = N.node_domains(tab, dt, None, nn, cutnb_prev=prev)
        assert np.array_equal(got_bits(cutnb, nn), rc)
        assert np.array_equal(dom.cpu().numpy(), rd)
    cutnb = torch.zeros((mesh.ne + 31) // 32, dtype=torch.int32, device="cuda")
    dom = torch.full((mesh.ne,), 7, dtype=torch.uint8, device="cuda")
    ctx.node_domains(None, d_dt, d_cr, mesh.ne, cutnb, dom)
    ctx.sync()
    rc, rd = N.node_domains(None, dt, cr, mesh.ne)
    assert np.array_equal(got_bits(cutnb, mesh.ne), rc)
    assert np.array_equal(dom.cpu().numpy(), rd)


def test_python_surface(X, orc):
    """the reference-shaped functions on a 2D and a 3D mesh"""
    for mesh, lset in (W.config1(), W.config2(8)):
        D = mesh.dim
        gf = X.GridFunction(X.H1(mesh, order=1))
        gf.vec[:] = lset
        ci = X.CutInfo(mesh, gf)
        dt, cr, _ = orc.mark(D, mesh.el2v, lset)
        hasneg = ci.GetElementsOfType(X.HASNEG)
        for p in (1, 2):
            V = X.H1(mesh, order=p)
            dofs, ndof = V.dofs()
            got = X.GetDofsOfElements(V, hasneg)
            assert len(got) == ndof
            assert np.array_equal(got.numpy(), N.dofs_of_items(dofs, dt != POS, ndof))
            fa = X.GetFacetsWithNeighborTypes(mesh, a=ci.GetElementsOfType(X.IF), b=hasneg)
            fd = mesh.facet_dofs(p)
            assert np.array_equal(X.GetDofsOfFacets(V, fa).numpy(), N.dofs_of_items(fd, fa.numpy(), ndof))
        f2el = mesh.facets()[2]
        fa = X.GetFacetsWithNeighborTypes(mesh, a=ci.GetElementsOfType(X.IF), b=hasneg)
        els = X.GetElementsWithNeighborFacets(mesh, fa)
        assert np.array_equal(els.numpy(), N.elements_with_neighbor_facets(f2el, fa.numpy(), mesh.ne))
        sv = X.GetElementsWithSharedVertex(mesh, ci.GetElementsOfType(X.IF))
        assert np.array_equal(sv.numpy(), N.elements_with_shared_vertex(mesh.el2v, dt == IF, mesh.nv))
        # node arrays of CutInformation::Update
        nts = [(X.NT_VERTEX, mesh.el2v, mesh.nv)]
        e2v, el2e = mesh.edges()
        nts.append((X.NT_EDGE, el2e, len(e2v)))
        if D == 3:
            nts.append((X.NT_FACE, mesh.facets()[1], mesh.GetNFacets()))
        for nt, tab, nn in nts:
            rc, rd = N.node_domains(tab, dt, None, nn)
            assert np.array_equal(ci.GetCutNeighboringNodes(nt).numpy(), rc)
            assert np.array_equal(ci.GetDomainOfNodes(nt), rd)
        rc, rd = N.node_domains(None, dt, cr, mesh.ne)
        assert np.array_equal(ci.GetCutNeighboringNodes(X.NT_ELEMENT).numpy(), rc)
        assert np.array_equal(ci.GetDomainOfNodes(X.NT_ELEMENT), rd)
        # a second Update with a shifted level set: cut-neighbour flags
        # accumulate, domains follow the latest level set
        gf2 = X.GridFunction(X.H1(mesh, order=1))
        gf2.vec[:] = lset + 0.05
        ci.Update(gf2)
        dt2, _, _ = orc.mark(D, mesh.el2v, lset + 0.05)
        rc2, rd2 = N.node_domains(mesh.el2v, dt2, None, mesh.nv, cutnb_prev=N.node_domains(mesh.el2v, dt, None,
                                                                                           mesh.nv)[0])
        assert np.array_equal(ci.GetCutNeighboringNodes(X.NT_VERTEX).numpy(), rc2)
        assert np.array_equal(ci.GetDomainOfNodes(X.NT_VERTEX), rd2)
