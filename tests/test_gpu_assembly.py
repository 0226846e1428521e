"""Linear forms and global assembly on the GPU (the steps after the element
matrices in every user flow: SymbolicCutLinearFormIntegrator::
T_CalcElementVector, xfem/symboliccutlfi.cpp:57-140, and the scatter of
BilinearForm::Assemble / LinearForm::Assemble).

Pins: the element vector of f = 1 is the row sum of the element mass matrix
(partition of unity), checked against the FP64 oracle's mass matrices; the
device scatter equals a host COO assembly of the same element matrices; the
assembled P1 and P2 matrices reproduce int_NEG 1 = 1^T M 1 = sum F and
K 1 = 0, and a P2 mass matrix integrates P2 functions exactly
(c_p^T M c_q = int p q for p, q in P2, by the interpolation property)."""
import math

import numpy as np
import pytest

from ngsxfem_amd import workloads as W

pytestmark = pytest.mark.gpu
NEG, POS, IF = 0, 1, 2


@pytest.fixture(scope="module")
def dev_abi():
    import torch
    assert torch.cuda.is_available()
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


@pytest.mark.parametrize("D,p", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_element_vector_is_mass_row_sum(orc, X, D, p):
    mesh, lset = W.config1() if D == 2 else W.config2(8)
    dto, _, _ = orc.mark(D, mesh.el2v, lset)
    els = np.nonzero(dto != POS)[0][:400].astype(np.int32)
    for dt in (NEG, IF):
        e, F = X.CutElementVectors(mesh, lset, dt, 2 * p, p, els)
        ref = orc.cut_pipeline(D, p, 2 * p, els, mesh.el2v, mesh.coords, lset, [dt], mask=1, nthreads=8)[:, 0]
        rs = ref.sum(axis=2)
        scale = np.abs(rs).max()
        assert np.abs(F - rs).max() <= 1e-12 * scale


@pytest.mark.parametrize("p", [1, 2])
def test_assembly_matches_host_scatter(X, p):
    from scipy.sparse import coo_matrix
    mesh = X.MakeStructured3DMesh(nx=6, ny=6, nz=6)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.sqrt((X.x - 0.5) ** 2 + (X.y - 0.5) ** 2 + (X.z - 0.5) ** 2) - 0.33, lset)
    ci = X.CutInfo(mesh, lset)
    V = X.H1(mesh, order=p)
    u, v = V.TnT()
    a = X.BilinearForm(V)
    a += (X.grad(u) * X.grad(v) + u * v) * X.dCut(lset, X.NEG, definedonelements=ci.GetElementsOfType(X.HASNEG))
    a.Assemble(precision=0)
    dofs, n = V.dofs()
    els, mats = a.elmats[0]
    nd = dofs.shape[1]
    d = dofs[els]
    host = coo_matrix((np.asarray(mats).ravel(), (np.repeat(d, nd, axis=1).ravel(), np.tile(d, (1, nd)).ravel())),
                      shape=(n, n)).tocsr()
    diff = abs(a.mat - host).max()
    assert diff <= 1e-13 * abs(host).max()
    # K 1 = 0 and 1^T M 1 = |NEG|
    m = X.BilinearForm(V)
    m += (u * v) * X.dCut(lset, X.NEG, definedonelements=ci.GetElementsOfType(X.HASNEG))
    m.Assemble(precision=0)
    one = np.ones(n)
    vol = X.Integrate({"levelset": lset, "domain_type": X.NEG}, cf=1, mesh=mesh, order=2)
    assert abs(one @ (m.mat @ one) - vol) < 1e-12 * vol
    k = a.mat - m.mat
    assert np.abs(k @ one).max() < 1e-11 * abs(k).max()
    # linear form: sum F = |NEG|, F = M 1
    f = X.LinearForm(V)
    f += 1.0 * v * X.dCut(lset, X.NEG, definedonelements=ci.GetElementsOfType(X.HASNEG))
    f.Assemble()
    assert abs(f.vec.sum() - vol) < 1e-12 * vol
    assert np.abs(f.vec - m.mat @ one).max() < 1e-13 * np.abs(f.vec).max()


def test_p2_mass_integrates_quadratics(X):
    """uncut P2 mass matrix on the unit cube (level set -1 everywhere):
    c_p^T M c_q = int p q for p = x^2, q = y (nodal interpolation is exact)"""
    mesh = X.MakeStructured3DMesh(nx=3, ny=3, nz=3)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.CoefficientFunction(-1.0), lset)
    V = X.H1(mesh, order=2)
    u, v = V.TnT()
    m = X.BilinearForm(V)
    m += (u * v) * X.dCut(lset, X.NEG)
    m.Assemble(precision=0)
    dofs, n = V.dofs()
    # nodal coordinates of the P2 dofs: vertices, then edge midpoints
    xyz = np.zeros((n, 3))
    xyz[:mesh.nv] = mesh.coords
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for k, (i, j) in enumerate(pairs):
        xyz[dofs[:, 4 + k]] = 0.5 * (mesh.coords[mesh.el2v[:, i]] + mesh.coords[mesh.el2v[:, j]])
    cp = xyz[:, 0] ** 2
    cq = xyz[:, 1]
    assert abs(cp @ (m.mat @ cq) - (1.0 / 3.0) * 0.5) < 1e-13
    assert abs(np.ones(n) @ (m.mat @ np.ones(n)) - 1.0) < 1e-13


def test_csr_pattern_device_matches_numpy(dev_abi):
    """xg_csr_pattern (radix sort + unique of the couplings) equals the sorted
    unique pattern from numpy, skipped (negative) dofs left out; P1 and P2
    dof tables of a 5^3 Kuhn mesh, and a dof >= n is reported"""
    torch, _abi, ctx = dev_abi
    from ngsxfem_amd.meshes import MakeStructured3DMesh
    mesh = MakeStructured3DMesh(nx=5, ny=5, nz=5)
    for p in (1, 2):
        dofs, n = mesh.dofs(p)
        dofs = dofs.astype(np.int32).copy()
        dofs[::7, 0] = -1  # skipped dofs
        nd = dofs.shape[1]
        rows = np.repeat(dofs, nd, axis=1).ravel().astype(np.int64)
        cols = np.tile(dofs, (1, nd)).ravel().astype(np.int64)
        ok = (rows >= 0) & (cols >= 0)
        key = np.unique(rows[ok] * n + cols[ok])
        ref_rp = np.zeros(n + 1, np.int64)
        np.add.at(ref_rp, key // n + 1, 1)
        ref_rp = np.cumsum(ref_rp)
        d = torch.from_numpy(dofs).cuda()
        rp = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
        ci = torch.full((dofs.shape[0] * nd * nd,), -1, dtype=torch.int32, device="cuda")
        nnz = torch.zeros(1, dtype=torch.int32, device="cuda")
        ctx.csr_pattern(d, n, rp, ci, nnz)
        ctx.sync()
        m = int(nnz.item())
        assert m == len(key)
        assert np.array_equal(rp.cpu().numpy(), ref_rp)
        assert np.array_equal(ci.cpu().numpy()[:m], (key % n).astype(np.int32))
    bad = torch.from_numpy(np.array([[0, 1, 2, 99]], np.int32)).cuda()
    ctx.csr_pattern(bad, 10, rp[:11], ci, nnz)
    with pytest.raises(RuntimeError):
        ctx.sync()


@pytest.mark.parametrize("p", [1, 2])
def test_restricted_bilinear_form(X, p):
    """RestrictedBilinearForm (python/__init__.py:973-1037, utils/
    restrictedblf.cpp:8-64,145-164): the matrix graph holds exactly the dof
    pairs of the active elements (checked against a numpy set of pairs),
    rows of dofs outside them are empty, and the values equal the full
    BilinearForm's on the same integrals; an integral on an element outside
    the graph is an error"""
    mesh = X.MakeStructured3DMesh(nx=6, ny=6, nz=6)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.sqrt((X.x - 0.5) ** 2 + (X.y - 0.5) ** 2 + (X.z - 0.5) ** 2) - 0.3, lset)
    ci = X.CutInfo(mesh, lset)
    act = ci.GetElementsOfType(X.HASNEG)
    V = X.H1(mesh, order=p)
    u, v = V.TnT()
    form = (X.grad(u) * X.grad(v) + u * v) * X.dCut(lset, X.NEG, definedonelements=act)
    a = X.RestrictedBilinearForm(V, element_restriction=act)
    a += form
    a.Assemble(precision=0)
    b = X.BilinearForm(V)
    b += form
    b.Assemble(precision=0)
    dofs, n = V.dofs()
    ae = np.nonzero(np.asarray(act.numpy(), dtype=bool))[0]
    d = dofs[ae]
    pairs = np.unique(np.repeat(d, d.shape[1], axis=1).ravel().astype(np.int64) * n +
                      np.tile(d, (1, d.shape[1])).ravel())
    A = a.mat.tocoo()
    got = np.unique(A.row.astype(np.int64) * n + A.col)
    assert np.array_equal(got, pairs)
    assert a.mat.nnz < b.mat.nnz
    inactive = np.setdiff1d(np.arange(n), np.unique(d))
    assert len(inactive) > 0 and np.all(np.diff(a.mat.indptr)[inactive] == 0)
    assert abs(a.mat - b.mat).max() < 1e-14 * abs(b.mat).max()
    c = X.RestrictedBilinearForm(V, element_restriction=act)
    c += (u * v) * X.dCut(lset, X.POS)  # POS elements lie outside the graph
    with pytest.raises(RuntimeError):
        c.Assemble(precision=0)
