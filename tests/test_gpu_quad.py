"""Quadrilateral / hexahedral straight cut rules on the GPU (csrc/xg_quad.hip)
vs the CPU oracle (or_quad_cut_rule / or_quad_mark, pinned in
test_quad_oracle_cpu.py): point counts exact, points and weights to 1e-14
for every direction policy; marks (domain types, the 8 bit arrays, cut
list) bit for bit, cut ratios to 1e-14; and the reference's known answers
(tests/pytests/test_straightcutrule.py, test_quadchallenges.py) through the
Python surface (Integrate, CutInfo, GetElementsWithSharedVertex)."""
import math

import numpy as np
import pytest

from ngsxfem_amd.meshes import MakeStructured2DMesh, MakeStructured3DMesh
from oracle import nodes as N

pytestmark = pytest.mark.gpu

NEG, POS, IF = 0, 1, 2
FIRST, OPTIMAL, FALLBACK = 0, 1, 2


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


def warped(D, n, amp=0.03):
    """quad / hex mesh with smoothly moved interior vertices (non-affine
    elements, so the interface weights use F at every point)"""
    if D == 2:
        f = lambda x, y: (x + amp * math.sin(math.pi * x) * math.sin(2 * math.pi * y),  # noqa: E731
                          y + amp * math.sin(math.pi * y) * math.sin(2 * math.pi * x + 0.4))
        return MakeStructured2DMesh(quads=True, nx=n, ny=n, mapping=f)
    f = lambda x, y, z: (x + amp * math.sin(math.pi * x) * math.sin(2 * math.pi * y + 0.2),  # noqa: E731
                         y + amp * math.sin(math.pi * y) * math.sin(2 * math.pi * z + 0.5),
                         z + amp * math.sin(math.pi * z) * math.sin(2 * math.pi * x + 0.9))
    return MakeStructured3DMesh(hexes=True, nx=n, ny=n, nz=n, mapping=f)


def level_sets(mesh, kind, seed=0):
    X = mesh.coords
    if kind == "random":  # no exact zeros
        return np.random.default_rng(seed).uniform(-1, 1, mesh.nv)
    if kind == "sphere":
        c = np.full(X.shape[1], 0.47)
        return np.sqrt(((X - c) ** 2).sum(1)) - 0.33
    if kind == "plane":
        g = np.array([0.7, -0.4, 0.59][:X.shape[1]])
        return X @ g - 0.21
    raise ValueError(kind)


def gpu_rule(dev, mesh, lset, dt, order, pol, elems=None):
    torch, _abi, ctx = dev
    D = mesh.dim
    el, xs, ls = up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, lset)
    ev = None if elems is None else up(torch, np.asarray(elems, dtype=np.int32))
    nel = mesh.ne if elems is None else len(elems)
    off = torch.empty(nel + 1, dtype=torch.int32, device="cuda")
    ctx.quad_rule_count(D, el, ls, dt, order, pol, off, elems=ev)
    ctx.sync()
    tot = int(off[-1].item())
    pts = torch.empty(max(tot, 1) * 4, dtype=torch.float64, device="cuda")
    ctx.quad_rule_fill(D, el, xs, ls, dt, order, pol, off, pts, elems=ev)
    ctx.sync()
    return off.cpu().numpy(), pts[:4 * tot].view(-1, 4).cpu().numpy()


def oracle_rules(orc, mesh, lset, dt, order, pol):
    """{element: rule or None} and the elements where the oracle (like the
    reference) raises"""
    rules, bad = {}, []
    for e in range(mesh.ne):
        vs = mesh.el2v[e]
        try:
            rules[e] = orc.quad_cut_rule(mesh.dim, lset[vs], dt, order, pol, V=mesh.coords[vs])
        except RuntimeError:
            bad.append(e)
    return rules, bad


@pytest.mark.parametrize("D,n", [(2, 9), (3, 4)])
@pytest.mark.parametrize("kind", ["random", "sphere", "plane"])
@pytest.mark.parametrize("pol", [FIRST, OPTIMAL, FALLBACK])
def test_rule_parity(dev, orc, D, n, kind, pol):
    """every element the oracle accepts, point by point; the elements where
    the reference throws ("if_scale_factor larger than bound" for strongly
    curved random level sets) make the GPU call fail as well"""
    torch, _abi, ctx = dev
    mesh = warped(D, n)
    lset = level_sets(mesh, kind, seed=D * 10 + pol)
    nbad = 0
    for dt in (NEG, POS, IF):
        for order in (0, 3, 6):
            rules, bad = oracle_rules(orc, mesh, lset, dt, order, pol)
            ok = sorted(rules)
            off, pts = gpu_rule(dev, mesh, lset, dt, order, pol, elems=ok)
            for i, e in enumerate(ok):
                r = rules[e]
                got = pts[off[i]:off[i + 1]]
                if r is None:
                    assert len(got) == 0, (e, dt, order)
                    continue
                assert len(got) == len(r[1]), (e, dt, order, len(got), len(r[1]))
                scale = max(1.0, np.abs(r[1]).max())
                assert np.abs(got[:, :3] - r[0]).max() <= 1e-14, (e, dt, order)
                assert np.abs(got[:, 3] - r[1]).max() <= 1e-14 * scale, (e, dt, order)
            for e in bad[:2]:
                nbad += 1
                with pytest.raises(_abi.XgError):
                    gpu_rule(dev, mesh, lset, dt, order, pol, elems=[e])
    if kind != "random" or pol == FALLBACK:
        assert nbad == 0


def test_rule_parity_tpmc_all_policies(dev, orc):
    """the single-hex TPMC level set (topology change inside the element)"""
    mesh = MakeStructured3DMesh(hexes=True, nx=1, ny=1, nz=1)
    lset = np.array([-4, 4, -1, -1, 2, -3, 5, -1], dtype=np.float64)
    for pol in (FIRST, OPTIMAL, FALLBACK):
        for dt in (NEG, POS, IF):
            off, pts = gpu_rule(dev, mesh, lset, dt, 4, pol)
            r = orc.quad_cut_rule(3, lset[mesh.el2v[0]], dt, 4, pol, V=mesh.coords[mesh.el2v[0]])
            assert len(pts) == len(r[1])
            assert np.abs(pts[:, :3] - r[0]).max() <= 1e-14
            assert np.abs(pts[:, 3] - r[1]).max() <= 1e-14 * max(1.0, np.abs(r[1]).max())


@pytest.mark.parametrize("D,n", [(2, 40), (3, 12)])
@pytest.mark.parametrize("kind", ["random", "sphere"])
@pytest.mark.parametrize("pol", [FIRST, OPTIMAL])
def test_mark_parity(dev, orc, D, n, kind, pol):
    torch, _abi, ctx = dev
    mesh = warped(D, n)
    lset = level_sets(mesh, kind, seed=7 + D)
    ne = mesh.ne
    nw = (ne + 31) // 32
    dt = torch.empty(ne, dtype=torch.uint8, device="cuda")
    cr = torch.empty(ne, dtype=torch.float64, device="cuda")
    bits = torch.empty(8 * nw, dtype=torch.int32, device="cuda")
    cut = torch.empty(ne, dtype=torch.int32, device="cuda")
    nc = torch.empty(1, dtype=torch.int32, device="cuda")
    ctx.quad_mark(D, up(torch, mesh.el2v), up(torch, lset), pol, dt, bits, cr, cut, nc)
    ctx.sync()
    rdt, rcr, rbits = orc.quad_mark(D, mesh.el2v, lset, pol)
    assert np.array_equal(dt.cpu().numpy(), rdt)
    assert np.array_equal(bits.cpu().numpy().view(np.uint32).reshape(8, nw), rbits)
    assert np.abs(cr.cpu().numpy() - rcr).max() <= 1e-14
    n_cut = int(nc.item())
    assert np.array_equal(cut[:n_cut].cpu().numpy(), np.nonzero(rdt == IF)[0])
    assert (rdt == IF).sum() > 0 and (rdt == NEG).sum() > 0 and (rdt == POS).sum() > 0


def p1(X, mesh, f):
    g = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(f, g)
    return g


# ---- tests/pytests/test_straightcutrule.py:106-126, 187-207, 209-240
@pytest.mark.parametrize("N_", [1, 10, 30])
@pytest.mark.parametrize("order", [2, 4, 8])
def test_integrate_straight_cut_quad2d(X, N_, order):
    mesh = MakeStructured2DMesh(quads=True, nx=N_, ny=N_)
    g = p1(X, mesh, 1 - 2 * X.x - 2 * X.y)
    for dt, ref in ((X.NEG, 7 / 8), (X.POS, 1 / 8), (X.IF, 1 / math.sqrt(2))):
        val = X.Integrate({"levelset": g, "domain_type": dt}, cf=X.CoefficientFunction(1), mesh=mesh, order=order)
        assert abs(val - ref) < 5e-15 * (order + 1) ** 2


@pytest.mark.parametrize("n", [1, 2, 5])
@pytest.mark.parametrize("pol", ["FIRST", "OPTIMAL", "FALLBACK"])
def test_integrate_straight_cut_hex(X, n, pol):
    mesh = MakeStructured3DMesh(hexes=True, nx=n, ny=n, nz=n)
    g = p1(X, mesh, 1 - 2 * X.x - 2 * X.y - 2 * X.z)
    policy = getattr(X, pol)
    for dt, ref in ((X.POS, 1 / 48), (X.NEG, 47 / 48), (X.IF, math.sqrt(3) / 8)):
        val = X.Integrate({"levelset": g, "domain_type": dt, "quad_dir_policy": policy}, cf=X.CoefficientFunction(1),
                          mesh=mesh, order=4)
        assert abs(val - ref) < 5e-15 * 25, (dt, val, ref)
    for alpha in (1, 2):
        for f in (X.x, X.y, X.z):
            val_pos = 2 ** (-alpha - 3) / (alpha ** 3 + 6 * alpha * alpha + 11 * alpha + 6)
            val = X.Integrate({"levelset": g, "domain_type": X.POS, "quad_dir_policy": policy}, cf=f ** alpha,
                              mesh=mesh, order=4)
            assert abs(val - val_pos) < 5e-15 * 25


@pytest.mark.parametrize("order", [2, 4])
def test_integrate_tpmc_trilinear(X, order):
    """test_quadchallenges.py:108-141 (10^3 hexes, trilinear level set)"""
    mesh = MakeStructured3DMesh(hexes=True, nx=10, ny=10, nz=10)
    g = p1(X, mesh, X.x * ((7 * X.y - 13) * X.z + 6) + X.y * (3 - 8 * X.z) + 8 * X.z - 4)
    one = X.CoefficientFunction(1)
    assert abs(X.Integrate({"levelset": g, "domain_type": X.IF}, cf=one, mesh=mesh, order=order) - 1.82169) < 5e-3
    assert abs(X.Integrate({"levelset": g, "domain_type": X.NEG}, cf=one, mesh=mesh, order=order) - 0.51681) < 1e-3
    assert abs(X.Integrate({"levelset": g, "domain_type": X.POS}, cf=one, mesh=mesh, order=order) - 0.48319) < 1e-3


def test_integrate_circle_eoc(X):
    """test_straightcutrule.py:33-68 on quads (levels 0..6)"""
    r = 0.6
    one = X.CoefficientFunction(1)
    for dt, ref in ((X.NEG, math.pi * r * r / 4), (X.POS, 1 - math.pi * r * r / 4), (X.IF, r * math.pi / 2)):
        errs = []
        for i in range(7):
            mesh = MakeStructured2DMesh(quads=True, nx=2 ** i, ny=2 ** i)
            g = p1(X, mesh, X.sqrt(X.x * X.x + X.y * X.y) - r)
            errs.append(abs(X.Integrate({"levelset": g, "domain_type": dt}, cf=one, mesh=mesh, order=4) - ref))
        eoc = [math.log(errs[i + 1] / errs[i]) / math.log(0.5) for i in range(len(errs) - 1)]
        assert np.mean(eoc[1:]) > 1.75, (dt, eoc)


@pytest.mark.parametrize("D,n", [(2, 16), (3, 6)])
def test_cutinfo_quads(X, orc, D, n):
    mesh = warped(D, n)
    if D == 2:
        f = X.sqrt((X.x - 0.47) * (X.x - 0.47) + (X.y - 0.52) * (X.y - 0.52)) - 0.3
    else:
        f = X.sqrt((X.x - 0.47) ** 2 + (X.y - 0.52) ** 2 + (X.z - 0.5) ** 2) - 0.33
    g = p1(X, mesh, f)
    lset = np.asarray(g.vec, dtype=np.float64)
    ci = X.CutInfo(mesh, g)
    rdt, rcr, _ = orc.quad_mark(D, mesh.el2v, lset, OPTIMAL)
    for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
        assert np.array_equal(ci.GetElementsOfType(cdt).numpy(), rdt == dt)
    assert np.abs(ci.GetCutRatios() - rcr).max() <= 1e-14
    assert (rdt == IF).sum() > 0
    # boundary elements: segments (2D, simplex rule) / quadrilateral faces (3D)
    bnd = mesh.bnd2v
    if D == 2:
        bdt, _, _ = orc.mark(1, bnd, lset)
    else:
        bdt, _, _ = orc.quad_mark(2, bnd, lset, OPTIMAL)
    for dt, cdt in ((NEG, X.NEG), (POS, X.POS), (IF, X.IF)):
        assert np.array_equal(ci.GetElementsOfType(cdt, X.BND).numpy(), bdt == dt)
    # GetElementsWithSharedVertex on the quad / hex vertex lists
    sv = X.GetElementsWithSharedVertex(mesh, ci.GetElementsOfType(X.IF))
    assert np.array_equal(sv.numpy(), N.elements_with_shared_vertex(mesh.el2v, rdt == IF, mesh.nv))


def test_simplex_only_paths_raise(X):
    """Q2 spaces and several level sets are not built on quads: they raise"""
    mesh = MakeStructured2DMesh(quads=True, nx=4, ny=4)
    g = p1(X, mesh, X.x - 0.5)
    V = X.H1(mesh, order=2)
    u, v = V.TnT()
    a = X.BilinearForm(V)
    a += u * v * X.dCut(g, X.NEG, order=4)
    with pytest.raises(Exception):
        a.Assemble()


QV2 = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)


def test_quad_elmat_uncut_known(dev, orc):
    """Q1 mass (1/36 [4 2 1 2; ...]) and stiffness (1/6 [4 -1 -2 -1; ...]) of
    the unit square, and of the unit cube (sum of all mass entries 1, K 1 = 0)"""
    torch, _abi, ctx = dev
    for D, n in ((2, 1), (3, 1)):
        mesh = MakeStructured2DMesh(quads=True, nx=n, ny=n) if D == 2 else \
            MakeStructured3DMesh(hexes=True, nx=n, ny=n, nz=n)
        lset = -np.ones(mesh.nv)
        off, pts = gpu_rule(dev, mesh, lset, NEG, 2, OPTIMAL)
        nv = 2 ** D
        for mask in (1, 2):
            out = torch.empty(nv * nv, dtype=torch.float64, device="cuda")
            ctx.quad_elmat(D, up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, off), up(torch, pts.reshape(-1)),
                           mask, _abi.XG_PREC_FP64, out)
            ctx.sync()
            m = out.cpu().numpy().reshape(nv, nv)
            if D == 2:
                ref = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36 if mask == 1 else \
                    np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]]) / 6
                assert np.abs(m - ref).max() < 1e-15
            elif mask == 1:
                assert abs(m.sum() - 1.0) < 1e-14 and abs(m[0, 0] - 1 / 27) < 1e-15
            else:
                assert np.abs(m.sum(axis=1)).max() < 1e-14 and abs(m[0, 0] - 1 / 3) < 1e-14


@pytest.mark.parametrize("D,n", [(2, 8), (3, 4)])
@pytest.mark.parametrize("pol", [FIRST, OPTIMAL, FALLBACK])
def test_quad_elmat_parity(dev, orc, D, n, pol):
    torch, _abi, ctx = dev
    mesh = warped(D, n)
    lset = level_sets(mesh, "sphere")
    nv = 2 ** D
    for dt in (NEG, POS, IF):
        off, pts = gpu_rule(dev, mesh, lset, dt, 2, pol)
        for mask in (1, 2, 3):
            out = torch.empty(mesh.ne * nv * nv, dtype=torch.float64, device="cuda")
            ctx.quad_elmat(D, up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, off), up(torch, pts.reshape(-1)),
                           mask, _abi.XG_PREC_FP64, out)
            ctx.sync()
            got = out.cpu().numpy().reshape(mesh.ne, nv, nv)
            for e in range(mesh.ne):
                q = pts[off[e]:off[e + 1]]
                ref = orc.quad_elmat(D, mesh.coords[mesh.el2v[e]], q[:, :3], q[:, 3], mask)
                assert np.abs(got[e] - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max()), (dt, mask, e)


@pytest.mark.parametrize("N_", [1, 10])
def test_dcut_forms_on_quads(X, N_):
    """a dCut bilinear form on quads: the assembled mass matrix sums to the
    measure of the part (7/8, 1/8 for 1 - 2x - 2y; test_straightcutrule.py
    :106-126), the stiffness matrix annihilates constants, and the interface
    mass sums to the interface length 1/sqrt(2)"""
    mesh = MakeStructured2DMesh(quads=True, nx=N_, ny=N_)
    g = p1(X, mesh, 1 - 2 * X.x - 2 * X.y)
    V = X.H1(mesh, order=1)
    u, v = V.TnT()
    for dt, ref in ((X.NEG, 7 / 8), (X.POS, 1 / 8), (X.IF, 1 / math.sqrt(2))):
        a = X.BilinearForm(V)
        a += u * v * X.dCut(g, dt, order=2)
        a.Assemble(precision=X.PREC_FP64)
        assert abs(a.mat.sum() - ref) < 1e-13
    a = X.BilinearForm(V)
    a += X.grad(u) * X.grad(v) * X.dCut(g, X.NEG, order=2)
    a.Assemble(precision=X.PREC_FP64)
    assert np.abs(a.mat @ np.ones(mesh.nv)).max() < 1e-13


@pytest.mark.parametrize("D,n", [(2, 8), (3, 4)])
def test_linear_form_on_quads(X, orc, D, n):
    """f * v * dCut on quads / hexes: the element vector of f = 1 equals the
    row sums of the oracle's mass matrix on the same rule (the Q1 basis sums
    to one), and the assembled vector sums to the measure of the part"""
    mesh = warped(D, n)
    lset = level_sets(mesh, "sphere")
    g = X.GridFunction(X.H1(mesh, order=1))
    g.vec = lset.copy()
    V = X.H1(mesh, order=1)
    u, v = V.TnT()
    for dt in (X.NEG, X.IF):
        f = X.LinearForm(V)
        f += 1.0 * v * X.dCut(g, dt, order=2)
        f.Assemble()
        vol = X.Integrate({"levelset": g, "domain_type": dt}, cf=X.CoefficientFunction(1), mesh=mesh, order=2)
        assert abs(f.vec.sum() - vol) < 1e-13 * max(1.0, vol)
        e_ids, vecs = f.integrals[0].element_vectors(mesh, 1)
        for k, e in enumerate(e_ids[:40]):
            r = orc.quad_cut_rule(D, lset[mesh.el2v[e]], int(dt), 2, 1, V=mesh.coords[mesh.el2v[e]])
            ref = np.zeros(2 ** D) if r is None else orc.quad_elmat(D, mesh.coords[mesh.el2v[e]], r[0], r[1], 1).sum(1)
            assert np.abs(vecs[k] - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
