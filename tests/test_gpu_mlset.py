"""Rules and element marks of several level sets on the GPU (csrc/xg_mlset.hip)
vs the CPU oracle (or_mlset_rule / or_mlset_mark, pinned in
test_mlset_cpu.py): point by point (1e-14), marks bit for bit, plus the
reference's known answers (tests/pytests/test_mlset_integration.py) through
the Python surface (Integrate with a tuple of level sets,
MultiLevelsetCutInfo, dCut forms)."""
import itertools
import math

import numpy as np
import pytest

from ngsxfem_amd.meshes import MakeStructured2DMesh, MakeStructured3DMesh
from oracle import nodes as N

pytestmark = pytest.mark.gpu

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
    return np.stack([np.array([f(*p) for p in mesh.coords]) for f in fs])


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


def gpu_rule(dev, mesh, lv, dtts, order):
    torch, _abi, ctx = dev
    D = mesh.dim
    ne = mesh.ne
    el = up(torch, mesh.el2v)
    co = up(torch, mesh.coords)
    ls = [up(torch, l) for l in lv]
    off = torch.zeros(ne + 1, dtype=torch.int32, device="cuda")
    ctx.mlset_rule_count(D, el, ls, dtts, order, off)
    ctx.sync()
    tot = int(off[-1].item())
    pts = torch.full((max(tot, 1) * 4,), float("nan"), dtype=torch.float64, device="cuda")
    ctx.mlset_rule_fill(D, el, co, ls, dtts, order, off, pts)
    ctx.sync()
    return off.cpu().numpy(), pts.cpu().numpy()[:tot * 4].reshape(tot, 4)


def compare(orc, mesh, lv, dtts, order, off, pts):
    D = mesh.dim
    nonempty = 0
    for e in range(mesh.ne):
        vs = mesh.el2v[e]
        p, w = orc.mlset_rule(D, lv[:, vs], dtts, order, V=mesh.coords[vs])
        g = pts[off[e]:off[e + 1]]
        assert len(g) == len(w), (e, len(g), len(w))
        if len(w) == 0:
            continue
        nonempty += 1
        scale = max(np.abs(w).max(), 1e-300)
        assert np.abs(g[:, :3] - p).max() <= 1e-14, e
        assert np.abs(g[:, 3] - w).max() <= 1e-14 * scale, e
    return nonempty


CASES = {
    "2d_triangle": (2, [(NEG, NEG, NEG)], [(IF, NEG, NEG), (NEG, IF, NEG), (NEG, NEG, IF)],
                    [(IF, IF, NEG), (IF, NEG, IF), (NEG, IF, IF)]),
}


@pytest.mark.parametrize("order", [0, 2, 5])
def test_rule_parity_2d(orc, dev, order):
    mesh = mesh2d(-1, 1, -0.5, 1.5, 9)
    lv = lsets_at(mesh, [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y])
    allt = list(itertools.product((NEG, POS), repeat=3))
    for dtts in ([(NEG, NEG, NEG)], allt, [(IF, NEG, NEG), (NEG, IF, NEG), (NEG, NEG, IF)],
                 [(IF, IF, NEG), (IF, NEG, IF), (NEG, IF, IF)], [(POS, IF, NEG), (IF, POS, POS)]):
        off, pts = gpu_rule(dev, mesh, lv, dtts, order)
        assert compare(orc, mesh, lv, dtts, order, off, pts) > 0


@pytest.mark.parametrize("order", [0, 3])
def test_rule_parity_3d(orc, dev, order):
    mesh = mesh3d(0.8, 4)
    fs = [lambda x, y, z: x - 0.5, lambda x, y, z: x + 0.5, lambda x, y, z: y - 0.5, lambda x, y, z: y + 0.5,
          lambda x, y, z: z - 0.5, lambda x, y, z: z + 0.5]
    lv = lsets_at(mesh, fs)
    cube = (NEG, POS, NEG, POS, NEG, POS)
    sides = [tuple(IF if i == j else cube[i] for i in range(6)) for j in range(6)]
    edges = [(IF, POS, IF, POS, NEG, POS), (NEG, IF, NEG, POS, IF, POS), (NEG, POS, NEG, IF, NEG, IF)]
    corners = [(IF, POS, IF, POS, IF, POS), (NEG, IF, NEG, IF, NEG, IF)]
    for dtts in ([cube], sides, edges, corners):
        off, pts = gpu_rule(dev, mesh, lv, dtts, order)
        assert compare(orc, mesh, lv, dtts, order, off, pts) > 0


def test_rule_parity_random_planes(orc, dev):
    """four random planes through a 3D mesh, all 16 volume tuples and the
    codim-1 tuples of the first plane: GPU = oracle point by point"""
    rng = np.random.default_rng(4)
    mesh = MakeStructured3DMesh(nx=3, ny=3, nz=3)
    fs = []
    for k in range(4):
        n = rng.standard_normal(3)
        c = rng.uniform(0.3, 0.7, 3)
        fs.append(lambda x, y, z, n=n, c=c: n[0] * (x - c[0]) + n[1] * (y - c[1]) + n[2] * (z - c[2]))
    lv = lsets_at(mesh, fs)
    allt = list(itertools.product((NEG, POS), repeat=4))
    off, pts = gpu_rule(dev, mesh, lv, allt, 2)
    compare(orc, mesh, lv, allt, 2, off, pts)
    # the volume tuples tile every element
    torch = dev[0]
    w = pts[:, 3]
    for e in range(mesh.ne):
        V = mesh.coords[mesh.el2v[e]]
        ad = abs(np.linalg.det((V[:3] - V[3]).T))
        assert abs(w[off[e]:off[e + 1]].sum() * ad - ad / 6) < 1e-14
    ift = [(IF,) + t for t in itertools.product((NEG, POS), repeat=3)]
    off, pts = gpu_rule(dev, mesh, lv, ift, 2)
    compare(orc, mesh, lv, ift, 2, off, pts)


def test_marks_match_oracle(orc, dev):
    torch, _abi, ctx = dev
    mesh = mesh2d(-1, 1, -0.5, 1.5, 12)
    lv = lsets_at(mesh, [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y])
    el, co = up(torch, mesh.el2v), up(torch, mesh.coords)
    ls = [up(torch, l) for l in lv]
    nw = (mesh.ne + 31) // 32
    for dtts in ([(NEG, NEG, NEG)], [(IF, NEG, NEG), (NEG, IF, NEG), (NEG, NEG, IF)],
                 [t for t in itertools.product((NEG, POS), repeat=3) if t != (NEG, NEG, NEG)]):
        ot = torch.full((nw,), -1, dtype=torch.int32, device="cuda")
        cb = torch.full((nw,), -1, dtype=torch.int32, device="cuda")
        ctx.mlset_mark(2, el, co, ls, dtts, ot, cb)
        ctx.sync()
        r_ot, r_cb = orc.mlset_mark(2, mesh.el2v, mesh.coords, lv, dtts)
        assert np.array_equal(N.bits_to_bool(ot.cpu().numpy(), mesh.ne), r_ot)
        assert np.array_equal(N.bits_to_bool(cb.cpu().numpy(), mesh.ne), r_cb)
        assert r_cb.any()


def gfs(X, mesh, fs):
    out = []
    for f in fs:
        g = X.GridFunction(X.H1(mesh, order=1))
        g.vec[:] = np.array([f(*p) for p in mesh.coords])
        out.append(g)
    return tuple(out)


def test_integrate_known_answers_2d(X):
    """test_mlset_integration.py:37-142 through Integrate with a tuple of
    level sets and DomainTypeArray domains"""
    from ngsxfem_amd.xfem.mlset import DomainTypeArray
    mesh = mesh2d(-1, 1, -0.5, 1.5, 9)
    ls = gfs(X, mesh, [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y])
    tri = DomainTypeArray((NEG, NEG, NEG))

    def I(dom, cf=1):
        return X.Integrate({"levelset": ls, "domain_type": dom}, cf=cf, mesh=mesh, order=0)
    assert abs(I(tri) - 0.5) < 1e-12
    assert abs(I(~tri) - 3.5) < 1e-12
    assert abs(I((IF, NEG, NEG)) - 1) < 1e-12
    assert abs(I((NEG, IF, NEG)) - math.sqrt(5 / 4)) < 1e-12
    assert abs(I(tri.Boundary()) - 1 - math.sqrt(5)) < 1e-12
    assert abs(I((IF, IF, NEG), X.x + X.y) - 1.5) < 1e-12
    assert abs(I(tri.Boundary().Boundary(), X.x + X.y) - 2.0) < 1e-12


def test_integrate_high_order_and_3d(X):
    from ngsxfem_amd.xfem.mlset import DomainTypeArray
    mesh = mesh2d(-0.2, 1.2, -0.2, 1.2, 7)
    ls = gfs(X, mesh, [lambda x, y: -y, lambda x, y: x - 1, lambda x, y: y - 1, lambda x, y: -x])
    r = X.Integrate({"levelset": ls, "domain_type": DomainTypeArray([(NEG, NEG, NEG, NEG)])},
                    cf=X.x * (1 - X.x) * X.y * (1 - X.y), mesh=mesh, order=4)
    assert abs(r - 1 / 36) < 1e-12
    mesh = mesh3d(0.8, 5)
    ls = gfs(X, mesh, [lambda x, y, z: x - 0.5, lambda x, y, z: x + 0.5, lambda x, y, z: y - 0.5,
                       lambda x, y, z: y + 0.5, lambda x, y, z: z - 0.5, lambda x, y, z: z + 0.5])
    cube = DomainTypeArray([(NEG, POS, NEG, POS, NEG, POS)])

    def I(dom, cf=1):
        return X.Integrate({"levelset": ls, "domain_type": dom}, cf=cf, mesh=mesh, order=0)
    assert abs(I(cube) - 1) < 1e-12
    assert abs(I(~cube) - 3.096) < 1e-12
    assert abs(I(cube.Boundary()) - 6) < 1e-12
    assert abs(I(cube.Boundary().Boundary()) - 12) < 1e-12
    assert abs(I((NEG, IF, NEG, IF, NEG, IF), X.x + X.y + X.z) + 1.5) < 1e-12


def test_multilevelset_cutinfo(X):
    """test_mlset_integration.py:62-77 and 250-300: marks, cached arrays,
    Update, safety checks"""
    from ngsxfem_amd.xfem.mlset import DomainTypeArray
    mesh = mesh2d(-1, 1, -0.5, 1.5, 9)
    ls = gfs(X, mesh, [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y])
    tri = DomainTypeArray((NEG, NEG, NEG))
    mlci = X.MultiLevelsetCutInfo(mesh, ls)
    neg = mlci.GetElementsOfType(tri).numpy()
    assert np.array_equal(neg, mlci.GetElementsOfType(tri.as_list).numpy())
    hasneg = mlci.GetElementsWithContribution(tri).numpy()
    if_ = mlci.GetElementsWithContribution(tri.Boundary()).numpy()
    notneg = mlci.GetElementsWithContribution(~tri).numpy()
    assert np.array_equal(~neg, notneg)
    assert np.array_equal(if_, hasneg & ~neg)
    # kept arrays: the level sets changed in place are only seen after Update
    mesh2 = mesh2d(-1, 1, -1, 1, 10)
    l2 = gfs(X, mesh2, [lambda x, y: x + 0.5, lambda x, y: x - 0.5])
    mc = X.MultiLevelsetCutInfo(mesh2, l2)
    b0 = mc.GetElementsOfType((POS, NEG)).numpy().copy()
    l2[0].vec[:] = mesh2.coords[:, 1] + 0.5
    l2[1].vec[:] = mesh2.coords[:, 1] - 0.5
    assert np.array_equal(mc.GetElementsOfType((POS, NEG)).numpy(), b0)
    mc.Update(l2)
    assert not np.array_equal(mc.GetElementsOfType((POS, NEG)).numpy(), b0)
    with pytest.raises(Exception):
        X.MultiLevelsetCutInfo(mesh2, (X.x, X.y))
    with pytest.raises(Exception):
        mc.GetElementsOfType((NEG, NEG, NEG))
    with pytest.raises(Exception):
        mc.GetElementsWithContribution([(NEG, NEG, NEG)])


def test_dcut_element_matrices(X, orc):
    """dCut forms on the rule of several level sets = the oracle's element
    matrices on the oracle's mlset rule; the mass matrices sum to the area"""
    from ngsxfem_amd.xfem.mlset import DomainTypeArray
    mesh = mesh2d(-1, 1, -0.5, 1.5, 8)
    fs = [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y]
    ls = gfs(X, mesh, fs)
    lv = np.stack([g.vec for g in ls])
    tri = DomainTypeArray((NEG, NEG, NEG))
    V = X.H1(mesh, order=2)
    u, v = V.TnT()
    a = X.BilinearForm(V)
    a += (u * v) * X.dCut(levelset=ls, domain_type=tri, order=4)
    a.Assemble(precision=X.PREC_FP64)
    one = np.ones(a.mat.shape[0])
    assert abs(one @ (a.mat @ one) - 0.5) < 1e-12
    els, mats = a.elmats[0]
    for k, e in enumerate(els):
        vs = mesh.el2v[e]
        p, w = orc.mlset_rule(2, lv[:, vs], tri.as_list, 4, V=mesh.coords[vs])
        ref = orc.elmat(2, 2, mesh.coords[vs], p, w, mask=1) if len(w) else np.zeros((6, 6))
        assert np.abs(mats[k] - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-300)


def test_dcut_coefficient_function(X, orc):
    """a coefficient function on the rule of several level sets: the
    oracle's element matrices of its mlset rule with the weights scaled by
    the coefficient at the physical points (numpy evaluation)"""
    from ngsxfem_amd.xfem.mlset import DomainTypeArray
    mesh = mesh2d(-1, 1, -0.5, 1.5, 8)
    fs = [lambda x, y: y - 1, lambda x, y: 2 * x - y, lambda x, y: -2 * x - y]
    ls = gfs(X, mesh, fs)
    lv = np.stack([g.vec for g in ls])
    tri = DomainTypeArray((NEG, NEG, NEG))
    V = X.H1(mesh, order=2)
    u, v = V.TnT()
    c = 1 + X.x * X.y + X.sin(X.y)
    integral = ((c * u * v) * X.dCut(levelset=ls, domain_type=tri, order=4)).parts[0]
    els, mats = integral.element_matrices(mesh, 2, X.PREC_FP64)
    assert len(els) > 0
    for k, e in enumerate(els):
        vs = mesh.el2v[e]
        Vx = mesh.coords[vs]
        p, w = orc.mlset_rule(2, lv[:, vs], tri.as_list, 4, V=Vx)
        if not len(w):
            assert np.abs(mats[k]).max() == 0.0
            continue
        xp = np.zeros((len(w), 3))
        xp[:, :2] = Vx[2] + p[:, :2] @ np.stack([Vx[0] - Vx[2], Vx[1] - Vx[2]], axis=1).T
        ref = orc.elmat(2, 2, Vx, p, w * c.evaluate(xp), mask=1)
        assert np.abs(mats[k] - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1e-300)
