"""Coefficient functions evaluated on the GPU (xg_rule_scale_cf, xg_elvec,
and the xfem surface: IntegrateX with non-polynomial integrands, c u v /
c grad u . grad v / f v forms with coefficient functions) against the FP64
oracle's straight cut rule of each element, mapped to the physical element,
with the coefficient evaluated by numpy (CoefficientFunction.evaluate, the
same expression tree): integrals per element to 1e-12, element matrices to
1e-12 (Frobenius, per element) against the oracle's element matrices of the
rule with scaled weights, element vectors against the row sums of those mass
matrices (the Lagrange basis sums to one).  Rule orders as the reference's
integrators (symboliccutbfi.cpp:107-118: 2p, 2p - 2 for gradients only;
symboliccutlfi.cpp:93-94: 2p)."""
import numpy as np
import pytest

from ngsxfem_amd import workloads as W

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
    from ngsxfem_amd import xfem
    return xfem


def up(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def oracle_rule(orc, D, V, lvals, dt, order):
    """oracle rule of one element mapped to physical points: (x (n x 3), w)"""
    F = np.stack([V[c] - V[D] for c in range(D)], axis=1)  # columns V_c - V_D
    _, Fi, _ = orc.jacobian(V)
    r = orc.straight_cut_rule(D, lvals, dt, order, Finv=Fi)
    if r is None:
        return np.zeros((0, 3)), np.zeros(0), None
    p, w, _ = r
    x = np.zeros((len(w), 3))
    x[:, :D] = V[D] + p[:, :D] @ F.T
    return x, w * abs(np.linalg.det(F)), (p, w)


def coefficient(X):
    return X.sin(3 * X.x) * X.exp(X.y) + X.IfPos(X.x - 0.4, X.sqrt(1 + X.y * X.y), 0.5) + X.z / (2 + X.x)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("dt", [NEG, POS, IF])
def test_rule_scale_cf_integrals(orc, dev, X, D, dt):
    torch, _abi, ctx = dev
    mesh, lset = W.config1() if D == 2 else W.config2(8)
    order = 4
    dto, _, _ = orc.mark(D, mesh.el2v, lset)
    cut = np.nonzero(dto == IF)[0]
    unc = np.nonzero(dto != IF)[0][:40]
    elems = np.concatenate([cut[:300], unc]).astype(np.int32)
    d_el, d_x, d_l, d_e = up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, lset), up(torch, elems)
    nel = len(elems)
    off = torch.zeros(nel + 1, dtype=torch.int32, device="cuda")
    ctx.cut_rule_count(D, d_el, d_l, dt, order, off, elems=d_e)
    ctx.sync()
    tot = int(off[-1].item())
    pts = torch.zeros(max(tot, 1) * 4, dtype=torch.float64, device="cuda")
    ctx.cut_rule_fill(D, d_el, d_x, d_l, dt, order, off, pts, elems=d_e)
    cf = coefficient(X)
    code, consts, gfs = cf.program()
    assert not gfs
    print("rule_scale_cf D=%d dt=%d nel=%d points=%d" % (D, dt, nel, tot), flush=True)
    ctx.rule_scale_cf(D, d_el, d_x, off, pts, code, consts, elems=d_e)
    total = torch.zeros(1, dtype=torch.float64, device="cuda")
    pe = torch.zeros(nel, dtype=torch.float64, device="cuda")
    ctx.integrate(D, d_el, d_x, off, pts, [(1.0, (0, 0, 0))], total, per_elem=pe, elems=d_e)
    ctx.sync()
    got = pe.cpu().numpy()
    ref = np.zeros(nel)
    for i, e in enumerate(elems):
        V = mesh.coords[mesh.el2v[e]]
        x, w, _ = oracle_rule(orc, D, V, lset[mesh.el2v[e]], dt, order)
        ref[i] = np.sum(w * cf.evaluate(x))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)


def test_integrate_nonpolynomial_on_gpu(orc, X):
    """IntegrateX with a non-polynomial integrand and a P1 grid function in
    it, element by element"""
    mesh = X.MakeStructured3DMesh(nx=5, ny=5, nz=5)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.sqrt((X.x - 0.5) ** 2 + (X.y - 0.5) ** 2 + (X.z - 0.5) ** 2) - 0.35, lset)
    g = X.GridFunction(X.H1(mesh, order=1))
    g.Set(1 + 2 * X.x - X.y + 0.5 * X.z)  # linear: its P1 interpolant is exact
    cf = coefficient(X) + g * X.cos(X.z)
    ref_cf = coefficient(X) + (1 + 2 * X.x - X.y + 0.5 * X.z) * X.cos(X.z)
    for dt in (X.NEG, X.IF):
        got = X.Integrate({"levelset": lset, "domain_type": dt}, cf=cf, mesh=mesh, order=4, element_wise=True)
        ref = np.zeros(mesh.ne)
        for e in range(mesh.ne):
            V = mesh.coords[mesh.el2v[e]]
            x, w, _ = oracle_rule(orc, 3, V, lset.vec[mesh.el2v[e]], int(dt), 4)
            ref[e] = np.sum(w * ref_cf.evaluate(x))
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


def rel_fro(a, b):
    return np.sqrt(np.sum((a - b) ** 2, axis=(-2, -1))) / np.maximum(np.sqrt(np.sum(b ** 2, axis=(-2, -1))), 1e-300)


@pytest.mark.parametrize("p", [1, 2])
def test_coefficient_forms(orc, X, p):
    mesh = X.MakeStructured3DMesh(nx=4, ny=4, nz=4)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.x + 0.7 * X.y - 0.3 * X.z - 0.55, lset)
    V = X.H1(mesh, order=p)
    u, v = V.TnT()
    c = 1 + X.x * X.y + X.sin(X.z)
    ci = X.CutInfo(mesh, lset)
    els = ci.GetElementsOfType(X.HASNEG)
    for form, mask, order in ((c * u * v, 1, 2 * p), (c * X.grad(u) * X.grad(v), 2, 2 * p - 2)):
        integral = (form * X.dCut(lset, X.NEG, definedonelements=els)).parts[0]
        e, m = integral.element_matrices(mesh, p, 0)  # FP64
        assert len(e) > 0
        for i, el in enumerate(e):
            Vx = mesh.coords[mesh.el2v[el]]
            x, _, r = oracle_rule(orc, 3, Vx, lset.vec[mesh.el2v[el]], NEG, order)
            if r is None:
                assert np.abs(m[i]).max() == 0.0
                continue
            pr, w = r
            ref = orc.elmat(3, p, Vx, pr, w * c.evaluate(x), mask=mask)
            assert rel_fro(m[i], ref) < 1e-12, (p, mask, el)
    # f v: rows of the mass matrix of the f-scaled rule
    f = X.exp(X.x) - X.y * X.z
    integral = (f * v * X.dCut(lset, X.NEG, definedonelements=els)).parts[0]
    e, vec = integral.element_vectors(mesh, p)
    for i, el in enumerate(e):
        Vx = mesh.coords[mesh.el2v[el]]
        x, _, r = oracle_rule(orc, 3, Vx, lset.vec[mesh.el2v[el]], NEG, 2 * p)
        pr, w = r
        ref = orc.elmat(3, p, Vx, pr, w * f.evaluate(x), mask=1).sum(axis=1)
        np.testing.assert_allclose(vec[i], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_coefficient_form_assembly(X):
    """a form with a coefficient function assembles (host scatter of the GPU
    element matrices) and equals the sum of its element matrices"""
    mesh = X.MakeStructured2DMesh(nx=8, ny=8)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.x - 0.37, lset)
    V = X.H1(mesh, order=1)
    u, v = V.TnT()
    a = X.BilinearForm(V)
    a += (1 + X.exp(X.y)) * u * v * X.dCut(lset, X.NEG)
    a.Assemble(precision=0)
    one = np.ones(mesh.nv)
    val = X.Integrate({"levelset": lset, "domain_type": X.NEG}, cf=1 + X.exp(X.y), mesh=mesh, order=2)
    assert abs(one @ (a.mat @ one) - val) < 1e-12 * abs(val)


def test_program_too_deep_is_an_error(X):
    mesh = X.MakeStructured2DMesh(nx=2, ny=2)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.x - 0.5, lset)
    cf = X.sin(X.x)
    for _ in range(20):
        cf = X.sin(X.x) + cf
    with pytest.raises(RuntimeError, match="stack"):
        X.Integrate({"levelset": lset, "domain_type": X.NEG}, cf=cf, mesh=mesh, order=2)


def test_interpolate_to_p1_on_gpu(X):
    """InterpolateToP1 of a coefficient function (utils/p1interpol.cpp:27-113)
    evaluates the coefficient's program at the vertices on the GPU: on C1's
    mesh, sqrt(x^2 + y^2) - 1 gives the numpy vertex values and the
    perturbation bit for bit (square roots and the other operations are
    correctly rounded in both), exactly 12 vertices at +1e-14, and 270 cut /
    2356 NEG triangles; a non-polynomial 3D level set with a P1 grid function
    in it matches numpy to rounding"""
    mesh = X.MakeStructured2DMesh(nx=60, ny=60, mapping=lambda x, y: (-1.5 + 3 * x, -1.5 + 3 * y))
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.sqrt(X.x * X.x + X.y * X.y) - 1.0, lset)
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    raw = np.sqrt(x * x + y * y) - 1.0
    ref = np.where(np.abs(raw) < 1e-14, 1e-14, raw)
    assert np.array_equal(lset.vec.view(np.int64), ref.view(np.int64))
    assert int(np.sum(lset.vec == 1e-14)) == 12
    ci = X.CutInfo(mesh, lset)
    assert ci.GetElementsOfType(X.IF).NumSet() == 270 and ci.GetElementsOfType(X.NEG).NumSet() == 2356
    m3 = X.MakeStructured3DMesh(nx=6, ny=6, nz=6)
    g = X.GridFunction(X.H1(m3, order=1))
    g.Set(0.5 * X.x - X.z)
    l3 = X.GridFunction(X.H1(m3, order=1))
    X.InterpolateToP1(X.exp(X.y) * X.cos(X.x) - 1.2 + g, l3)
    X3 = m3.coords
    ref3 = np.exp(X3[:, 1]) * np.cos(X3[:, 0]) - 1.2 + (0.5 * X3[:, 0] - X3[:, 2])
    np.testing.assert_allclose(l3.vec, np.where(np.abs(ref3) < 1e-14, 1e-14, ref3), rtol=1e-14, atol=1e-15)


def test_piecewise_constant_coefficients(X):
    """CutRatioGF / IndicatorCF (python/__init__.py:441-462) as coefficients:
    the integral of the cut ratio over each element's NEG part is the ratio
    times that part's volume; the indicator of the cut elements keeps their
    mass matrices and zeroes the others; kappa and IsCut agree with the
    cut ratios and the IF elements; vertex interpolation rejects them"""
    mesh = X.MakeStructured3DMesh(nx=5, ny=5, nz=5)
    lset = X.GridFunction(X.H1(mesh, order=1))
    X.InterpolateToP1(X.sqrt((X.x - 0.5) ** 2 + (X.y - 0.45) ** 2 + (X.z - 0.5) ** 2) - 0.33, lset)
    ci = X.CutInfo(mesh, lset)
    kap = X.CutRatioGF(ci)
    lsd = {"levelset": lset, "domain_type": X.NEG}
    vol = X.Integrate(lsd, cf=1, mesh=mesh, order=2, element_wise=True)
    got = X.Integrate(lsd, cf=kap * 1.0 + 0.0 * X.x, mesh=mesh, order=2, element_wise=True)
    np.testing.assert_allclose(got, kap.vec * vol, rtol=1e-13, atol=1e-16)
    V = X.H1(mesh, order=1)
    u, v = V.TnT()
    ifel = ci.GetElementsOfType(X.IF)
    ind = X.IndicatorCF(mesh, ifel)
    e1, m1 = ((ind * u * v) * X.dCut(lset, X.NEG)).parts[0].element_matrices(mesh, 1, 0)
    e0, m0 = ((u * v) * X.dCut(lset, X.NEG)).parts[0].element_matrices(mesh, 1, 0)
    assert np.array_equal(e1, e0)
    isif = np.asarray(ifel.numpy(), dtype=bool)[e1]
    assert np.abs(m1[~isif]).max() == 0.0
    np.testing.assert_allclose(m1[isif], m0[isif], rtol=1e-13, atol=1e-18)
    k1, k2 = X.kappa(mesh, lset)
    assert np.allclose(k1.vec + k2.vec, 1.0) and np.array_equal(k1.vec, kap.vec)
    assert int(X.IsCut(mesh, lset).vec.sum()) == ifel.NumSet()
    g = X.GridFunction(X.H1(mesh, order=1))
    with pytest.raises(RuntimeError):
        X.InterpolateToP1(kap * X.x, g)


def test_set_p1_from_grid_function_expression(X):
    """GridFunction.Set on P1 evaluates the coefficient's program at the
    vertices on the GPU, so expressions of grid functions interpolate too"""
    mesh = X.MakeStructured3DMesh(nx=4, ny=4, nz=4)
    V = X.H1(mesh, order=1)
    h, g = X.GridFunction(V), X.GridFunction(V)
    h.vec = np.arange(mesh.nv, dtype=np.float64) / mesh.nv
    g.Set(2 * h + X.sin(X.x) - X.IfPos(X.y - 0.5, 1.0, 0.0))
    xc = mesh.coords
    ref = 2 * h.vec + np.sin(xc[:, 0]) - (xc[:, 1] > 0.5)
    np.testing.assert_allclose(g.vec, ref, rtol=1e-15, atol=1e-15)
