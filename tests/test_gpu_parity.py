"""HIP path vs the FP64 oracle, through the C ABI (include/xg_abi.h).

Bars: marking (dt, cut ratio bits, the 8 bit arrays, the compacted cut list)
and sub-simplex counts are bit-exact; rule points/weights agree to 1e-14
(identical FP64 operation order, base-rule tables within 1 ulp); FP64
element matrices to 1e-12 relative; FP32 element matrices to 1e-6 relative
(Frobenius norm per element), the tolerance the element-matrix path states.
"""
import math

import numpy as np
import pytest

from ngsxfem_amd import workloads as W
from ngsxfem_amd.meshes import MakeStructured2DMesh, MakeStructured3DMesh, Make1DMesh

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


def up(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def gpu_mark(dev, D, el2v, lset):
    torch, _abi, ctx = dev
    ne = el2v.shape[0]
    nw = (ne + 31) // 32
    d_el = up(torch, el2v.astype(np.int32))
    d_l = up(torch, lset)
    dt = torch.zeros(ne, dtype=torch.uint8, device="cuda")
    cr = torch.zeros(ne, dtype=torch.float64, device="cuda")
    bits = torch.zeros(8 * nw, dtype=torch.int32, device="cuda")
    cl = torch.full((ne,), -1, dtype=torch.int32, device="cuda")
    nc = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.mark(D, d_el, d_l, dt, cr, bits, cl, nc)
    ctx.sync()
    n = int(nc.item())
    return (dt.cpu().numpy(), cr.cpu().numpy(), bits.cpu().numpy().view(np.uint32).reshape(8, nw),
            cl.cpu().numpy()[:n])


def assert_mark_parity(orc, dev, D, el2v, lset):
    dt_o, cr_o, bits_o = orc.mark(D, el2v, lset)
    dt_g, cr_g, bits_g, cl_g = gpu_mark(dev, D, el2v, lset)
    assert np.array_equal(dt_o, dt_g)
    assert np.array_equal(cr_o.view(np.int64), cr_g.view(np.int64)), "cut ratio not bit-exact"
    assert np.array_equal(bits_o, bits_g)
    assert np.array_equal(np.nonzero(dt_o == IF)[0].astype(np.int32), cl_g)
    return dt_o


def test_mark_config1(orc, dev):
    mesh, lset = W.config1()
    dt = assert_mark_parity(orc, dev, 2, mesh.el2v, lset)
    assert int(np.sum(dt == IF)) == 270
    assert int(np.sum(dt == NEG)) == 2356


def test_p1_perturbation_config1(orc, dev):
    """InterpolateToP1's perturbation (utils/p1interpol.cpp:104-107) through
    xg_p1_perturb on C1's raw vertex values |x| - 1: exactly the 12 vertices
    with |v| < 1e-14 (FP64 gives |v| ~ 2.2e-16 at points such as (0.6, 0.8))
    become +1e-14, bit for bit; marking the perturbed values gives 270 cut
    and 2356 NEG triangles, as the oracle does"""
    torch, _abi, ctx = dev
    mesh = MakeStructured2DMesh(nx=60, ny=60, mapping=lambda x, y: (-1.5 + 3 * x, -1.5 + 3 * y))
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    raw = np.sqrt(x * x + y * y) - 1.0
    d = up(torch, raw)
    ctx.p1_perturb(d, 1e-14)
    ctx.sync()
    g = d.cpu().numpy()
    changed = np.nonzero(g.view(np.int64) != raw.view(np.int64))[0]
    assert len(changed) == 12
    assert np.all(np.abs(raw[changed]) < 1e-14) and np.all(g[changed] == 1e-14)
    assert np.sum(np.abs(raw) < 1e-14) == 12
    dt, cr, bits, cl = gpu_mark(dev, 2, mesh.el2v, g)
    assert int(np.sum(dt == IF)) == 270 and int(np.sum(dt == NEG)) == 2356
    dt_o, cr_o, bits_o = orc.mark(2, mesh.el2v, g)
    assert np.array_equal(dt, dt_o) and np.array_equal(bits, bits_o)


def test_mark_config2_full(orc, dev):
    mesh, lset = W.config2(64)
    dt = assert_mark_parity(orc, dev, 3, mesh.el2v, lset)
    assert 40000 < int(np.sum(dt == IF)) < 46000


@pytest.mark.parametrize("D", [1, 2, 3])
def test_mark_fuzz_special_values(orc, dev, D):
    """random sign patterns with denormals, huge and tiny magnitudes and
    +0 (no exact zero next to a negative value: that raises, see below)"""
    rng = np.random.default_rng(7 + D)
    ne = 20000
    el2v = np.arange(ne * (D + 1), dtype=np.int32).reshape(ne, D + 1)
    choices = np.array([1.0, -1.0, 5e-324, -5e-324, 1e-300, -1e-300, 1e300, -1e300, 0.3, -0.7,
                        1e-14, -1e-14, 2.2e-308, -2.2e-308])
    l = rng.choice(choices, size=ne * (D + 1)) * rng.uniform(0.5, 2.0, size=ne * (D + 1))
    l = l.reshape(ne, D + 1)
    # exact zeros only in elements without negative values
    # (an all-zero element raises as well, so one entry stays positive)
    z = rng.random((ne, D + 1)) < 0.1
    z[:, 0] = False
    l[:, 0] = np.abs(l[:, 0])
    nonneg = (l >= 0).all(axis=1)
    l[z & nonneg[:, None]] = 0.0
    l = l.reshape(-1)
    assert_mark_parity(orc, dev, D, el2v, l)


@pytest.mark.parametrize("ne", [1, 5, 63, 64, 65, 255, 1023, 1024, 1025, 4097])
def test_mark_small_and_ragged_sizes(orc, dev, ne):
    """element counts around the wave (64), block (1024 elements: 256
    threads x 4) and bit-array word boundaries, with a cut in every block"""
    rng = np.random.default_rng(ne)
    D = 3
    el2v = rng.integers(0, 50, size=(ne, D + 1)).astype(np.int32)
    lset = rng.uniform(-1, 1, 50)
    assert_mark_parity(orc, dev, D, el2v, lset)


def test_mark_degenerate_raises(orc, dev):
    torch, _abi, ctx = dev
    el2v = np.array([[0, 1, 2, 3]], np.int32)
    lset = np.array([0.0, 1.0, -1.0, 1.0])
    with pytest.raises(RuntimeError):
        orc.mark(3, el2v, lset)
    with pytest.raises(_abi.XgError):
        gpu_mark(dev, 3, el2v, lset)
    # the context recovers after the error
    gpu_mark(dev, 3, el2v, np.array([1.0, 1.0, -1.0, 1.0]))


def gpu_rule(dev, D, mesh, lset, dt, order, elems=None):
    torch, _abi, ctx = dev
    d_el = up(torch, mesh.el2v)
    d_x = up(torch, mesh.coords)
    d_l = up(torch, lset)
    d_e = up(torch, np.asarray(elems, np.int32)) if elems is not None else None
    nel = len(elems) if elems is not None else mesh.ne
    off = torch.zeros(nel + 1, dtype=torch.int32, device="cuda")
    ns = torch.zeros(nel, dtype=torch.uint8, device="cuda")
    ctx.cut_rule_count(D, d_el, d_l, dt, order, off, elems=d_e, nsub=ns)
    ctx.sync()
    tot = int(off[-1].item())
    pts = torch.zeros(max(tot, 1) * 4, dtype=torch.float64, device="cuda")
    ctx.cut_rule_fill(D, d_el, d_x, d_l, dt, order, off, pts, elems=d_e)
    ctx.sync()
    return off.cpu().numpy(), ns.cpu().numpy(), pts.cpu().numpy()[: 4 * tot].reshape(tot, 4)


@pytest.mark.parametrize("cfg", ["c1", "c2_16"])
@pytest.mark.parametrize("dt", [NEG, POS, IF])
@pytest.mark.parametrize("order", [0, 2, 4, 6])
def test_rule_parity(orc, dev, cfg, dt, order):
    if cfg == "c1":
        mesh, lset = W.config1()
        D = 2
    else:
        mesh, lset = W.config2(16)
        D = 3
    dto, _, _ = orc.mark(D, mesh.el2v, lset)
    cut = np.nonzero(dto == IF)[0]
    rng = np.random.default_rng(1)
    unc = rng.choice(np.nonzero(dto != IF)[0], 50, replace=False)
    elems = np.concatenate([cut[:400], unc]).astype(np.int32)
    off, ns, pts = gpu_rule(dev, D, mesh, lset, dt, order, elems)
    for i, e in enumerate(elems):
        V = mesh.coords[mesh.el2v[e]]
        _, Fi, _ = orc.jacobian(V)
        r = orc.straight_cut_rule(D, lset[mesh.el2v[e]], dt, order, Finv=Fi)
        g = pts[off[i]:off[i + 1]]
        if r is None:
            assert len(g) == 0
            assert ns[i] == 0
            continue
        p, w, nsub = r
        assert ns[i] == nsub
        assert len(g) == len(w)
        np.testing.assert_allclose(g[:, :3], p, rtol=0, atol=1e-14)
        np.testing.assert_allclose(g[:, 3], w, rtol=1e-14, atol=1e-300)


@pytest.mark.parametrize("dt", [NEG, POS, IF])
def test_integrate_config1(orc, dev, dt):
    torch, _abi, ctx = dev
    mesh, lset = W.config1()
    order = 2
    off, ns, pts = gpu_rule(dev, 2, mesh, lset, dt, order)
    d_off = up(torch, off)
    d_pts = up(torch, pts.reshape(-1))
    tot = torch.zeros(1, dtype=torch.float64, device="cuda")
    pe = torch.zeros(mesh.ne, dtype=torch.float64, device="cuda")
    terms = [(1.0, (0, 0, 0)), (0.5, (2, 1, 0))]
    ctx.integrate(2, up(torch, mesh.el2v), up(torch, mesh.coords), d_off, d_pts, terms, tot, per_elem=pe)
    ctx.sync()
    ref_pe = orc.integrate_mesh(mesh, lset, dt, order, terms=terms, element_wise=True)
    np.testing.assert_allclose(pe.cpu().numpy(), ref_pe, rtol=1e-12, atol=1e-15)
    assert abs(tot.item() - ref_pe.sum()) < 1e-12
    if dt == NEG:
        assert abs(orc.integrate_mesh(mesh, lset, NEG, 2) - math.pi) < 5e-3


def rel_fro(a, b):
    num = np.sqrt(np.sum((a - b) ** 2, axis=(-2, -1)))
    den = np.sqrt(np.sum(b ** 2, axis=(-2, -1)))
    return num / np.maximum(den, 1e-300)


@pytest.mark.parametrize("D,p,n", [(2, 1, 0), (2, 2, 0), (2, 3, 0), (3, 1, 8), (3, 2, 8), (3, 3, 8), (3, 2, 64),
                                   (3, 3, 64)])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_pipeline_elmat_parity(orc, dev, D, p, n, prec):
    """FP64 MFMA 1e-12; FP32 MFMA and split-BF16 MFMA (both with FP64 sums of
    8-point groups) 1e-6 -- on the full 64^3 3D plain mesh too, P2 and P3"""
    torch, _abi, ctx = dev
    if D == 2:
        mesh, lset = W.config1()
    else:
        mesh, lset = W.config2(n)
    order = 2 * p
    dto, _, _ = orc.mark(D, mesh.el2v, lset)
    cut = np.nonzero(dto == IF)[0].astype(np.int32)
    neg = np.nonzero(dto == NEG)[0][:64].astype(np.int32)
    elems = np.concatenate([cut, neg])
    dts = [NEG, IF, POS]
    ref = orc.cut_pipeline(D, p, order, elems, mesh.el2v, mesh.coords, lset, dts, mask=3, nthreads=8)
    nd = ref.shape[-1]
    P = {"fp64": _abi.XG_PREC_FP64, "fp32": _abi.XG_PREC_FP32, "bf16x3": _abi.XG_PREC_BF16X3}[prec]
    out = torch.zeros(len(elems) * len(dts) * nd * nd,
                      dtype=torch.float64 if prec == "fp64" else torch.float32, device="cuda")
    ctx.cut_pipeline(D, p, order, up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, lset), dts, 3, P,
                     out, elems=up(torch, elems))
    ctx.sync()
    g = out.cpu().numpy().astype(np.float64).reshape(ref.shape)
    nz = np.sqrt(np.sum(ref ** 2, axis=(-2, -1))) > 0
    assert np.all(g[~nz] == 0)
    err = rel_fro(g[nz], ref[nz])
    tol = 1e-12 if prec == "fp64" else 1e-6
    assert err.max() < tol, "max rel err %g" % err.max()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_csr_elmat_matches_pipeline(orc, dev, prec):
    torch, _abi, ctx = dev
    mesh, lset = W.config2(8)
    D, p, order = 3, 2, 4
    dto, _, _ = orc.mark(D, mesh.el2v, lset)
    elems = np.nonzero(dto == IF)[0].astype(np.int32)
    off, ns, pts = gpu_rule(dev, D, mesh, lset, NEG, order, elems)
    P = {"fp64": _abi.XG_PREC_FP64, "fp32": _abi.XG_PREC_FP32, "bf16x3": _abi.XG_PREC_BF16X3}[prec]
    tdt = torch.float64 if prec == "fp64" else torch.float32
    out = torch.zeros(len(elems) * 100, dtype=tdt, device="cuda")
    d_el, d_x, d_e = up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, elems)
    ctx.elmat(D, p, d_el, d_x, up(torch, off), up(torch, pts.reshape(-1)), 3, P, out, elems=d_e)
    out2 = torch.zeros_like(out)
    ctx.cut_pipeline(D, p, order, d_el, d_x, up(torch, lset), [NEG], 3, P, out2, elems=d_e)
    ctx.sync()
    # same rule, built in FP64 by different kernels (materialised vs staged)
    a = out.cpu().numpy().astype(np.float64).reshape(-1, 10, 10)
    b = out2.cpu().numpy().astype(np.float64).reshape(-1, 10, 10)
    assert rel_fro(a, b).max() < (1e-13 if prec == "fp64" else 1e-6)


def test_facets_with_neighbor_types(dev):
    torch, _abi, ctx = dev
    mesh, lset = W.config2(8)
    from oracle import oracle as O
    dt, _, bits = O.mark(3, mesh.el2v, lset)
    f2v, el2f, f2el = mesh.facets()
    nf = f2el.shape[0]
    a = bits[4]  # IF
    b = bits[5]  # HASNEG
    out = torch.zeros((nf + 31) // 32, dtype=torch.int32, device="cuda")
    ctx.facets_with_neighbor_types(up(torch, f2el), up(torch, a.view(np.int32)), up(torch, b.view(np.int32)),
                                   True, False, True, out)
    ctx.sync()
    got = out.cpu().numpy().view(np.uint32)
    ea = (dt == IF)
    eb = (dt == IF) | (dt == NEG)
    l, r = f2el[:, 0], f2el[:, 1]
    al, bl = ea[l], eb[l]
    ar = np.where(r >= 0, ea[np.maximum(r, 0)], True)
    br = np.where(r >= 0, eb[np.maximum(r, 0)], False)
    ref = (al & br) | (ar & bl)
    refbits = np.zeros((nf + 31) // 32, np.uint32)
    for f in np.nonzero(ref)[0]:
        refbits[f >> 5] |= np.uint32(1 << (f & 31))
    assert np.array_equal(got, refbits)


@pytest.mark.parametrize("L", [0, 1, 2, 3, 4])
def test_subdiv_weights_parity(orc, dev, L):
    """GPU subdivision rule (contracted to P4 lattice weights) vs the FP64
    oracle: identical cut decisions (element domain), weights to 1e-12
    relative to the element's weight scale (different summation order)."""
    torch, _abi, ctx = dev
    nel = 256 if L <= 3 else 64
    vals = W.random_quadric_p2(nel, seed=11 + L)
    rng = np.random.default_rng(L)
    # random affine maps for the interface-weight transform
    Fs = np.eye(3)[None] + 0.3 * rng.standard_normal((nel, 3, 3))
    Finv = np.linalg.inv(Fs)
    ref_n = np.zeros((nel, 35))
    ref_i = np.zeros((nel, 35))
    ref_dom = np.zeros(nel, np.uint8)
    for e in range(nel):
        d, mn, mi = orc.subdiv_moments3(vals[e], 2, L, 4, 4, Finv=Finv[e])
        ref_n[e], ref_i[e], ref_dom[e] = mn, mi, d
    on = torch.zeros(nel * 35, dtype=torch.float64, device="cuda")
    oi = torch.zeros(nel * 35, dtype=torch.float64, device="cuda")
    dom = torch.zeros(nel, dtype=torch.uint8, device="cuda")
    ctx.subdiv_weights(up(torch, vals), 2, L, 4, 4, on, oi, dom, finv=up(torch, Finv.reshape(nel, 9)))
    ctx.sync()
    gn = on.cpu().numpy().reshape(nel, 35)
    gi = oi.cpu().numpy().reshape(nel, 35)
    assert np.array_equal(dom.cpu().numpy(), ref_dom)
    sn = np.abs(ref_n).max(axis=1, keepdims=True) + 1e-300
    si = np.abs(ref_i).max(axis=1, keepdims=True) + 1e-300
    assert (np.abs(gn - ref_n) / sn).max() < 1e-12
    assert (np.abs(gi - ref_i) / si).max() < 1e-12


def test_subdiv_p1_levelset_equals_straight_rule(orc, dev):
    """P1 nodal input: moments equal the straight cut rule's (any L)"""
    torch, _abi, ctx = dev
    rng = np.random.default_rng(5)
    l = rng.standard_normal((128, 4))
    l = l[(l < 0).any(1) & (l > 0).any(1)]
    nel = l.shape[0]
    on = torch.zeros(nel * 35, dtype=torch.float64, device="cuda")
    ctx.subdiv_weights(up(torch, l), 1, 2, 4, 4, on, None, None)
    ctx.sync()
    g = on.cpu().numpy().reshape(nel, 35)
    for e in range(nel):
        p, w, _ = orc.straight_cut_rule(3, l[e], NEG, 4)
        psi = np.array([orc.h1_shape(3, 4, x)[0] for x in p])
        np.testing.assert_allclose(g[e], psi.T @ w, rtol=0, atol=1e-13)


def _moments_from_rule(pts, w, mons):
    if len(w) == 0:
        return np.zeros(len(mons))
    lam = np.hstack([pts, 1.0 - pts.sum(axis=1, keepdims=True)])
    E = np.array(mons)
    return (np.prod(lam[:, None, :] ** E[None, :, :], axis=2) * w[:, None]).sum(axis=0)


@pytest.mark.parametrize("L,p", [(0, 2), (1, 1), (2, 3), (3, 2), (4, 2), (4, 3), (5, 1)])
def test_subdiv_moments_parity(orc, dev, L, p):
    """xg_subdiv_moments vs the moments of the oracle's materialised
    subdivision rule (or_subdiv_rule3): identical element domains, moments
    of degree 2p to 1e-12 relative to the element's largest moment."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "ngsxfem_amd", "csrc"))
    import gen_moments as G
    torch, _abi, ctx = dev
    mons = G.monomials(3, 2 * p)
    nel = {0: 64, 1: 64, 2: 48, 3: 32, 4: 16, 5: 4}[L]
    vals = W.random_quadric_p2(nel, seed=101 + L)
    rng = np.random.default_rng(L + 10 * p)
    Fs = np.eye(3)[None] + 0.3 * rng.standard_normal((nel, 3, 3))
    Finv = np.linalg.inv(Fs)
    nm = len(mons)
    on = torch.zeros(nel * nm, dtype=torch.float64, device="cuda")
    oi = torch.zeros(nel * nm, dtype=torch.float64, device="cuda")
    dom = torch.zeros(nel, dtype=torch.uint8, device="cuda")
    ctx.subdiv_moments(up(torch, vals), 2, L, 2 * p, p, on, oi, dom, finv=up(torch, Finv.reshape(nel, 9)))
    ctx.sync()
    gn = on.cpu().numpy().reshape(nel, nm)
    gi = oi.cpu().numpy().reshape(nel, nm)
    gd = dom.cpu().numpy()
    for e in range(nel):
        d, rules = orc.subdiv_rule3(vals[e], 2, L, 2 * p, Finv=Finv[e], cap=400000)
        assert gd[e] == d, e
        for dt, g in ((NEG, gn[e]), (IF, gi[e])):
            ref = _moments_from_rule(rules[dt][0], rules[dt][1], mons)
            scale = np.abs(ref).max() + 1e-300
            assert np.abs(g - ref).max() / scale < 1e-12, (e, dt)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_elmat_from_subdiv_moments(orc, dev, p):
    """subdivision rule (L = 2) -> degree-2p moments -> element matrices on
    physical tets, vs the oracle's direct quadrature on the materialised rule
    (FP64: 1e-12)"""
    torch, _abi, ctx = dev
    from math import comb
    nel = 24
    vals = W.random_quadric_p2(nel, seed=300 + p)
    rng = np.random.default_rng(p)
    V = np.zeros((nel, 4, 3))
    V[:, :3] = np.eye(3)[None] + 0.2 * rng.standard_normal((nel, 3, 3))
    V[:, 3] = 0.1 * rng.standard_normal((nel, 3))
    coords = V.reshape(-1, 3)
    el2v = np.arange(4 * nel, dtype=np.int32).reshape(nel, 4)
    F = np.stack([V[:, c] - V[:, 3] for c in range(3)], axis=2)  # column c = V_c - V_D
    Finv = np.linalg.inv(F)
    nm = comb(2 * p + 3, 3)
    mom = torch.zeros(nel * nm, dtype=torch.float64, device="cuda")
    ctx.subdiv_moments(up(torch, vals), 2, 2, 2 * p, p, mom, None, None, finv=up(torch, Finv.reshape(nel, 9)))
    nd = comb(p + 3, 3)
    out = torch.zeros(nel * nd * nd, dtype=torch.float64, device="cuda")
    ctx.elmat_moments(3, p, mom, 3, _abi.XG_PREC_FP64, out, el2v=up(torch, el2v), coords=up(torch, coords),
                      nel=nel)
    ctx.sync()
    g = out.cpu().numpy().reshape(nel, nd, nd)
    for e in range(nel):
        d, rules = orc.subdiv_rule3(vals[e], 2, 2, 2 * p, Finv=Finv[e])
        pts, w = rules[NEG]
        if len(w) == 0:
            assert np.all(g[e] == 0)
            continue
        ref = orc.elmat(3, p, V[e], np.ascontiguousarray(pts), w, mask=3)
        assert np.abs(g[e] - ref).max() <= 1e-12 * np.abs(ref).max(), e


@pytest.mark.parametrize("L", [0, 2, 4])
def test_mark_p2_matches_subdivision_domain(orc, dev, L):
    """P2 marking (Bernstein certificate + lattice check) == the domain of the
    oracle's subdivision rule, on cut, barely cut, uncut and zero-valued
    elements; bit arrays and cut list consistent with the domains"""
    torch, _abi, ctx = dev
    vals = W.random_quadric_p2(96, seed=500 + L)
    rng = np.random.default_rng(L)
    shifts = np.concatenate([np.zeros(32), rng.uniform(-3, 3, 32), rng.uniform(-0.05, 0.05, 32)])
    vals = vals + shifts[:, None]
    extra = np.array([np.zeros(10), np.r_[np.zeros(4), np.ones(6)], np.r_[np.ones(4), -np.ones(6) * 1e-3],
                      np.r_[-np.ones(9), 0.0]])
    vals = np.ascontiguousarray(np.vstack([vals, extra]))
    ne = vals.shape[0]
    ref = np.array([orc.subdiv_moments3(v, 2, L, 0, 4)[0] for v in vals], np.uint8)
    nw = (ne + 31) // 32
    dt = torch.zeros(ne, dtype=torch.uint8, device="cuda")
    bits = torch.zeros(8 * nw, dtype=torch.int32, device="cuda")
    cl = torch.full((ne,), -1, dtype=torch.int32, device="cuda")
    nc = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.mark_p2(up(torch, vals), L, dt, bits, cl, nc)
    ctx.sync()
    g = dt.cpu().numpy()
    assert np.array_equal(g, ref)
    n = int(nc.item())
    assert np.array_equal(cl.cpu().numpy()[:n], np.nonzero(ref == IF)[0])
    b = bits.cpu().numpy().view(np.uint32).reshape(8, nw)
    for c in range(8):
        want = np.array([(c == 7) or bool(c & (4 if d == IF else (1 if d == NEG else 2))) for d in ref])
        got = np.array([(b[c, e >> 5] >> (e & 31)) & 1 for e in range(ne)], bool)
        assert np.array_equal(got, want), c


@pytest.mark.parametrize("L", [0, 2, 4, 5])
def test_mark_p2_trig_matches_oracle(orc, dev, L):
    """P2 marking of boundary triangles (xg_mark_p2_trig) == the oracle's
    order-0 triangle subdivision rule: domains from the part areas bit for
    bit, cut ratios to 1e-12, bit arrays and cut list; cut, barely cut,
    uncut, zero-valued and one-sided lattice-cut elements, a ragged count"""
    torch, _abi, ctx = dev
    rng = np.random.default_rng(900 + L)
    vals = rng.normal(size=(150, 6))
    vals[50:100] += rng.uniform(-3, 3, (50, 1))
    vals[100:] = vals[100:] * 0.02 + rng.uniform(-0.01, 0.01, (50, 1))
    extra = np.array([np.zeros(6), np.r_[np.zeros(3), np.ones(3)], np.r_[-np.ones(5), 0.0],
                      np.r_[np.ones(5), -1e-17], np.r_[np.ones(3), -np.ones(3) * 1e-3]])
    vals = np.ascontiguousarray(np.vstack([vals, extra]))
    ne = vals.shape[0]
    ref_dt, ref_cr = np.zeros(ne, np.uint8), np.zeros(ne)
    for e in range(ne):
        _, an, ap = orc.subdiv_trig_parts(vals[e], 2, L)
        ref_dt[e] = (IF if ap > 0 else NEG) if an > 0 else POS
        ref_cr[e] = an / (an + ap)
    nw = (ne + 31) // 32
    dt = torch.zeros(ne, dtype=torch.uint8, device="cuda")
    cr = torch.zeros(ne, dtype=torch.float64, device="cuda")
    bits = torch.zeros(8 * nw, dtype=torch.int32, device="cuda")
    cl = torch.full((ne,), -1, dtype=torch.int32, device="cuda")
    nc = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.mark_p2_trig(up(torch, vals), L, dt, cr, bits, cl, nc)
    ctx.sync()
    g = dt.cpu().numpy()
    assert np.array_equal(g, ref_dt)
    assert (ref_dt == IF).sum() > 10 and (ref_dt == NEG).sum() > 0 and (ref_dt == POS).sum() > 0
    assert np.abs(cr.cpu().numpy() - ref_cr).max() <= 1e-12
    n = int(nc.item())
    assert np.array_equal(cl.cpu().numpy()[:n], np.nonzero(ref_dt == IF)[0])
    b = bits.cpu().numpy().view(np.uint32).reshape(8, nw)
    for c in range(8):
        want = np.array([(c == 7) or bool(c & (4 if d == IF else (1 if d == NEG else 2))) for d in ref_dt])
        got = np.array([(b[c, e >> 5] >> (e & 31)) & 1 for e in range(ne)], bool)
        assert np.array_equal(got, want), c
