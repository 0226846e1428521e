"""Moment-form element matrices (xg_cut_pipeline | XG_METHOD_MOMENTS) vs the
FP64 oracle's quadrature sum (oracle/xg_oracle.c or_cut_pipeline, which
follows symboliccutbfi.cpp:229-287 point by point).

The moment form accumulates sum_i w_i lam(x_i)^alpha in FP64 over the same
cut rule and applies exact polynomial identities, so it must reproduce the
quadrature sum for ANY rule order (also orders below 2p, where the rule is
not exact and a closed-form integral would differ).  Bars: FP64 output 1e-12
relative (Frobenius norm per element matrix); FP32 output 2e-7 (one FP32
rounding of each entry on top of the FP64 result).
"""
import numpy as np
import pytest

from ngsxfem_amd import workloads as W
from ngsxfem_amd.meshes import Make1DMesh

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


def up(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel_fro(a, b):
    num = np.sqrt(np.sum((a - b) ** 2, axis=(-2, -1)))
    den = np.sqrt(np.sum(b ** 2, axis=(-2, -1)))
    return num / np.maximum(den, 1e-300)


def mesh_for(D):
    if D == 1:
        mesh = Make1DMesh(40, lambda v: 2.0 * v - 1.0)
        lset = np.abs(mesh.coords[:, 0]) - 0.37
        return mesh, lset
    if D == 2:
        return W.config1()
    return W.config2(8)


def run_moments(dev, D, p, order, mesh, lset, elems, dts, mask, fp64):
    torch, _abi, ctx = dev
    nd = {1: p + 1, 2: (p + 1) * (p + 2) // 2, 3: (p + 1) * (p + 2) * (p + 3) // 6}[D]
    prec = (_abi.XG_PREC_FP64 if fp64 else _abi.XG_PREC_FP32) | _abi.XG_METHOD_MOMENTS
    out = torch.full((len(elems) * len(dts) * nd * nd,), float("nan"),
                     dtype=torch.float64 if fp64 else torch.float32, device="cuda")
    ctx.cut_pipeline(D, p, order, up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, lset), dts, mask, prec,
                     out, elems=up(torch, elems))
    ctx.sync()
    return out.cpu().numpy().astype(np.float64).reshape(len(elems), len(dts), nd, nd)


def elements(orc, D, mesh, lset, nuncut=48):
    dto, _, _ = orc.mark(D, mesh.el2v, lset)
    cut = np.nonzero(dto == IF)[0]
    neg = np.nonzero(dto == NEG)[0][:nuncut]
    pos = np.nonzero(dto == POS)[0][:nuncut]
    return np.concatenate([cut, neg, pos]).astype(np.int32)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("fp64", [True, False])
def test_moments_match_quadrature(orc, dev, D, p, fp64):
    mesh, lset = mesh_for(D)
    elems = elements(orc, D, mesh, lset)
    dts = [NEG, IF, POS]
    order = 2 * p
    ref = orc.cut_pipeline(D, p, order, elems, mesh.el2v, mesh.coords, lset, dts, mask=3, nthreads=8)
    g = run_moments(dev, D, p, order, mesh, lset, elems, dts, 3, fp64)
    nz = np.sqrt(np.sum(ref ** 2, axis=(-2, -1))) > 0
    assert np.all(g[~nz] == 0), "nonzero matrix where the rule is empty"
    err = rel_fro(g[nz], ref[nz])
    assert err.max() < (1e-12 if fp64 else 2e-7), "max rel err %g" % err.max()


@pytest.mark.parametrize("order", [0, 1, 2, 3, 5, 8])
@pytest.mark.parametrize("mask", [1, 2, 3])
def test_moments_any_rule_order(orc, dev, order, mask):
    """below 2p the rule is inexact: the moment form must still equal the
    quadrature sum of that rule, not the exact integral"""
    D, p = 3, 2
    mesh, lset = mesh_for(D)
    elems = elements(orc, D, mesh, lset, nuncut=8)
    dts = [NEG, IF]
    ref = orc.cut_pipeline(D, p, order, elems, mesh.el2v, mesh.coords, lset, dts, mask=mask, nthreads=8)
    g = run_moments(dev, D, p, order, mesh, lset, elems, dts, mask, True)
    nz = np.sqrt(np.sum(ref ** 2, axis=(-2, -1))) > 0
    assert np.all(g[~nz] == 0)
    assert rel_fro(g[nz], ref[nz]).max() < 1e-12


def test_moments_match_mfma_pipeline(orc, dev):
    """the moment form and the MFMA contraction (FP64) compute the same sum"""
    torch, _abi, ctx = dev
    D, p, order = 3, 2, 4
    mesh, lset = mesh_for(D)
    elems = elements(orc, D, mesh, lset)
    dts = [NEG, IF]
    g = run_moments(dev, D, p, order, mesh, lset, elems, dts, 3, True)
    out = torch.zeros(g.size, dtype=torch.float64, device="cuda")
    ctx.cut_pipeline(D, p, order, up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, lset), dts, 3,
                     _abi.XG_PREC_FP64, out, elems=up(torch, elems))
    ctx.sync()
    h = out.cpu().numpy().reshape(g.shape)
    nz = np.sqrt(np.sum(h ** 2, axis=(-2, -1))) > 0
    assert rel_fro(g[nz], h[nz]).max() < 1e-12


def test_moments_full_config2_properties(dev):
    """64^3 Kuhn mesh (1.57M tets, 43k cut): symmetry, K 1 = 0, mass sum =
    cut volume, and NEG + POS = the uncut element matrix"""
    torch, _abi, ctx = dev
    mesh, lset = W.config2(64)
    D, p, order = 3, 2, 4
    ne = mesh.ne
    nw = (ne + 31) // 32
    d_el, d_x, d_l = up(torch, mesh.el2v), up(torch, mesh.coords), up(torch, lset)
    dt = torch.zeros(ne, dtype=torch.uint8, device="cuda")
    cl = torch.zeros(ne, dtype=torch.int32, device="cuda")
    nc = torch.zeros(1, dtype=torch.int32, device="cuda")
    bits = torch.zeros(8 * nw, dtype=torch.int32, device="cuda")
    ctx.mark(D, d_el, d_l, dt, None, bits, cl, nc)
    ctx.sync()
    n = int(nc.item())
    assert n > 40000
    cut = cl[:n].contiguous()
    prec = _abi.XG_PREC_FP64 | _abi.XG_METHOD_MOMENTS
    out = torch.zeros(n * 3 * 100, dtype=torch.float64, device="cuda")
    ctx.cut_pipeline(D, p, order, d_el, d_x, d_l, [NEG, POS, IF], 3, prec, out, elems=cut)
    mass = torch.zeros(n * 1 * 100, dtype=torch.float64, device="cuda")
    ctx.cut_pipeline(D, p, order, d_el, d_x, d_l, [NEG], 1, prec, mass, elems=cut)
    ctx.sync()
    # the same elements with the level set made positive: the uncut matrix
    one = torch.ones_like(d_l)
    full = torch.zeros(n * 100, dtype=torch.float64, device="cuda")
    ctx.cut_pipeline(D, p, order, d_el, d_x, one, [POS], 3, prec, full, elems=cut)
    ctx.sync()
    A = out.cpu().numpy().reshape(n, 3, 10, 10)
    Mn = mass.cpu().numpy().reshape(n, 10, 10)
    F = full.cpu().numpy().reshape(n, 10, 10)
    assert np.abs(A - A.transpose(0, 1, 3, 2)).max() <= 1e-15 * np.abs(A).max()
    scale = np.abs(F).max()
    assert np.abs(A[:, 0] + A[:, 1] - F).max() < 1e-12 * scale
    # the mass matrix sums to the NEG volume (partition of unity); with the
    # stiffness part, the row sums equal those of the mass part (K 1 = 0)
    vol = Mn.sum(axis=(1, 2))
    assert np.all(vol >= -1e-18)
    rs = A[:, 0].sum(axis=2) - Mn.sum(axis=2)
    assert np.abs(rs).max() < 1e-12 * scale
