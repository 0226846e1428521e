"""xg_mark_noalloc (marking with a fixed device scratch array, for callers
that cannot allocate) against xg_mark: every output bit for bit, on two calls
in a row (the second with fewer elements, so the array still holds the first
call's block counts), on an empty mesh, and the capacity check."""
import ctypes

import numpy as np
import pytest

from ngsxfem_amd import workloads as W
from ngsxfem_amd.meshes import MakeStructured3DMesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    from ngsxfem_amd import _abi
    ctx = _abi.Context(0)
    yield torch, _abi, ctx
    ctx.close()


def run(dev, fn, D, el2v, lset):
    torch, _abi, ctx = dev
    ne = el2v.shape[0]
    nw = (ne + 31) // 32
    dt = torch.full((ne,), 255, dtype=torch.uint8, device="cuda")
    cr = torch.full((ne,), -1.0, dtype=torch.float64, device="cuda")
    bits = torch.full((8 * nw,), -1, dtype=torch.int32, device="cuda")
    cl = torch.full((ne,), -7, dtype=torch.int32, device="cuda")
    nc = torch.full((1,), -3, dtype=torch.int32, device="cuda")
    fn(D, el2v, lset, dt, cr, bits, cl, nc)
    ctx.sync()
    n = int(nc.item())
    return dt.cpu().numpy(), cr.cpu().numpy(), bits.cpu().numpy(), cl.cpu().numpy()[:n], n


def test_noalloc_matches_mark_two_calls(dev):
    torch, _abi, ctx = dev
    sizes = [(48, 3, W.config2), (20, 3, W.config2)]   # 663,552 then 48,000 tets
    for n, D, make in sizes:
        mesh, lset = make(n)
        el2v = torch.from_numpy(mesh.el2v).cuda()
        ls = torch.from_numpy(lset).cuda()
        a = run(dev, ctx.mark, D, el2v, ls)
        b = run(dev, ctx.mark_noalloc, D, el2v, ls)
        assert a[4] == b[4] and a[4] > 0
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    # 2D, a size that is not a multiple of the block
    mesh, lset = W.config1()
    el2v = torch.from_numpy(mesh.el2v).cuda()
    ls = torch.from_numpy(lset).cuda()
    a = run(dev, ctx.mark, 2, el2v, ls)
    b = run(dev, ctx.mark_noalloc, 2, el2v, ls)
    assert a[4] == b[4] == 270
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_noalloc_empty_and_capacity(dev):
    torch, _abi, ctx = dev
    nc = torch.full((1,), 5, dtype=torch.int32, device="cuda")
    bits = torch.zeros(8, dtype=torch.int32, device="cuda")
    el = torch.zeros((1, 4), dtype=torch.int32, device="cuda")
    ls = torch.zeros(4, dtype=torch.float64, device="cuda")
    L = _abi.lib()
    assert L.xg_mark_noalloc(ctx.h, 3, 0, _abi.ptr(el), _abi.ptr(ls), None, None, _abi.ptr(bits), None,
                             _abi.ptr(nc)) == 0
    ctx.sync()
    assert int(nc.item()) == 0
    # more elements than the fixed scratch holds: rejected before any launch
    rc = L.xg_mark_noalloc(ctx.h, 3, (1 << 28) + 1, _abi.ptr(el), _abi.ptr(ls), None, None, _abi.ptr(bits), None,
                           _abi.ptr(nc))
    assert rc == -4
    assert b"XG_MARK_NOALLOC_MAX_NE" in L.xg_last_error()
