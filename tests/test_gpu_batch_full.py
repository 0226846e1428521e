"""The batched uneven workload at its full size (256 cube meshes, n in
[8, 48], 32.6 M tets, p in {1, 2, 3}) through xg_batch, in one captured
graph:

* the cut count of every mesh equals the single-mesh GPU marking's;
* for two meshes of each degree the batch's element matrices equal the
  single-mesh fused pipeline's bit for bit (MFMA and moment methods);
* 64 cut elements of each of those meshes against the FP64 oracle, 1e-6
  (Frobenius, per element; the FP32 store's bar).
"""
import numpy as np
import pytest

from ngsxfem_amd import workloads as W

pytestmark = pytest.mark.gpu
NEG, IF = 0, 2


@pytest.fixture(scope="module")
def full():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import _abi
    coords, el2v, lset, off, ps, ns, axes = W.config3(256)
    print("batched uneven: %d meshes, %d tets" % (len(ps), el2v.shape[0]), flush=True)
    ctx = _abi.Context(0)
    d = dict(x=torch.from_numpy(coords).cuda(), e=torch.from_numpy(el2v).cuda(), l=torch.from_numpy(lset).cuda())
    b = _abi.Batch(ctx, d["x"], d["e"], d["l"], off, ps)
    yield torch, _abi, ctx, b, (coords, el2v, lset, off, ps), d
    b.close()
    ctx.close()


def mesh_cut_list(torch, ctx, d, a0, a1):
    """single-mesh GPU marking of the elements a0..a1 (global vertex ids)"""
    ne = a1 - a0
    nw = (ne + 31) // 32
    dt = torch.zeros(ne, dtype=torch.uint8, device="cuda")
    cr = torch.zeros(ne, dtype=torch.float64, device="cuda")
    bits = torch.zeros(8 * nw, dtype=torch.int32, device="cuda")
    cl = torch.zeros(ne, dtype=torch.int32, device="cuda")
    nc = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.mark(3, d["e"][a0:a1], d["l"], dt, cr, bits, cl, nc)
    ctx.sync()
    return cl[:int(nc.item())].cpu().numpy() + a0


@pytest.mark.parametrize("method", ["mfma", "moments"])
def test_full_batch(orc, full, method):
    torch, _abi, ctx, b, (coords, el2v, lset, off, ps), d = full
    b.method(_abi.XG_METHOD_MOMENTS if method == "moments" else 0)
    flags = _abi.XG_PREC_FP32 | (_abi.XG_METHOD_MOMENTS if method == "moments" else 0)
    cnt, oo = b.plan()
    out = torch.full((int(oo[-1]) + 7,), -1.0, dtype=torch.float32, device="cuda")
    b.run(out, use_graph=True)
    ctx.sync()
    got = out.cpu().numpy()
    assert got[oo[-1]:].tolist() == [-1.0] * 7
    print("%s: %d cut elements" % (method, int(np.sum(cnt))), flush=True)
    lists = {}
    for m in range(len(ps)):
        cl = mesh_cut_list(torch, ctx, d, int(off[m]), int(off[m + 1]))
        assert cnt[m] == len(cl), m
        lists[m] = cl
    rng = np.random.default_rng(11)
    for p in (1, 2, 3):
        ms = [m for m in range(len(ps)) if ps[m] == p and cnt[m] > 0]
        for m in rng.choice(ms, 2, replace=False):
            cl = lists[m].astype(np.int32)
            nd = {1: 4, 2: 10, 3: 20}[p]
            single = torch.zeros(len(cl) * 2 * nd * nd, dtype=torch.float32, device="cuda")
            ctx.cut_pipeline(3, p, 2 * p, d["e"], d["x"], d["l"], [NEG, IF], 3, flags, single,
                             elems=torch.from_numpy(cl).cuda())
            ctx.sync()
            seg = got[oo[m]:oo[m + 1]]
            assert np.array_equal(seg, single.cpu().numpy()), "mesh %d differs from the single-mesh pipeline" % m
            pick = np.sort(rng.choice(len(cl), min(64, len(cl)), replace=False))
            ref = orc.cut_pipeline(3, p, 2 * p, cl[pick], el2v, coords, lset, [NEG, IF], mask=3, nthreads=8)
            g = seg.reshape(len(cl), 2, nd, nd)[pick].astype(np.float64)
            nz = np.sqrt((ref ** 2).sum(axis=(-2, -1))) > 0
            err = np.sqrt(((g - ref) ** 2).sum(axis=(-2, -1)))[nz] / np.sqrt((ref ** 2).sum(axis=(-2, -1)))[nz]
            assert err.max() < 1e-6, (m, p, err.max())
