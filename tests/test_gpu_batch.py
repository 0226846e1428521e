"""Batched multi-mesh pipeline (xg_batch) on the GPU.

* cut counts per mesh equal the oracle's marking of each mesh;
* every element matrix of the batch equals the single-mesh fused pipeline's
  (same kernels: bit for bit) and the FP64 oracle's to 1e-6;
* the captured hipGraph replays give the same result as the eager run,
  with eager work and synchronisations between the replays;
* every autotuner launch shape gives the same result.
"""
import numpy as np
import pytest

from ngsxfem_amd import workloads as W

pytestmark = pytest.mark.gpu
NEG, IF = 0, 2


@pytest.fixture(scope="module")
def setup():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import _abi
    ctx = _abi.Context(0)
    coords, el2v, lset, off, ps, ns, axes = W.config3(nmesh=10, seed0=77, nmin=3, nmax=9)
    d = dict(x=torch.from_numpy(coords).cuda(), e=torch.from_numpy(el2v).cuda(), l=torch.from_numpy(lset).cuda())
    b = _abi.Batch(ctx, d["x"], d["e"], d["l"], off, ps)
    yield torch, _abi, ctx, b, (coords, el2v, lset, off, ps), d
    b.close()
    ctx.close()


def test_batch_counts_and_matrices(orc, setup):
    torch, _abi, ctx, b, (coords, el2v, lset, off, ps), d = setup
    cnt, oo = b.plan()
    out = torch.full((int(oo[-1]) + 7,), -1.0, dtype=torch.float32, device="cuda")
    b.run(out, use_graph=False)
    ctx.sync()
    got = out.cpu().numpy()
    for m in range(len(ps)):
        a0, a1 = off[m], off[m + 1]
        dt, cr, bits = orc.mark(3, el2v[a0:a1] - el2v[a0:a1].min(), lset[el2v[a0:a1].min():el2v[a0:a1].max() + 1])
        assert cnt[m] == int(np.sum(dt == IF)), m
        cut_m = (np.nonzero(dt == IF)[0] + a0).astype(np.int32)
        p = int(ps[m])
        nd = {1: 4, 2: 10, 3: 20}[p]
        if cnt[m] == 0:
            continue
        ref_gpu = torch.zeros(len(cut_m) * 2 * nd * nd, dtype=torch.float32, device="cuda")
        ctx.cut_pipeline(3, p, 2 * p, d["e"], d["x"], d["l"], [NEG, IF], 3, _abi.XG_PREC_FP32, ref_gpu,
                         elems=torch.from_numpy(cut_m).cuda())
        ctx.sync()
        seg = got[oo[m]:oo[m + 1]]
        assert np.array_equal(seg, ref_gpu.cpu().numpy()), "mesh %d differs from the single-mesh pipeline" % m
        ref = orc.cut_pipeline(3, p, 2 * p, cut_m, el2v, coords, lset, [NEG, IF], mask=3, nthreads=4)
        g = seg.astype(np.float64).reshape(ref.shape)
        nz = np.sqrt((ref ** 2).sum(axis=(-2, -1))) > 0
        err = np.sqrt(((g - ref) ** 2).sum(axis=(-2, -1)))[nz] / np.sqrt((ref ** 2).sum(axis=(-2, -1)))[nz]
        # FP32 MFMA products, FP64 sums of 8-point groups: 1e-6 for every degree
        assert err.max() < 1e-6, (m, p, err.max())
    assert got[oo[-1]:].tolist() == [-1.0] * 7  # nothing written past the plan


def test_batch_graph_replays(setup):
    torch, _abi, ctx, b, host, d = setup
    cnt, oo = b.plan()
    n = int(oo[-1])
    eager = torch.zeros(n, dtype=torch.float32, device="cuda")
    b.run(eager, use_graph=False)
    ctx.sync()
    ref = eager.clone()
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    for k in range(4):
        out.zero_()
        b.run(out, use_graph=True)
        ctx.sync()
        assert torch.equal(out, ref), "replay %d" % k
        # unrelated eager work between replays
        tmp = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 8, 4), (1, 4, 2)])
def test_batch_tuning_shapes_agree(setup, shape):
    torch, _abi, ctx, b, host, d = setup
    cnt, oo = b.plan()
    n = int(oo[-1])
    ref = torch.zeros(n, dtype=torch.float32, device="cuda")
    b.run(ref, use_graph=False)
    ctx.sync()
    for p in (1, 2, 3):
        b.tune(p, *shape)
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    b.run(out, use_graph=True)
    ctx.sync()
    assert torch.equal(out, ref)
    for p in (1, 2, 3):
        b.tune(p, 2, 4, 1)


def test_batch_capacity_error(setup):
    torch, _abi, ctx, b, host, d = setup
    cnt, oo = b.plan()
    small = torch.zeros(max(int(oo[-1]) // 2, 1), dtype=torch.float32, device="cuda")
    b.run(small, use_graph=False)
    with pytest.raises(_abi.XgError):
        ctx.sync()


def test_batch_moments_match_single_mesh(orc, setup):
    """moment form in the batch: bit-equal to the single-mesh moment kernel
    (same device function) and within 2e-7 of the FP64 oracle"""
    torch, _abi, ctx, b, (coords, el2v, lset, off, ps), d = setup
    b.method(_abi.XG_METHOD_MOMENTS)
    try:
        cnt, oo = b.plan()
        out = torch.full((int(oo[-1]),), -1.0, dtype=torch.float32, device="cuda")
        b.run(out, use_graph=True)
        ctx.sync()
        got = out.cpu().numpy()
        for m in range(len(ps)):
            if cnt[m] == 0:
                continue
            a0 = off[m]
            dt, _, _ = orc.mark(3, el2v[a0:off[m + 1]] - el2v[a0:off[m + 1]].min(),
                                lset[el2v[a0:off[m + 1]].min():el2v[a0:off[m + 1]].max() + 1])
            cut_m = (np.nonzero(dt == IF)[0] + a0).astype(np.int32)
            p = int(ps[m])
            nd = {1: 4, 2: 10, 3: 20}[p]
            single = torch.zeros(len(cut_m) * 2 * nd * nd, dtype=torch.float32, device="cuda")
            ctx.cut_pipeline(3, p, 2 * p, d["e"], d["x"], d["l"], [NEG, IF], 3,
                             _abi.XG_PREC_FP32 | _abi.XG_METHOD_MOMENTS, single, elems=torch.from_numpy(cut_m).cuda())
            ctx.sync()
            seg = got[oo[m]:oo[m + 1]]
            assert np.array_equal(seg, single.cpu().numpy()), m
            ref = orc.cut_pipeline(3, p, 2 * p, cut_m, el2v, coords, lset, [NEG, IF], mask=3, nthreads=4)
            g = seg.astype(np.float64).reshape(ref.shape)
            nz = np.sqrt((ref ** 2).sum(axis=(-2, -1))) > 0
            err = np.sqrt(((g - ref) ** 2).sum(axis=(-2, -1)))[nz] / np.sqrt((ref ** 2).sum(axis=(-2, -1)))[nz]
            assert err.max() < 2e-7, (m, p, err.max())
        for shape in [(1, 1, 1), (2, 8, 4), (1, 4, 2)]:
            for p in (1, 2, 3):
                b.tune(p, *shape)
            o2 = torch.zeros_like(out)
            b.run(o2, use_graph=True)
            ctx.sync()
            assert torch.equal(o2, out), shape
    finally:
        for p in (1, 2, 3):
            b.tune(p, 2, 4, 1)
        b.method(0)
