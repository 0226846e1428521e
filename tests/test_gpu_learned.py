"""Learned quadrature and fixed-node element matrices on the GPU, through the
C ABI, checked against the FP64 oracle.

* MLP gradients of the library's FP32 backward kernels vs the oracle's FP64
  backward (itself pinned by finite differences): 1e-4 relative.
* BF16 MFMA inference vs the FP32 forward: BF16 rounding of activations and
  weights, 3e-2 relative (Frobenius) -- also validates the MFMA lane layout.
* The trained correction network brings the level-1 subdivision moments
  closer to the level-4 ones than level 1 alone (> 2x) and than an untrained
  network (> 10x); the degree-2 check passes the plain rule itself to 1e-12;
  threshold 0 makes every element fall back to the level-4 rule.
* Element matrices from fixed-node weights equal those on the materialised
  subdivision rule (exactness of the degree-4 lattice for P2 integrands).
"""
import numpy as np
import pytest

from ngsxfem_amd import workloads as W

pytestmark = pytest.mark.gpu
NEG, POS, IF = 0, 1, 2


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from ngsxfem_amd import _abi
    ctx = _abi.Context(0)
    yield torch, _abi, ctx
    ctx.close()


def test_mlp_gradient_vs_oracle(orc, dev):
    torch, _abi, ctx = dev
    I, H, O, B = 10, 128, 35, 64
    rng = np.random.default_rng(0)
    npar = orc.mlp_nparams(I, H, O)
    P = (0.2 * rng.standard_normal(npar)).astype(np.float32)
    x = rng.standard_normal((B, I)).astype(np.float32)
    t = rng.standard_normal((B, O)).astype(np.float32)
    loss_ref, g_ref = orc.mlp_loss_grad(I, H, O, P.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    dP = torch.from_numpy(P).cuda()
    g = torch.zeros(npar, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    work = torch.zeros(4 * B * H + 2 * B * O, dtype=torch.float32, device="cuda")
    dx = torch.from_numpy(x).cuda()  # keep the device tensors alive across the call
    dt = torch.from_numpy(t).cuda()
    _abi.check(_abi.lib().xg_mlp_train_step(ctx.h, I, H, O, _abi.ptr(dP), None, None, 1, 0.0, B,
                                            _abi.ptr(dx), _abi.ptr(dt), _abi.ptr(loss), _abi.ptr(g),
                                            _abi.ptr(work)))
    ctx.sync()
    gg = g.cpu().numpy().astype(np.float64)
    assert abs(loss.item() - loss_ref) < 1e-5 * abs(loss_ref)
    assert np.linalg.norm(gg - g_ref) < 1e-4 * np.linalg.norm(g_ref)
    # parameters untouched when lr = 0
    assert np.array_equal(dP.cpu().numpy(), P)


@pytest.mark.parametrize("O", [35, 84])
def test_mlp_forward_fp32_and_bf16(orc, dev, O):
    """FP32 and BF16 inference vs the FP64 oracle, for the P2 (35) and P3 (84)
    moment corrections"""
    torch, _abi, ctx = dev
    I, H, B = 10, 128, 1000
    rng = np.random.default_rng(1)
    npar = orc.mlp_nparams(I, H, O)
    P = (0.3 * rng.standard_normal(npar)).astype(np.float32)
    x = rng.uniform(-1, 1, (B, I)).astype(np.float32)
    ref = orc.mlp_forward(I, H, O, P.astype(np.float64), x.astype(np.float64))
    dP = torch.from_numpy(P).cuda()
    dx = torch.from_numpy(x).cuda()
    y32 = torch.zeros(B * O, dtype=torch.float32, device="cuda")
    y16 = torch.zeros(B * O, dtype=torch.float32, device="cuda")
    work = torch.zeros(2 * B * H, dtype=torch.float32, device="cuda")
    lib = _abi.lib()
    _abi.check(lib.xg_mlp_forward(ctx.h, I, H, O, _abi.ptr(dP), B, _abi.ptr(dx), _abi.ptr(y32), _abi.ptr(work),
                                  _abi.XG_PREC_FP32))
    _abi.check(lib.xg_mlp_forward(ctx.h, I, H, O, _abi.ptr(dP), B, _abi.ptr(dx), _abi.ptr(y16), None,
                                  _abi.XG_PREC_BF16))
    ctx.sync()
    a32 = y32.cpu().numpy().reshape(B, O)
    a16 = y16.cpu().numpy().reshape(B, O)
    assert np.abs(a32 - ref).max() < 1e-4 * max(1, np.abs(ref).max())
    assert np.linalg.norm(a16 - ref) < 3e-2 * np.linalg.norm(ref)


def test_learned_rule_gain(orc, dev):
    """the correction network, trained on the GPU, brings the level-1
    subdivision moments closer to the level-4 ones than the level-1 rule
    alone (and far closer than an untrained network); the degree-2 check is
    exact for the plain rule itself; threshold 0 sends every element to the
    plain fallback, an infinite one accepts all"""
    torch, _abi, ctx = dev
    from ngsxfem_amd.learned import LearnedRule, random_cut_quadrics
    lr_ = LearnedRule(ctx, p=2, base_L=1, L=4, seed=0)
    untrained = LearnedRule(ctx, p=2, base_L=1, L=4, seed=0)
    untrained.fold()
    lr_.train(n_train=40000, epochs=8, batch=512, lr=2e-3, seed=3)
    xv = random_cut_quadrics(4000, 99)
    m4 = lr_.plain_moments(xv, 4)
    m1 = lr_.plain_moments(xv, 1)
    c = lr_.correction(xv).to(torch.float64)
    cu = untrained.correction(xv).to(torch.float64)
    ctx.sync()

    def err(a):
        return float(torch.linalg.norm(a - m4) / torch.linalg.norm(m4))
    e_nn, e_un, e_l1 = err(m1 + c), err(m1 + cu), err(m1)
    print("rel err vs L4: corrected %.3e untrained %.3e plain L1 %.3e" % (e_nn, e_un, e_l1))
    # measured: corrected 5.9e-2, level 1 alone 1.7e-1, untrained 79
    assert e_nn < e_l1 / 2
    assert e_nn < e_un / 10
    # the check lowers degree-4 moments to degree 2 through sum lam = 1: the
    # plain level-4 moments pass it against the level-4 degree-2 moments
    n = xv.shape[0]
    ref2 = lr_.plain_moments(xv, 4, p=1)
    out = torch.empty_like(m4)
    rel = torch.empty(n, dtype=torch.float64, device="cuda")
    acc = torch.empty(n, dtype=torch.uint8, device="cuda")
    zero = torch.zeros(n, 35, dtype=torch.float32, device="cuda")
    _abi.check(_abi.lib().xg_learned_moments(ctx.h, n, 2, _abi.ptr(m4), _abi.ptr(zero), _abi.ptr(ref2), 1e-12,
                                             _abi.ptr(out), _abi.ptr(rel), _abi.ptr(acc)))
    ctx.sync()
    assert float(rel.max()) < 1e-12 and int(acc.sum()) == n
    assert torch.equal(out, m4)
    # fallback extremes through LearnedRule.moments
    out0, acc0, _ = lr_.moments(xv, thr=0.0)
    ctx.sync()
    assert int(acc0.sum().item()) == 0 and torch.equal(out0, m4)
    out1, acc1, _ = lr_.moments(xv, thr=float("inf"))
    ctx.sync()
    assert int(acc1.sum().item()) == n
    assert torch.allclose(out1, m1 + lr_.correction(xv).to(torch.float64), rtol=0, atol=1e-14)


def test_learned_moments_elmat(orc, dev):
    """element matrices from learned-with-fallback moments at threshold 0
    are the plain level-4 matrices (every element falls back)"""
    torch, _abi, ctx = dev
    from ngsxfem_amd.learned import LearnedRule, random_cut_quadrics
    lr_ = LearnedRule(ctx, p=2, base_L=1, L=4, seed=1)
    lr_.fold()
    xv = random_cut_quadrics(512, 7)
    m, acc, rel = lr_.moments(xv, thr=0.0)
    a = torch.empty(512 * 100, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    ctx.elmat_moments(3, 2, m, 3, _abi.XG_PREC_FP64, a, nel=512)
    ctx.elmat_moments(3, 2, lr_.plain_moments(xv, 4), 3, _abi.XG_PREC_FP64, b, nel=512)
    ctx.sync()
    assert torch.equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fixed_node_elmat_matches_materialised_rule(orc, dev, prec):
    torch, _abi, ctx = dev
    vals = W.random_quadric_p2(24, seed=5)
    n = vals.shape[0]
    wn = torch.zeros(n * 35, dtype=torch.float64, device="cuda")
    ctx.subdiv_weights(torch.from_numpy(vals).cuda(), 2, 2, 4, 4, wn, None, None)
    P = _abi.XG_PREC_FP64 if prec == "fp64" else _abi.XG_PREC_FP32
    out = torch.zeros(n * 100, dtype=torch.float64 if prec == "fp64" else torch.float32, device="cuda")
    _abi.check(_abi.lib().xg_elmat_weights(ctx.h, 2, 4, n, None, None, None, _abi.ptr(wn), 3, P, _abi.ptr(out)))
    ctx.sync()
    g = out.cpu().numpy().astype(np.float64).reshape(n, 10, 10)
    Vref = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]], float)
    for e in range(n):
        dom, rules = orc.subdiv_rule3(vals[e], 2, 2, 4)
        p, w = rules[NEG]
        ref = orc.elmat(3, 2, Vref, p, w, mask=3)
        err = np.linalg.norm(g[e] - ref) / np.linalg.norm(ref)
        assert err < (1e-12 if prec == "fp64" else 1e-6), (e, err)
