"""Learned quadrature for P2-cut tetrahedra (host side).

The plain rule of a P2 level set is the subdivision rule at level L
(cutint/xintegration.cpp:1626-1754, 617-1063), contracted on the GPU into the
degree-2p barycentric moments m_alpha = sum_points w lam^alpha, from which
every degree-p mass and stiffness matrix follows exactly (xg_elmat_moments).

The learned rule corrects a cheap plain rule instead of replacing it: a small
MLP (10 -> 128 -> 128 -> NM, tanh) maps the normalised P2 nodal values of a
cut tet to the difference between the level-L moments and the moments of the
same subdivision rule at a low level (base_L, a few percent of the cost).

* Labels: both plain rules on the GPU (xg_subdiv_moments).
* Training: the library's FP32 forward / backward / Adam kernels
  (xg_mlp_train_step); inference: the fused BF16 MFMA kernel (xg_mlp_forward,
  FP32 accumulation).  The output standardisation is folded into the last
  layer after training, so the raw output is the correction.
* Check without the full plain rule: the learned moments are lowered through
  sum lam = 1 to degree-2 moments and compared with the exact degree-2
  moments of the level-L rule (xg_subdiv_moments with p = 1: the same tree,
  4-point instead of 14 / 24-point rules) in xg_learned_moments; elements
  above the threshold fall back to the level-L rule itself.  The plain rule
  also judges the learned rule's error (bench and tests compute both).
"""
import numpy as np

from . import _abi

I_IN, H = 10, 128
NMOM = {1: 10, 2: 35, 3: 84}


def normalise_inputs(lset_nodal):
    """phi -> phi / max|phi| per element (the cut region is scale invariant)"""
    import torch
    m = lset_nodal.abs().amax(dim=1, keepdim=True).clamp_min(1e-300)
    return (lset_nodal / m).to(torch.float32).contiguous()


def random_cut_quadrics(n, seed, device="cuda"):
    """C4 distribution: x^T A x + b^T x + c on the reference tet, A symmetric
    N(0,1), b N(0,1), c shifted by minus the median nodal value so the tet is
    cut.  Returns P2 nodal values (n x 10, FP64)."""
    import torch
    from .workloads import p2_nodes
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = torch.randn(n, 3, 3, generator=g, dtype=torch.float64)
    A = 0.5 * (A + A.transpose(1, 2))
    b = torch.randn(n, 3, generator=g, dtype=torch.float64)
    X = torch.from_numpy(p2_nodes())
    vals = torch.einsum("ki,nij,kj->nk", X, A, X) + b @ X.T
    c = -vals.median(dim=1).values
    return (vals + c[:, None]).to(device).contiguous()


class LearnedRule:
    """MLP correction of the level-base_L subdivision moments towards level L
    (NEG part of P2-cut reference tets; degree-2p moments)."""

    def __init__(self, ctx, p=2, base_L=1, L=4, seed=0, device="cuda"):
        import torch
        self.ctx = ctx
        self.device = device
        self.p, self.base_L, self.L = p, base_L, L
        self.O = NMOM[p]
        npar = _abi.lib().xg_mlp_nparams(I_IN, H, self.O)
        g = torch.Generator(device="cpu").manual_seed(seed)
        prm = torch.zeros(npar, dtype=torch.float32)
        o = 0
        for fan_in, n_w, n_b in ((I_IN, H * I_IN, H), (H, H * H, H), (H, self.O * H, self.O)):
            bound = 1.0 / np.sqrt(fan_in)
            prm[o:o + n_w] = (torch.rand(n_w, generator=g) * 2 - 1) * bound
            o += n_w + n_b
        self.params = prm.to(device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.mean = torch.zeros(self.O, dtype=torch.float32, device=device)
        self.std = torch.ones(self.O, dtype=torch.float32, device=device)
        self.folded = False

    # ------------------------------------------------------------- plain rules
    def plain_moments(self, lset_nodal, L, p=None, out=None):
        """NEG moments (n x NM FP64) of the level-L subdivision rule with exact
        base rules (order 2p), on the GPU"""
        import torch
        p = self.p if p is None else p
        n = lset_nodal.shape[0]
        m = out if out is not None else torch.empty(n, NMOM[p], dtype=torch.float64, device=self.device)
        self.ctx.subdiv_moments(lset_nodal, 2, L, 2 * p, p, m, None, None)
        return m

    # ------------------------------------------------------------- training
    def train(self, n_train=200000, epochs=20, batch=1024, lr=1e-3, seed=1, log=None, x_raw=None):
        """Adam on the FP32 kernels; samples: random cut quadrics (n_train), or
        the given P2 nodal values x_raw (n x 10 FP64, device)"""
        import torch
        if x_raw is None:
            x_raw = random_cut_quadrics(n_train, seed, self.device)
        else:
            n_train = x_raw.shape[0]
        full = self.plain_moments(x_raw, self.L)
        base = self.plain_moments(x_raw, self.base_L)
        self.ctx.sync()
        x = normalise_inputs(x_raw)
        y = (full - base).to(torch.float32)
        self.mean = y.mean(dim=0)
        self.std = y.std(dim=0).clamp_min(1e-20)
        t = ((y - self.mean) / self.std).contiguous()
        work = torch.empty(4 * batch * H + 2 * batch * self.O, dtype=torch.float32, device=self.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        nb = n_train // batch
        g = torch.Generator(device="cpu").manual_seed(seed + 7)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n_train, generator=g).to(self.device)
            lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)) * 0.9 + 0.1)
            tot = 0.0
            for b in range(nb):
                idx = perm[b * batch:(b + 1) * batch]
                xb = x[idx].contiguous()
                tb = t[idx].contiguous()
                self.step += 1
                _abi.check(_abi.lib().xg_mlp_train_step(
                    self.ctx.h, I_IN, H, self.O, _abi.ptr(self.params), _abi.ptr(self.m), _abi.ptr(self.v),
                    self.step, float(lr_ep), batch, _abi.ptr(xb), _abi.ptr(tb), _abi.ptr(loss), None,
                    _abi.ptr(work)))
                if b % 50 == 0:
                    self.ctx.sync()
                    tot = float(loss.item())
            hist.append(tot)
            if log:
                log("epoch %d loss %.3e" % (ep, tot))
        self.fold()
        return hist

    def fold(self):
        """fold the output standardisation into W3, b3: raw output = correction"""
        if self.folded:
            return
        o3 = H * I_IN + H + H * H + H
        W3 = self.params[o3:o3 + self.O * H].view(self.O, H)
        b3 = self.params[o3 + self.O * H:o3 + self.O * H + self.O]
        W3.mul_(self.std[:, None])
        b3.mul_(self.std).add_(self.mean)
        self.folded = True

    # ------------------------------------------------------------- inference
    def correction(self, lset_nodal, precision=_abi.XG_PREC_BF16, x_norm=None, out=None):
        """the MLP's correction (n x NM, FP32)"""
        import torch
        assert self.folded, "train() or load() first"
        x = normalise_inputs(lset_nodal) if x_norm is None else x_norm
        n = x.shape[0]
        y = out if out is not None else torch.empty(n, self.O, dtype=torch.float32, device=self.device)
        work = None
        if precision == _abi.XG_PREC_FP32:
            work = torch.empty(2 * n * H, dtype=torch.float32, device=self.device)
        _abi.check(_abi.lib().xg_mlp_forward(self.ctx.h, I_IN, H, self.O, _abi.ptr(self.params), n, _abi.ptr(x),
                                             _abi.ptr(y), _abi.ptr(work), precision))
        return y

    def moments(self, lset_nodal, thr=1e-6, bufs=None):
        """learned NEG moments with the degree-2 check and the plain
        fallback: returns (moments n x NM FP64, accepted uint8, residual FP64)"""
        import torch
        n = lset_nodal.shape[0]
        if bufs is None:
            bufs = {}
        get = lambda k, shape, dt: bufs.setdefault(k, torch.empty(shape, dtype=dt, device=self.device))  # noqa
        base = self.plain_moments(lset_nodal, self.base_L, out=get("base", (n, self.O), torch.float64))
        corr = self.correction(lset_nodal, out=get("corr", (n, self.O), torch.float32))
        ref2 = self.plain_moments(lset_nodal, self.L, p=1, out=get("ref2", (n, 10), torch.float64))
        out = get("out", (n, self.O), torch.float64)
        rel = get("rel", (n,), torch.float64)
        acc = get("acc", (n,), torch.uint8)
        _abi.check(_abi.lib().xg_learned_moments(self.ctx.h, n, self.p, _abi.ptr(base), _abi.ptr(corr), _abi.ptr(ref2),
                                                 float(thr), _abi.ptr(out), _abi.ptr(rel), _abi.ptr(acc)))
        rej = torch.nonzero(acc == 0).flatten()
        if rej.numel():
            fb = self.plain_moments(lset_nodal.index_select(0, rej).contiguous(), self.L)
            out.index_copy_(0, rej, fb)
        return out, acc, rel

    # ------------------------------------------------------------- persistence
    def save(self, path):
        np.savez(path, params=self.params.detach().cpu().numpy(), folded=np.array(self.folded),
                 cfg=np.array([self.p, self.base_L, self.L]))

    def load(self, path):
        import torch
        d = np.load(path, allow_pickle=False)
        self.params = torch.from_numpy(d["params"]).to(self.device)
        self.folded = bool(d["folded"])
        self.p, self.base_L, self.L = (int(v) for v in d["cfg"])
        self.O = NMOM[self.p]
