"""Learned quadrature for P2-cut tetrahedra (host side).

A small MLP (10 -> 128 -> 128 -> 35, tanh) maps the normalised P2 nodal
level-set values of a cut tet to the NEG-part weights W_k on the 35 nodes of
the degree-4 Lagrange lattice (W_k = integral over {phi < 0} of psi_k).  These
weights integrate every polynomial of degree <= 4 -- hence every P2 mass and
stiffness integrand -- exactly like the rule they approximate.

* Labels come from the plain rule: the subdivision rule at level L
  (cutint/xintegration.cpp:1626-1754) contracted to fixed-node weights on the
  GPU (xg_subdiv_weights).
* Training runs on the GPU with the library's own FP32 forward/backward and
  Adam kernels (xg_mlp_train_step); inference uses the fused BF16 MFMA kernel
  (xg_mlp_forward, FP32 accumulation).
* The moment-residual check (xg_moment_residual, FP64) compares a prediction
  with the plain rule's weights and falls back to the plain weights above a
  threshold; the plain rule also judges the learned rule's error.

The output standardisation (per-output mean / std of the training labels) is
folded into the last layer after training, so the network's raw output is W.
"""
import numpy as np

from . import _abi

I_IN, H, O_OUT, Q = 10, 128, 35, 4


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
    def __init__(self, ctx, seed=0, device="cuda"):
        import torch
        self.ctx = ctx
        self.device = device
        npar = _abi.lib().xg_mlp_nparams(I_IN, H, O_OUT)
        g = torch.Generator(device="cpu").manual_seed(seed)
        p = torch.zeros(npar, dtype=torch.float32)
        o = 0
        for fan_in, n_w, n_b in ((I_IN, H * I_IN, H), (H, H * H, H), (H, O_OUT * H, O_OUT)):
            bound = 1.0 / np.sqrt(fan_in)
            p[o:o + n_w] = (torch.rand(n_w, generator=g) * 2 - 1) * bound
            o += n_w + n_b
        self.params = p.to(device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.mean = torch.zeros(O_OUT, dtype=torch.float32, device=device)
        self.std = torch.ones(O_OUT, dtype=torch.float32, device=device)
        self.folded = False

    # ------------------------------------------------------------- labels
    def plain_weights(self, lset_nodal, L=4, order=4, want_if=False):
        """subdivision rule at level L contracted to fixed-node weights (GPU)"""
        import torch
        n = lset_nodal.shape[0]
        wn = torch.empty(n, O_OUT, dtype=torch.float64, device=self.device)
        wi = torch.empty(n, O_OUT, dtype=torch.float64, device=self.device) if want_if else None
        dom = torch.empty(n, dtype=torch.uint8, device=self.device)
        self.ctx.subdiv_weights(lset_nodal, 2, L, order, Q, wn, wi, dom)
        return wn, wi, dom

    # ------------------------------------------------------------- training
    def train(self, n_train=200000, epochs=20, batch=1024, lr=1e-3, L=4, seed=1, log=None):
        import torch
        x_raw = random_cut_quadrics(n_train, seed, self.device)
        w, _, _ = self.plain_weights(x_raw, L=L)
        self.ctx.sync()
        x = normalise_inputs(x_raw)
        y = w.to(torch.float32)
        self.mean = y.mean(dim=0)
        self.std = y.std(dim=0).clamp_min(1e-12)
        t = ((y - self.mean) / self.std).contiguous()
        work = torch.empty(4 * batch * H + 2 * batch * O_OUT, dtype=torch.float32, device=self.device)
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
                    self.ctx.h, I_IN, H, O_OUT, _abi.ptr(self.params), _abi.ptr(self.m), _abi.ptr(self.v),
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
        """fold the output standardisation into W3, b3: raw output = W"""
        if self.folded:
            return
        o3 = H * I_IN + H + H * H + H
        W3 = self.params[o3:o3 + O_OUT * H].view(O_OUT, H)
        b3 = self.params[o3 + O_OUT * H:o3 + O_OUT * H + O_OUT]
        W3.mul_(self.std[:, None])
        b3.mul_(self.std).add_(self.mean)
        self.folded = True

    # ------------------------------------------------------------- inference
    def predict(self, lset_nodal, precision=_abi.XG_PREC_BF16, x_norm=None, out=None):
        """fixed-node weights (n x 35, FP32) of the NEG part"""
        import torch
        assert self.folded, "train() or load() first"
        x = normalise_inputs(lset_nodal) if x_norm is None else x_norm
        n = x.shape[0]
        y = out if out is not None else torch.empty(n, O_OUT, dtype=torch.float32, device=self.device)
        work = None
        if precision == _abi.XG_PREC_FP32:
            work = torch.empty(2 * n * H, dtype=torch.float32, device=self.device)
        _abi.check(_abi.lib().xg_mlp_forward(self.ctx.h, I_IN, H, O_OUT, _abi.ptr(self.params), n, _abi.ptr(x),
                                             _abi.ptr(y), _abi.ptr(work), precision))
        return y

    def with_fallback(self, pred, plain, thr=1e-6):
        """moment-residual check against the plain weights; returns
        (weights FP64, accepted mask, relative residual)"""
        import torch
        n = pred.shape[0]
        rel = torch.empty(n, dtype=torch.float64, device=self.device)
        acc = torch.empty(n, dtype=torch.uint8, device=self.device)
        out = torch.empty(n, O_OUT, dtype=torch.float64, device=self.device)
        _abi.check(_abi.lib().xg_moment_residual(self.ctx.h, n, O_OUT, _abi.ptr(pred), 1.0, _abi.ptr(plain), thr,
                                                 _abi.ptr(rel), _abi.ptr(acc), _abi.ptr(out)))
        return out, acc, rel

    # ------------------------------------------------------------- persistence
    def state(self):
        return {"params": self.params.detach().cpu().numpy(), "folded": self.folded}

    def save(self, path):
        np.savez(path, params=self.params.detach().cpu().numpy(), folded=np.array(self.folded))

    def load(self, path):
        import torch
        d = np.load(path, allow_pickle=False)
        self.params = torch.from_numpy(d["params"]).to(self.device)
        self.folded = bool(d["folded"])
