"""Benchmark of the cut-element hot path on one MI355X.

Workload (default): the 3D plain configuration -- unit cube, 64^3 cells x 6
Kuhn tets = 1,572,864 tets, P1 level set |x - (0.5,0.5,0.5)| - 0.35 +
0.01 sin(7x) sin(7y) sin(7z) (perturbed as InterpolateToP1 does), order-4
straight cut rules on NEG and IF, P2 H1 mass + stiffness element matrices.

One step = the fused rule + element-matrix pass (xg_cut_pipeline) over all cut
elements for both domain types, inputs resident in HBM.  value = cut
elements / second.  Marking (xg_mark) is timed separately.

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64]
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cut-elem/s (rule+elmat) plain vs learned/mixed at ≤1e-6 rel err; %HBM/%MFMA"
PEAK_FP32_MFMA = 157.3e12
PEAK_FP64_MFMA = 78.6e12
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="c2", choices=["c2", "c3", "c4"],
                    help="c2: 3D plain straight-cut pipeline (headline); c4: learned quadrature on P2-cut tets")
    ap.add_argument("--c4-n", type=int, default=2000000)
    ap.add_argument("--c4-train", type=int, default=100000)
    return ap.parse_args()


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) from the committed summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def bench_c4(args):
    """Learned quadrature (P2 level sets on single tets): plain subdivision
    rule (level 4, FP64) vs the learned rule (BF16 MLP), both followed by the
    P2 mass + stiffness element matrices on the fixed-node rule (FP32)."""
    import torch

    from ngsxfem_amd import _abi
    from ngsxfem_amd.learned import LearnedRule, random_cut_quadrics

    stream = torch.cuda.current_stream()
    ctx = _abi.Context(0, stream=stream)
    n = args.c4_n
    t0 = time.time()
    lr_ = LearnedRule(ctx, seed=0)
    hist = lr_.train(n_train=args.c4_train, epochs=10, batch=1024, lr=2e-3, L=4, seed=3)
    t_train = time.time() - t0
    x = random_cut_quadrics(n, 42)
    xn = None
    from ngsxfem_amd.learned import normalise_inputs
    xn = normalise_inputs(x)
    wp = torch.empty(n, 35, dtype=torch.float64, device="cuda")
    wl = torch.empty(n, 35, dtype=torch.float32, device="cuda")
    wl64 = torch.empty(n, 35, dtype=torch.float64, device="cuda")
    em = torch.empty(n * 100, dtype=torch.float32, device="cuda")

    def plain():
        ctx.subdiv_weights(x, 2, 4, 4, 4, wp, None, None)
        _abi.check(_abi.lib().xg_elmat_weights(ctx.h, 2, 4, n, None, None, None, _abi.ptr(wp), 3,
                                               _abi.XG_PREC_FP32, _abi.ptr(em)))

    def learned():
        lr_.predict(x, x_norm=xn, out=wl)
        wl64.copy_(wl)
        _abi.check(_abi.lib().xg_elmat_weights(ctx.h, 2, 4, n, None, None, None, _abi.ptr(wl64), 3,
                                               _abi.XG_PREC_FP32, _abi.ptr(em)))

    def timed(f, reps):
        f()
        ctx.sync()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            f()
        b.record(stream)
        ctx.sync()
        return a.elapsed_time(b) / reps

    ms_plain = timed(plain, 2)
    ms_learned = timed(learned, 10)
    err = (torch.linalg.norm(wl.to(torch.float64) - wp, dim=1) / torch.linalg.norm(wp, dim=1)).cpu().numpy()
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = min(16, os.cpu_count() or 1)
        ns = 4000
        xs = x[:ns].cpu().numpy()
        t1 = time.perf_counter()
        O.subdiv_moments3_batch(xs, 2, 4, 4, 4, nthreads=cores)
        dt_ = time.perf_counter() - t1
        cpu = {"value": ns / dt_, "unit": "cut-elem/s", "cores": cores, "kind": "port",
               "sample": "%d of the %d tets, subdivision level 4 weights only (FP64 oracle, %d threads, %.1fs)"
                         % (ns, n, cores, dt_)}
    res = {"metric": METRIC, "config": {"workload": "learned quadrature: %d cut tets, random quadric P2 level "
                                                    "sets, plain = subdivision L=4 FP64 + P2 elmat FP32, learned = "
                                                    "MLP 10-128-128-35 BF16 + P2 elmat FP32" % n},
           "plain": {"ms": ms_plain, "cut_elem_per_s": n / (ms_plain * 1e-3)},
           "learned": {"ms": ms_learned, "cut_elem_per_s": n / (ms_learned * 1e-3),
                       "rel_err_vs_plain_median": float(np.median(err)),
                       "rel_err_vs_plain_p99": float(np.percentile(err, 99)),
                       "frac_within_1e-6": float(np.mean(err <= 1e-6)),
                       "frac_within_1e-3": float(np.mean(err <= 1e-3))},
           "train_s": t_train, "train_loss_last": hist[-1], "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)
    ctx.close()


def bench_c3(args):
    """Batched uneven: 256 cube meshes (n log-uniform in [8, 48], own ellipsoid
    level set, own degree p in {1,2,3}, rule order 2p) in one captured launch
    sequence (marking + plan + persistent fused rule/elmat kernels), FP32."""
    import torch

    from ngsxfem_amd import _abi, autotune
    from ngsxfem_amd import workloads as W

    stream = torch.cuda.current_stream()
    ctx = _abi.Context(0, stream=stream)
    t0 = time.time()
    coords, el2v, lset, off, ps, ns, axes = W.config3(256)
    print("# c3 workload built in %.1fs: %d tets" % (time.time() - t0, off[-1]), file=sys.stderr, flush=True)
    d_x = torch.from_numpy(coords).cuda()
    d_e = torch.from_numpy(el2v).cuda()
    d_l = torch.from_numpy(lset).cuda()
    batch = _abi.Batch(ctx, d_x, d_e, d_l, off, ps)
    tuning = autotune.load()
    autotune.apply(batch, tuning)
    cnt, oo = batch.plan()
    n_cut = int(cnt.sum())
    out = torch.empty(int(oo[-1]), dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        batch.run(out, use_graph=True)
    ctx.sync()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(args.steps):
        batch.run(out, use_graph=True)
    b.record(stream)
    ctx.sync()
    ms = a.elapsed_time(b) / args.steps
    # algorithmic FLOPs: 2 npts ndof^2 (1 + D) per (cut element, dt)
    flops = 0.0
    starts = np.concatenate([[0], np.cumsum(cnt)])
    cut = batch.cut_list_host(n_cut)
    for p in (1, 2, 3):
        sel = np.concatenate([cut[starts[m]:starts[m + 1]] for m in range(len(ps)) if ps[m] == p] or [np.zeros(0, np.int32)])
        if len(sel) == 0:
            continue
        nd = {1: 4, 2: 10, 3: 20}[p]
        dsel = torch.from_numpy(sel.astype(np.int32)).cuda()
        for dtv in (_abi.XG_NEG, _abi.XG_IF):
            o = torch.zeros(len(sel) + 1, dtype=torch.int32, device="cuda")
            ctx.cut_rule_count(3, d_e, d_l, dtv, 2 * p, o, elems=dsel)
            ctx.sync()
            flops += 2.0 * int(o[-1].item()) * nd * nd * 4
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = min(16, os.cpu_count() or 1)
        t_cpu, done = 0.0, 0
        for m in range(len(ps)):
            if t_cpu > args.cpu_seconds:
                break
            seg = cut[starts[m]:starts[m + 1]]
            if len(seg) == 0:
                continue
            t1 = time.perf_counter()
            O.cut_pipeline(3, int(ps[m]), 2 * int(ps[m]), seg, el2v, coords, lset, [0, 2], mask=3, nthreads=cores)
            t_cpu += time.perf_counter() - t1
            done += len(seg)
        cpu = {"value": done / t_cpu, "unit": "cut-elem/s", "cores": cores, "kind": "port",
               "sample": "the first meshes' %d cut tets (mixed p, FP64 oracle, %d threads, %.1fs)" % (done, cores, t_cpu)}
    res = {"metric": METRIC, "value": n_cut / (ms * 1e-3), "unit": "cut-elem/s", "ms_per_step": ms,
           "config": {"workload": "batched uneven: 256 cube meshes n in [8,48] log-uniform (%d tets), ellipsoid "
                                  "P1 level sets, p in {1,2,3}, rule order 2p, NEG+IF, FP32 MFMA, one hipGraph "
                                  "(marking + plan + persistent kernels)" % int(off[-1]),
                      "cut_elements": n_cut, "meshes_per_p": [int(np.sum(ps == q)) for q in (1, 2, 3)],
                      "tuning": tuning},
           "roofline": {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
                        "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / PEAK_FP32_MFMA,
                        "note": "whole graph time (marking of all tets included)"},
           "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)
    batch.close()
    ctx.close()


def main():
    args = parse()
    if args.config == "c4":
        return bench_c4(args)
    if args.config == "c3":
        return bench_c3(args)
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by element "
              "ranges; see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from ngsxfem_amd import _abi
    from ngsxfem_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    ctx = _abi.Context(0, stream=stream)

    D, p, order = 3, 2, 4
    t0 = time.time()
    mesh, lset = W.config2(args.n)
    print("# workload built in %.1fs: %d tets, %d vertices" % (time.time() - t0, mesh.ne, mesh.nv),
          file=sys.stderr, flush=True)
    d_el = torch.from_numpy(mesh.el2v).to(dev)
    d_x = torch.from_numpy(mesh.coords).to(dev)
    d_l = torch.from_numpy(lset).to(dev)
    ne = mesh.ne
    nw = (ne + 31) // 32
    dt = torch.empty(ne, dtype=torch.uint8, device=dev)
    cr = torch.empty(ne, dtype=torch.float64, device=dev)
    bits = torch.empty(8 * nw, dtype=torch.int32, device=dev)
    cut = torch.empty(ne, dtype=torch.int32, device=dev)
    ncut = torch.zeros(1, dtype=torch.int32, device=dev)

    # ---- marking (timed separately) ----
    def mark():
        ctx.mark(D, d_el, d_l, dt, cr, bits, cut, ncut)

    for _ in range(3):
        mark()
    ctx.sync()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    nmark = 20
    ev[0].record(stream)
    for _ in range(nmark):
        mark()
    ev[1].record(stream)
    ctx.sync()
    mark_ms = ev[0].elapsed_time(ev[1]) / nmark
    n_cut = int(ncut.item())
    cut_list = cut[:n_cut].clone()

    # ---- algorithmic work per step ----
    npts = {}
    for dtv in (_abi.XG_NEG, _abi.XG_IF):
        off = torch.zeros(n_cut + 1, dtype=torch.int32, device=dev)
        ctx.cut_rule_count(D, d_el, d_l, dtv, order, off, elems=cut_list)
        ctx.sync()
        npts[dtv] = int(off[-1].item())
    nd = 10
    flops = sum(2.0 * npts[k] * nd * nd * (1 + D) for k in npts)
    bytes_alg = n_cut * (4 * 4 + 8 * 4 + 24 * 4 + 4) + n_cut * 2 * nd * nd * (4 if args.precision == "fp32" else 8)

    prec = _abi.XG_PREC_FP32 if args.precision == "fp32" else _abi.XG_PREC_FP64
    out = torch.empty(n_cut * 2 * nd * nd, dtype=torch.float32 if prec == _abi.XG_PREC_FP32 else torch.float64,
                      device=dev)
    dts = [_abi.XG_NEG, _abi.XG_IF]

    def step():
        ctx.cut_pipeline(D, p, order, d_el, d_x, d_l, dts, 3, prec, out, elems=cut_list)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    g0.record(stream)
    for k in range(args.steps):
        evs[k][0].record(stream)
        step()
        evs[k][1].record(stream)
    g1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    total_ms = g0.elapsed_time(g1)
    launch_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    ms_per_step = total_ms / args.steps
    value = n_cut / (ms_per_step * 1e-3)

    peak = PEAK_FP32_MFMA if prec == _abi.XG_PREC_FP32 else PEAK_FP64_MFMA
    achieved = flops / (launch_ms * 1e-3)
    roof = {"bound": "mfma", "kernel": "k_pipeline<3,2,%s>" % ("float" if prec == 1 else "double"),
            "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": read_traffic("pipeline_c2_pmc.txt"),
            "algorithmic_flops_per_launch": flops, "algorithmic_bytes_per_launch": bytes_alg,
            "hbm_frac_algorithmic": bytes_alg / (launch_ms * 1e-3) / PEAK_HBM,
            "avg_launch_ms": launch_ms}

    # ---- CPU baseline: FP64 restatement of the reference algorithm ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = min(16, os.cpu_count() or 1)
        cl = cut_list.cpu().numpy()
        reps, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and reps < 100:
            t1 = time.perf_counter()
            O.cut_pipeline(D, p, order, cl, mesh.el2v, mesh.coords, lset, dts, mask=3, nthreads=cores)
            t_cpu += time.perf_counter() - t1
            reps += 1
        cpu = {"value": reps * n_cut / t_cpu, "unit": "cut-elem/s", "cores": cores, "kind": "port",
               "sample": "%d passes over all %d cut tets (FP64 oracle, %d threads, %.1fs)" % (reps, n_cut, cores, t_cpu)}

    res = {
        "metric": METRIC, "value": value, "unit": "cut-elem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "3D plain: unit cube %d^3 x 6 Kuhn tets (%d tets), P1 sphere r=0.35 + 0.01 "
                               "sin(7x)sin(7y)sin(7z), order-4 straight cut rule NEG+IF, P2 mass+stiffness "
                               "elmat (%s MFMA)" % (args.n, ne, args.precision),
                   "cut_elements": n_cut, "points_neg": npts[0], "points_if": npts[2]},
        "roofline": roof, "cpu_baseline": cpu,
        "marking": {"ms": mark_ms, "tets_per_s": ne / (mark_ms * 1e-3),
                    "hbm_frac_algorithmic": ne * 26.4 / (mark_ms * 1e-3) / PEAK_HBM},
    }
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
