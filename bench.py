"""Benchmark of the cut-element hot path on one MI355X.

Workload (default): the large mixed configuration -- unit cube, 128^3 cells x
6 Kuhn tets = 12,582,912 tets, P2 level set of the union (min) of 64 spheres
(radii U[0.05, 0.15], centres U[0.2, 0.8]^3, MT19937 seed 7); P2 marking at
subdivision level 4 (timed on its own); one step = the level-4 subdivision
rule of every cut element on NEG and IF contracted into FP64 degree-6
moments, and the P3 mass + stiffness element matrices from them
(xg_subdiv_pipeline, FP32 store), inputs resident in HBM.  value = cut
elements / second.  The 3D plain configuration (64^3 x 6 Kuhn tets, P1 level
set, order-4 straight cut rules, P2 matrices) is measured in the same run and
reported under "c2".

    python bench.py [--steps K] [--warmup W] [--config c5|c2|c3|c4|extras]
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
PEAK_FP64_VALU = 78.6e12


def fp64_ops_moments(npts, n_cut, D=3, p=2):
    """FP64 flops of the moment path (FMA = 2).  The order-4 rule counts npts
    (27-point tets, 9-point triangles) give the numbers of pieces.  With exact
    rules the moment form takes each piece's moments from its vertices (the
    complete homogeneous recursion of degree 4: per vertex 4 adds and 136
    FMAs, plus 7 operations to scale its barycentrics, and 16 per piece for
    its volume and weight); per (element, domain type) the alpha! factors
    (34 products), the degree-2 lowering (90 adds) and the Bernstein-form
    identities (508 coefficient terms, one FMA each, 20 scalings)."""
    per_vertex = 7 + 4 + 2 * 136
    pieces_tet, pieces_tri = npts[0] // 27, npts[2] // 9
    per_el = 34 + 90 + 2 * 508 + 20
    return pieces_tet * (4 * per_vertex + 16) + pieces_tri * (3 * per_vertex + 16) + 2 * n_cut * per_el


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64, help="cells per direction")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"], help="element-matrix store type")
    ap.add_argument("--method", default="moments", choices=["moments", "mfma"],
                    help="path reported as value (both are measured)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="launch each path 3 times plus a 256 MiB copy, for rocprofv3 --pmc runs")
    ap.add_argument("--config", default="c5", choices=["c2", "c3", "c4", "c5", "extras"],
                    help="c5 (default): large mixed P2-level-set pipeline, with the 3D plain straight-cut "
                         "pipeline (c2) beside it; c2 / c3 / c4 / extras: one configuration alone")
    ap.add_argument("--c4-n", type=int, default=2000000)
    ap.add_argument("--c4-train", type=int, default=100000)
    ap.add_argument("--c5-n", type=int, default=128)
    ap.add_argument("--c5-L", type=int, default=4)
    ap.add_argument("--no-c5-learned", action="store_true", help="skip the learned leg of the large mixed line")
    return ap.parse_args()


def read_profile_value(name, key):
    """a "key: value" figure from a committed profile summary, or None"""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(re.escape(key) + r"\s*[:=]\s*([0-9.eE+-]+)", open(path).read())
    return float(m.group(1)) if m else None


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) from the committed summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def bench_c4(args):
    """Learned quadrature (P2 level sets on single reference tets, the NEG
    part): the plain rule is the level-4 subdivision rule contracted into
    FP64 degree-4 moments, then the P2 mass + stiffness element matrices
    (FP32 store).  The learned path corrects the level-1 moments with the MLP
    (BF16 MFMA inference), checks the result against the exact degree-2
    moments of the level-4 rule and falls back to the level-4 rule where the
    check fails; its time includes the check and the fallback.  Plain rules
    at levels 1-3 are timed beside it, with their error against level 4, to
    compare at equal error."""
    import torch

    from ngsxfem_amd import _abi
    from ngsxfem_amd.learned import LearnedRule, random_cut_quadrics

    stream = torch.cuda.current_stream()
    ctx = _abi.Context(0, stream=stream)
    n = args.c4_n
    t0 = time.time()
    lr_ = LearnedRule(ctx, p=2, base_L=1, L=4, seed=0)
    hist = lr_.train(n_train=args.c4_train, epochs=10, batch=1024, lr=2e-3, seed=3)
    t_train = time.time() - t0
    x = random_cut_quadrics(n, 42)
    em = torch.empty(n * 100, dtype=torch.float32, device="cuda")

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

    mom = {L: torch.empty(n, 35, dtype=torch.float64, device="cuda") for L in (1, 2, 3, 4)}

    def plain(L):
        lr_.plain_moments(x, L, out=mom[L])
        ctx.elmat_moments(3, 2, mom[L], 3, _abi.XG_PREC_FP32, em, nel=n)

    ms_plain = {L: timed(lambda L=L: plain(L), 3) for L in (4, 3, 2, 1)}
    m4 = mom[4]
    em4 = torch.empty(n * 100, dtype=torch.float64, device="cuda")
    ctx.elmat_moments(3, 2, m4, 3, _abi.XG_PREC_FP64, em4, nel=n)
    ctx.sync()
    e4 = em4.view(n, 100)
    n4 = torch.linalg.vector_norm(e4, dim=1)

    def elmat_err(m):
        e = torch.empty(n * 100, dtype=torch.float64, device="cuda")
        ctx.elmat_moments(3, 2, m, 3, _abi.XG_PREC_FP64, e, nel=n)
        ctx.sync()
        return (torch.linalg.vector_norm(e.view(n, 100) - e4, dim=1) / n4).cpu().numpy()

    bufs = {}
    res_thr = {}
    for thr in (1e-6, 1e-4, 1e-3, 1e-2):
        out, acc, rel = lr_.moments(x, thr=thr, bufs=bufs)
        ctx.sync()
        a = acc.bool()
        err_acc = elmat_err(out)
        res_thr["%g" % thr] = {"accepted_frac": float(a.float().mean().item()),
                               "max_rel_err_elmat_accepted": float(err_acc[a.cpu().numpy()].max()) if a.any() else None,
                               "max_rel_err_elmat_all": float(err_acc.max())}

    def learned():
        out, acc, rel = lr_.moments(x, thr=1e-6, bufs=bufs)
        ctx.elmat_moments(3, 2, out, 3, _abi.XG_PREC_FP32, em, nel=n)

    ms_learned = timed(learned, 3)
    # the network's own accuracy (before the check): level-1 moments + correction
    corr = lr_.correction(x).to(torch.float64)
    err_nn = elmat_err(mom[1] + corr)
    err_plain = {L: elmat_err(mom[L]) for L in (1, 2, 3)}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = cpu_threads()
        ns = 2000
        xs = x[:ns].cpu().numpy()
        V = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [0, 0, 0]])
        coords = np.tile(V, (ns, 1))
        el2v = np.arange(4 * ns, dtype=np.int32).reshape(ns, 4)
        t1 = time.perf_counter()
        O.subdiv_pipeline(2, 4, 4, np.arange(ns, dtype=np.int32), el2v, coords, xs, [0], 3, nthreads=cores)
        dt_ = time.perf_counter() - t1
        cpu = {"value": ns / dt_, "unit": "cut-elem/s", "cores": cores, "kind": "port",
               "sample": "%d of the %d tets: level-4 subdivision rule materialised + P2 mass+stiffness elmat per "
                         "point (FP64 oracle, %d threads, %.1fs)" % (ns, n, cores, dt_)}
    q = lambda a: {"median": float(np.median(a)), "p99": float(np.percentile(a, 99)), "max": float(a.max())}  # noqa
    res = {"metric": METRIC, "config": {"workload": "learned quadrature: %d cut tets, random quadric P2 level sets "
                                                    "(seed 42), NEG part, plain = subdivision L=4 -> FP64 degree-4 "
                                                    "moments -> P2 elmat (FP32 store); learned = L=1 moments + MLP "
                                                    "10-128-128-35 correction (BF16 MFMA) + degree-2 check against "
                                                    "the L=4 rule + L=4 fallback" % n},
           "plain": {"ms": ms_plain[4], "cut_elem_per_s": n / (ms_plain[4] * 1e-3)},
           "plain_lower_levels": {"L%d" % L: {"ms": ms_plain[L], "cut_elem_per_s": n / (ms_plain[L] * 1e-3),
                                              "rel_err_elmat_vs_L4": q(err_plain[L])} for L in (1, 2, 3)},
           "learned": {"ms": ms_learned, "cut_elem_per_s": n / (ms_learned * 1e-3), "threshold": 1e-6,
                       "includes": "L1 moments + MLP + L4 degree-2 moments + check + L4 fallback of the rejected "
                                   "+ elmat",
                       "network_rel_err_elmat_vs_L4": q(err_nn),
                       "by_threshold": res_thr},
           "train_s": t_train, "train_loss_last": hist[-1], "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)
    ctx.close()


def c5_fp64_ops(stats, launches, p=3):
    """FP64 operations (FMA = 2) of the exact-rule subdivision kernel from its
    per-element counts (stats[..., 0..3]: table rows, volume tets and
    interface triangles integrated point by point, cut leaves) and the number
    of launches per element (each evaluates the lattice once):
      lattice      969 P2 evaluations (10 shape functions: 22 products for
                   the shapes, 10 FMA) = 42 flops each;
      table rows   NM additions each;
      volume tets  24 points x (mapping 27 + degree-3 monomials 30 + weighted
                   monomials 21 + NM FMA) + determinant 17;
      IF triangles 12 points x (mapping 18 + 30 + 21 + 2 NM) + 40 (cross
                   product, normal, interface transform);
      cut leaves   6 edges x 8 (cut points) + 4 P2 evaluations x 42."""
    nm = {1: 10, 2: 35, 3: 84}[p]
    nq_v, nq_i = {1: (4, 3), 2: (14, 6), 3: (24, 12)}[p]
    rows, vol, itf, leaves = (float(stats[..., k].sum()) for k in range(4))
    return (launches * 969 * 42 + rows * nm + vol * (nq_v * (27 + 30 + 21 + 2 * nm) + 17)
            + itf * (nq_i * (18 + 30 + 21 + 2 * nm) + 40) + leaves * (48 + 4 * 42))


def c5_leaf_ops(st_vol, st_if, p=3):
    """FP64 operations (FMA = 2) of the cut-leaf kernel k_sdx_leaves (volume
    side and interface in one launch) from the per-element counts of the two
    parts (stats[..., 0..3] as c5_fp64_ops): the volume tets and interface
    triangles point by point, the cut leaves' pieces once."""
    nm = {1: 10, 2: 35, 3: 84}[p]
    nq_v, nq_i = {1: (4, 3), 2: (14, 6), 3: (24, 12)}[p]
    vol = float(st_vol[..., 1].sum())
    itf = float(st_if[..., 2].sum())
    leaves = float(st_vol[..., 3].sum())
    if p == 3:
        # centred affine moments: per point the map (25 volume / 18 interface),
        # 16 monomials, 19 weighted monomials, 83 FMAs + 1 add; per element the
        # 84 x 84 map to the homogeneous moments for each of the two parts
        nel = float(st_vol[..., 0].size)
        return (vol * (nq_v * (25 + 16 + 19 + 167) + 17) + itf * (nq_i * (18 + 16 + 19 + 167) + 40)
                + leaves * (48 + 4 * 42) + nel * 2 * 2 * nm * nm)
    return (vol * (nq_v * (27 + 30 + 21 + 2 * nm) + 17) + itf * (nq_i * (18 + 30 + 21 + 2 * nm) + 40)
            + leaves * (48 + 4 * 42))


def cpu_threads():
    """host threads for the CPU baseline: every CPU this process may run on"""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure_c5(args):
    """Large mixed: 128^3 cells x 6 Kuhn tets (12.6M tets), P2 level set of the
    union of 64 random spheres (MT19937 seed 7), P3 element matrices.
    Marking (P2, level 4) is timed on its own; one step is the fused
    subdivision rule (level 4, order 6) + P3 mass + stiffness matrices on NEG
    and IF of every cut element (xg_subdiv_pipeline: FP64 moments and
    identities, FP32 store)."""
    import torch

    from ngsxfem_amd import _abi
    from ngsxfem_amd import workloads as W
    from ngsxfem_amd.meshes import MakeStructured3DMesh

    stream = torch.cuda.current_stream()
    ctx = _abi.Context(0, stream=stream)
    dev = torch.device("cuda:0")
    n, L, p, order = args.c5_n, args.c5_L, 3, 6
    nd, nm = 20, 84
    t0 = time.time()
    mesh = MakeStructured3DMesh(nx=n, ny=n, nz=n)
    centres, radii = W.config5_spheres(7)
    d_el = torch.from_numpy(mesh.el2v).to(dev)
    d_x = torch.from_numpy(mesh.coords).to(dev)
    lp2 = W.p2_nodal_values_torch(d_el, d_x, centres, radii)
    torch.cuda.synchronize()
    print("# large mixed workload built in %.1fs: %d tets" % (time.time() - t0, mesh.ne), file=sys.stderr, flush=True)
    ne = mesh.ne
    nw = (ne + 31) // 32
    dt = torch.empty(ne, dtype=torch.uint8, device=dev)
    bits = torch.empty(8 * nw, dtype=torch.int32, device=dev)
    cut = torch.empty(ne, dtype=torch.int32, device=dev)
    ncut = torch.zeros(1, dtype=torch.int32, device=dev)

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

    ms_mark = timed(lambda: ctx.mark_p2(lp2, L, dt, bits, cut, ncut), 5)
    n_cut = int(ncut.item())
    cl = cut[:n_cut].clone()
    counts = np.bincount(dt.cpu().numpy(), minlength=3)
    dts = [_abi.XG_NEG, _abi.XG_IF]
    out = torch.empty(n_cut * 2 * nd * nd, dtype=torch.float32, device=dev)

    def step():
        ctx.subdiv_pipeline(p, L, order, cl, d_el, d_x, lp2, dts, 3, _abi.XG_PREC_FP32, out)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    g0.record(stream)
    for _ in range(args.steps):
        step()
    g1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    ms = g0.elapsed_time(g1) / args.steps
    print("# large mixed step %.3f ms over %d steps" % (ms, args.steps), file=sys.stderr, flush=True)

    # the subdivision kernels of the step (walk, NEG leaf pieces, interface
    # leaf pieces, table rows; the one-kernel path for list overflows), each
    # launch between its own pair of events on the context's stream
    # (xg_timing_*), 3 repetitions of the step's subdivision part
    mom2 = torch.empty(n_cut * 2 * nm, dtype=torch.float64, device=dev)
    stats = torch.zeros(n_cut * 2 * 4, dtype=torch.int32, device=dev)
    ms_sub = timed(lambda: ctx.subdiv_moments_mesh(p, L, order, cl, d_el, d_x, lp2, dts, mom2), 3)
    ctx.timing_enable(True)
    for _ in range(3):
        ctx.subdiv_moments_mesh(p, L, order, cl, d_el, d_x, lp2, dts, mom2)
    kt = {}
    for k in ("sdx_tree", "sdx_leaves_both", "sdx_rows", "subdiv_exact"):
        tot, nl = ctx.timing_read(k)
        kt[k] = {"ms_per_step": tot / 3, "launches_per_step": nl / 3}
    ctx.timing_enable(False)
    ctx.subdiv_moments_mesh(p, L, order, cl, d_el, d_x, lp2, dts, mom2, stats)
    ctx.sync()
    st = stats.cpu().numpy().reshape(n_cut, 2, 4)
    ops = c5_leaf_ops(st[:, 0], st[:, 1])
    ms_kern = kt["sdx_leaves_both"]["ms_per_step"]
    rows_bytes = float(st[:, 0, 0].sum()) * nm * 8
    if args.pmc_probe:
        # counter runs (rocprofv3 --pmc): a 256 MiB device copy of known size
        # (calibration of FETCH_SIZE / WRITE_SIZE), then 3 kernel launches
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        for _ in range(3):
            ctx.subdiv_moments_mesh(p, L, order, cl, d_el, d_x, lp2, dts, mom2)
        ctx.sync()
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": 2 * a.numel() * 4, "cut_elements": n_cut,
                          "leaves_flops_per_step": ops,
                          "kernels": kt}), flush=True)
        ctx.close()
        return None

    # error of the FP32 step against the same pipeline with FP64 output
    ref = torch.empty(n_cut * 2 * nd * nd, dtype=torch.float64, device=dev)
    ctx.subdiv_pipeline(p, L, order, cl, d_el, d_x, lp2, dts, 3, _abi.XG_PREC_FP64, ref)
    ctx.sync()
    r64 = ref.view(-1, nd * nd)
    rn = torch.linalg.vector_norm(r64, dim=1)
    nzr = rn > 0
    dn = torch.linalg.vector_norm(out.view(-1, nd * nd).double() - r64, dim=1)
    err32 = float((dn[nzr] / rn[nzr]).max().item())

    roof = {"bound": "fp64-valu",
            "kernel": "k_sdx_leaves<2,3,NEG,3> (NEG and interface cut-leaf pieces, the step's largest)",
            "achieved": ops / (ms_kern * 1e-3) / 1e12, "peak": PEAK_FP64_VALU / 1e12, "unit": "TFLOP/s",
            "frac": ops / (ms_kern * 1e-3) / PEAK_FP64_VALU, "traffic": read_traffic("subdiv_c5_pmc.txt"),
            "traffic_source": "profiles/subdiv_c5_pmc.txt (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the same "
                              "kernel, copy-calibrated, summed over its launches of one step)",
            "algorithmic_flops_per_launch": ops, "avg_launch_ms": ms_kern,
            "launch_note": "one launch per step (a launch takes up to 524,288 elements); flops and ms are per step",
            "counts_per_cut_element": {"table_rows": float(st[:, 0, 0].mean()),
                                       "volume_tets": float(st[:, 0, 1].mean()),
                                       "interface_triangles": float(st[:, 1, 2].mean()),
                                       "cut_leaves": float(st[:, 0, 3].mean())},
            "kernels": {
                "k_sdx_tree<2,3,NEG> (refinement walk, node lists, the kept side's table rows)": {
                    "ms": kt["sdx_tree"]["ms_per_step"], "bound": "VALU issue (integer lattice checks)"},
                "k_sdx_rows<3> (complemented leaves' table rows)": {
                    "ms": kt["sdx_rows"]["ms_per_step"], "bound": "L2 (the depth <= 4 rows, 3.1 MB)"},
                "table_rows_bytes_per_step": rows_bytes,
                "k_subdiv_exact (list overflows)": {"ms": kt["subdiv_exact"]["ms_per_step"],
                                                    "launches": kt["subdiv_exact"]["launches_per_step"]}},
            "note": "flops counted from the kernels' own per-element piece counts (bench.py c5_leaf_ops); each "
                    "kernel time is the sum of its launches' event pairs on the stream (xg_timing_read)"}
    res = {"metric": METRIC, "value": n_cut / (ms * 1e-3), "unit": "cut-elem/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp64+fp32", "data": "synthetic",
           "config": {"workload": "large mixed: unit cube %d^3 x 6 Kuhn tets (%d tets), P2 level set = min of 64 "
                                  "spheres r U[0.05,0.15] centres U[0.2,0.8]^3 (MT19937 seed 7), subdivision rule "
                                  "L=%d order 6 on NEG+IF -> FP64 degree-6 moments -> P3 mass+stiffness elmat "
                                  "(FP64 identities, FP32 store)" % (n, ne, L),
                      "cut_elements": n_cut, "neg": int(counts[0]), "pos": int(counts[1])},
           "roofline": roof,
           "max_rel_err_vs_fp64": err32,
           "paths": {"subdivision_ms": ms_sub, "elmat_and_rest_ms": ms - ms_sub},
           "marking": {"ms": ms_mark, "tets_per_s": ne / (ms_mark * 1e-3), "kernel": "xg_mark_p2 (L=4)"}}

    if not args.no_c5_learned:
        res["learned"] = c5_learned(args, ctx, stream, cl, d_el, d_x, lp2, L, ms)

    # ---- CPU baseline: the oracle (reference algorithm: materialised level-4
    # subdivision rule + P3 element matrices per point, FP64) on a bounded
    # sample of the cut elements, all host threads
    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        cln = cl.cpu().numpy()
        vals = lp2.cpu().numpy()
        rng = np.random.default_rng(11)
        order_ = rng.permutation(n_cut)
        probe = np.sort(cln[order_[:threads]]).astype(np.int32)
        t1 = time.perf_counter()
        O.subdiv_pipeline(p, L, order, probe, mesh.el2v, mesh.coords, vals, dts, 3, nthreads=threads)
        per = (time.perf_counter() - t1) / len(probe)
        ns = int(min(n_cut, max(threads, args.cpu_seconds / max(per, 1e-9))))
        sample = np.sort(cln[order_[:ns]]).astype(np.int32)
        t1 = time.perf_counter()
        ref_s = O.subdiv_pipeline(p, L, order, sample, mesh.el2v, mesh.coords, vals, dts, 3, nthreads=threads)
        dt_ = time.perf_counter() - t1
        # the GPU's FP32 step output on the same elements against the oracle
        pos = np.searchsorted(cln, sample)
        g = out.view(n_cut, 2, nd, nd)[torch.from_numpy(pos).to(dev).long()].double().cpu().numpy()
        scale = np.abs(ref_s).max(axis=(-2, -1)) + 1e-300
        err_o = float((np.abs(g - ref_s).max(axis=(-2, -1)) / scale).max())
        res["cpu_baseline"] = {"value": ns / dt_, "unit": "cut-elem/s", "cores": threads, "kind": "port",
                               "sample": "%d random cut elements of %d (oracle: level-4 subdivision rule "
                                         "materialised, P3 elmat per point, NEG+IF, FP64, %d threads, %.1fs; %s, "
                                         "os.cpu_count()=%s)" % (ns, n_cut, threads, dt_, cpu_model(),
                                                                 os.cpu_count())}
        res["max_rel_err_vs_oracle_sample"] = err_o
    ctx.close()
    return res


def c5_learned(args, ctx, stream, cl, d_el, d_x, lp2, L, ms_plain):
    """The learned rule on the large mixed workload: the NEG moments of the cut
    elements from the level-1 moments + the MLP correction (p = 3, 84
    outputs, BF16 MFMA), checked on degree 2 against the level-4 rule, the
    rejected elements from the level-4 rule; the interface moments from the
    plain (staged) path; then the P3 element matrices of both.  The network is
    trained on every other cut element of this workload and evaluated on all."""
    import torch

    from ngsxfem_amd import _abi
    from ngsxfem_amd.learned import LearnedRule

    n = cl.shape[0]
    lc = lp2[cl.long()].contiguous()
    t0 = time.time()
    lr_ = LearnedRule(ctx, p=3, base_L=1, L=L, seed=0)
    hist = lr_.train(epochs=8, batch=1024, lr=2e-3, seed=3, x_raw=lc[::2].contiguous())
    ctx.sync()
    t_train = time.time() - t0
    nd, nm = 20, 84
    mif = torch.empty(n * nm, dtype=torch.float64, device="cuda")
    e_neg = torch.empty(n * nd * nd, dtype=torch.float32, device="cuda")
    e_if = torch.empty_like(e_neg)
    bufs = {}

    def step():
        out, acc, rel = lr_.moments(lc, thr=1e-6, bufs=bufs)
        ctx.subdiv_moments_mesh(3, L, 6, cl, d_el, d_x, lp2, [_abi.XG_IF], mif)
        ctx.elmat_moments(3, 3, out, 3, _abi.XG_PREC_FP32, e_neg, nel=n, elems=cl, el2v=d_el, coords=d_x)
        ctx.elmat_moments(3, 3, mif, 3, _abi.XG_PREC_FP32, e_if, nel=n, elems=cl, el2v=d_el, coords=d_x)

    step()
    ctx.sync()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(3):
        step()
    b.record(stream)
    ctx.sync()
    ms = a.elapsed_time(b) / 3
    # accuracy of the network (before the check) and of the plain level-1
    # rule, element matrices against level 4 (NEG, FP64)
    m4 = lr_.plain_moments(lc, L)
    m1 = lr_.plain_moments(lc, 1)
    corr = lr_.correction(lc).to(torch.float64)

    def em(m):
        e = torch.empty(n * nd * nd, dtype=torch.float64, device="cuda")
        ctx.elmat_moments(3, 3, m, 3, _abi.XG_PREC_FP64, e, nel=n, elems=cl, el2v=d_el, coords=d_x)
        ctx.sync()
        return e.view(n, nd * nd)

    e4 = em(m4)
    n4 = torch.linalg.vector_norm(e4, dim=1).clamp_min(1e-300)
    q = lambda t: {"median": float(t.median()), "p99": float(torch.quantile(t.float(), 0.99)),  # noqa
                   "max": float(t.max())}
    err_nn = torch.linalg.vector_norm(em(m1 + corr) - e4, dim=1) / n4
    err_l1 = torch.linalg.vector_norm(em(m1) - e4, dim=1) / n4
    by_thr = {}
    for thr in (1e-6, 1e-4, 1e-3, 1e-2):
        out, acc, rel = lr_.moments(lc, thr=thr, bufs={})
        ctx.sync()
        by_thr["%g" % thr] = float(acc.float().mean())
    return {"ms": ms, "cut_elem_per_s": n / (ms * 1e-3), "plain_ms": ms_plain,
            "includes": "L1 moments (p = 3) + MLP 10-128-128-84 (BF16) + L4 degree-2 moments + check + L4 "
                        "fallback of the rejected (NEG) + plain interface moments + both P3 elmat",
            "network_rel_err_elmat_vs_L4": q(err_nn), "level1_rel_err_elmat_vs_L4": q(err_l1),
            "accepted_frac_by_threshold": by_thr, "train_s": t_train, "train_loss_last": hist[-1],
            "train_samples": int((n + 1) // 2)}


def bench_c5(args):
    res = measure_c5(args)
    if res is not None:
        print(json.dumps(res), flush=True)


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
    nds = np.array([{1: 4, 2: 10, 3: 20}[int(q)] for q in ps])
    bytes_alg = 26.4 * float(off[-1]) + float(np.sum(cnt * (148 + 2 * nds * nds * 4)))
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = cpu_threads()
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
                                  "P1 level sets, p in {1,2,3}, rule order 2p, NEG+IF, FP32 store, per-degree method "
                                  "and launch shape from the autotuner, one hipGraph (marking + plan + persistent "
                                  "kernels)" % int(off[-1]),
                      "cut_elements": n_cut, "meshes_per_p": [int(np.sum(ps == q)) for q in (1, 2, 3)],
                      "tuning": tuning},
           "roofline": {"bound": "hbm", "achieved": bytes_alg / (ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9,
                        "unit": "GB/s", "frac": bytes_alg / (ms * 1e-3) / PEAK_HBM,
                        "algorithmic_bytes": bytes_alg,
                        "mfma_equivalent_tflops": flops / (ms * 1e-3) / 1e12,
                        "note": "whole graph time (marking of all tets included); bytes = marking 26.4 B/tet + "
                                "148 B in and ndt*ndof^2*4 B out per cut element"},
           "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)
    batch.close()
    ctx.close()


def bench_extras(args):
    """The paths beside the headline one, each timed on its own with events
    (inputs resident, 10 launches): node and dof marking on the 64^3 Kuhn
    mesh (1.57M tets), several level sets on the 32^3 Kuhn mesh, straight cuts
    of a 64^3 hexahedral mesh.  Memory-bound ones carry the HBM fraction of
    their algorithmic bytes."""
    import torch

    from ngsxfem_amd import _abi
    from ngsxfem_amd import workloads as W
    from ngsxfem_amd.meshes import MakeStructured3DMesh

    stream = torch.cuda.current_stream()
    ctx = _abi.Context(0, stream=stream)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731

    def timed(f, reps=10):
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

    out = {}
    # ---- node and dof marking (cutinfo.cpp:135-224, 598-647) on the 3D plain mesh
    mesh, lset = W.config2(64)
    ne, nv = mesh.ne, mesh.nv
    e2v, el2e = mesh.edges()
    f2v, el2f, f2el = mesh.facets()
    d_el, d_l = up(mesh.el2v), up(lset)
    dt = torch.empty(ne, dtype=torch.uint8, device="cuda")
    cr = torch.empty(ne, dtype=torch.float64, device="cuda")
    bits = torch.empty(8 * ((ne + 31) // 32), dtype=torch.int32, device="cuda")
    ctx.mark(3, d_el, d_l, dt, cr, bits)
    tabs = {"vertex": (d_el, nv), "edge": (up(el2e), len(e2v)), "face": (up(el2f), len(f2v))}
    nodes = {}
    for name, (tab, nn) in tabs.items():
        cut = torch.zeros((nn + 31) // 32, dtype=torch.int32, device="cuda")
        dom = torch.empty(nn, dtype=torch.uint8, device="cuda")
        ms = timed(lambda: ctx.node_domains(tab, dt, None, nn, cut, dom))
        # element table + dt read, one flag bit and one domain byte per node written
        alg = ne * (4 * tab.shape[1] + 1) + nn * (1 + 1 / 8)
        nodes[name] = {"nodes": nn, "ms": ms, "hbm_frac_algorithmic": alg / (ms * 1e-3) / PEAK_HBM}
    dofs, ndof = mesh.dofs(2)
    d_dofs = up(dofs)
    ifbits = bits[4 * ((ne + 31) // 32):5 * ((ne + 31) // 32)]
    dmark = torch.empty((ndof + 31) // 32, dtype=torch.int32, device="cuda")
    ms = timed(lambda: ctx.mark_from_items(d_dofs, ifbits, ndof, dmark))
    nodes["p2_dofs_of_if_elements"] = {"dofs": ndof, "ms": ms,
                                       "hbm_frac_algorithmic": (ne * (40 + 1 / 8) + ndof / 8) / (ms * 1e-3) / PEAK_HBM}
    out["node_and_dof_marking"] = {"mesh": "64^3 x 6 Kuhn tets (%d)" % ne, "paths": nodes}
    # ---- several level sets (mlsetintegration.cpp) on 32^3 Kuhn tets
    m2, _ = W.config2(32)
    X = m2.coords
    l1 = W.sphere_lset(X)
    l2 = W.p1_perturb(X @ np.array([0.6, -0.5, 0.62]) - 0.31)
    d2e, d2x = up(m2.el2v), up(X)
    L = [up(l1), up(l2)]
    ml = {}
    for name, dtts in (("NEG_NEG", [[0, 0]]), ("IF_NEG", [[2, 0]]), ("IF_IF", [[2, 2]])):
        off = torch.empty(m2.ne + 1, dtype=torch.int32, device="cuda")
        ctx.mlset_rule_count(3, d2e, L, dtts, 4, off)
        ctx.sync()
        tot = int(off[-1].item())
        pts = torch.empty(max(tot, 1) * 4, dtype=torch.float64, device="cuda")

        def rule():
            ctx.mlset_rule_count(3, d2e, L, dtts, 4, off)
            ctx.mlset_rule_fill(3, d2e, d2x, L, dtts, 4, off, pts)
        ml[name] = {"points": tot, "ms": timed(rule)}
    ofb = torch.empty((m2.ne + 31) // 32, dtype=torch.int32, device="cuda")
    cob = torch.empty((m2.ne + 31) // 32, dtype=torch.int32, device="cuda")
    ml["mark_IF_IF"] = {"ms": timed(lambda: ctx.mlset_mark(3, d2e, d2x, L, [[2, 2]], of_type=ofb, contrib=cob))}
    out["several_level_sets"] = {"mesh": "32^3 x 6 Kuhn tets (%d), sphere + plane, order 4" % m2.ne, "paths": ml}
    # ---- quad / hex straight cuts on 64^3 hexahedra
    h = MakeStructured3DMesh(hexes=True, nx=64, ny=64, nz=64)
    hl = W.sphere_lset(h.coords)
    he, hx, hls = up(h.el2v), up(h.coords), up(hl)
    hdt = torch.empty(h.ne, dtype=torch.uint8, device="cuda")
    hcr = torch.empty(h.ne, dtype=torch.float64, device="cuda")
    hbits = torch.empty(8 * ((h.ne + 31) // 32), dtype=torch.int32, device="cuda")
    hcut = torch.empty(h.ne, dtype=torch.int32, device="cuda")
    hnc = torch.empty(1, dtype=torch.int32, device="cuda")
    qd = {"mark": {"ms": timed(lambda: ctx.quad_mark(3, he, hls, 1, hdt, hbits, hcr, hcut, hnc))}}
    ncut = int(hnc.item())
    cl = hcut[:ncut].clone()
    for name, dtv in (("rule_IF", 2), ("rule_NEG", 0)):
        off = torch.empty(ncut + 1, dtype=torch.int32, device="cuda")
        ctx.quad_rule_count(3, he, hls, dtv, 4, 1, off, elems=cl)
        ctx.sync()
        tot = int(off[-1].item())
        pts = torch.empty(max(tot, 1) * 4, dtype=torch.float64, device="cuda")

        def rule():
            ctx.quad_rule_count(3, he, hls, dtv, 4, 1, off, elems=cl)
            ctx.quad_rule_fill(3, he, hx, hls, dtv, 4, 1, off, pts, elems=cl)
        ms = timed(rule)
        qd[name] = {"cut_elements": ncut, "points": tot, "ms": ms, "cut_elem_per_s": ncut / (ms * 1e-3)}
    out["quad_hex"] = {"mesh": "64^3 hexahedra (%d), the 3D plain level set, order 4, FIND_OPTIMAL" % h.ne,
                       "paths": qd}
    # ---- boundary triangles with a P2 level set (a sphere centred on the
    # face z = 0), subdivision level 4
    b2v = mesh.bnd2v
    P = mesh.coords[b2v]                                                   # nse x 3 x 3
    Q = np.concatenate([P, 0.5 * (P[:, [0, 0, 1]] + P[:, [1, 2, 2]])], axis=1)
    phi = ((Q - np.array([0.5, 0.5, 0.0])) ** 2).sum(axis=2) - 0.35 ** 2   # nse x 6 nodal values
    nse = phi.shape[0]
    d_phi = up(phi)
    bdt = torch.empty(nse, dtype=torch.uint8, device="cuda")
    bcr = torch.empty(nse, dtype=torch.float64, device="cuda")
    bbits = torch.empty(8 * ((nse + 31) // 32), dtype=torch.int32, device="cuda")
    bcut = torch.empty(nse, dtype=torch.int32, device="cuda")
    bnc = torch.empty(1, dtype=torch.int32, device="cuda")
    ms = timed(lambda: ctx.mark_p2_trig(d_phi, 4, bdt, bcr, bbits, bcut, bnc))
    out["p2_boundary_marking"] = {"mesh": "boundary of the 64^3 Kuhn mesh (%d triangles), P2 sphere on a face, "
                                          "subdivision level 4" % nse,
                                  "cut": int(bnc.item()), "ms": ms, "triangles_per_s": nse / (ms * 1e-3)}
    print(json.dumps({"metric": METRIC, "config": {"workload": "secondary paths: node / dof marking, several "
                                                               "level sets, quad / hex straight cuts, P2 "
                                                               "boundary marking"},
                      "paths": out}), flush=True)
    ctx.close()


def measure_c2(args):
    """3D plain configuration (C2 of the header): fused straight-cut rule +
    P2 element matrices; returns the result dict"""
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
    odt = torch.float32 if prec == _abi.XG_PREC_FP32 else torch.float64
    dts = [_abi.XG_NEG, _abi.XG_IF]

    def timed(precision, out):
        # one step = one enqueued launch of the fused pipeline (arguments
        # built once); the K steps run back to back between two events
        step = ctx.prepare_cut_pipeline(D, p, order, d_el, d_x, d_l, dts, 3, precision, out, elems=cut_list)
        for _ in range(args.warmup):
            step()
        ctx.sync()
        torch.cuda.synchronize()
        g0 = torch.cuda.Event(enable_timing=True)
        g1 = torch.cuda.Event(enable_timing=True)
        g0.record(stream)
        for k in range(args.steps):
            step()
        g1.record(stream)
        ctx.sync()
        torch.cuda.synchronize()
        total = g0.elapsed_time(g1) / args.steps
        # a step is exactly one launch of the kernel on `stream`, so `total` is
        # its average launch duration over the timed region (the roofline's
        # figure); events around each launch in isolation add their own
        # timestamp packets to every interval and are reported beside it
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        for k in range(args.steps):
            evs[k][0].record(stream)
            step()
            evs[k][1].record(stream)
        ctx.sync()
        torch.cuda.synchronize()
        return total, float(np.mean([a.elapsed_time(b) for a, b in evs]))

    if args.pmc_probe:
        # counter runs (rocprofv3 --pmc): a 256 MiB device copy of known size
        # (calibration of FETCH_SIZE / WRITE_SIZE), then 3 launches of each path
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        for precision in (prec, prec | _abi.XG_METHOD_MOMENTS):
            o = torch.empty(n_cut * 2 * nd * nd, dtype=odt, device=dev)
            for _ in range(3):
                ctx.cut_pipeline(D, p, order, d_el, d_x, d_l, dts, 3, precision, o, elems=cut_list)
        ctx.sync()
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": 2 * a.numel() * 4, "cut_elements": n_cut,
                          "algorithmic_bytes_per_launch": bytes_alg}), flush=True)
        return

    # FP64 MFMA pipeline: the accuracy yardstick (matches the FP64 oracle to
    # 1e-12 in the parity tests)
    ref64 = torch.empty(n_cut * 2 * nd * nd, dtype=torch.float64, device=dev)
    ctx.cut_pipeline(D, p, order, d_el, d_x, d_l, dts, 3, _abi.XG_PREC_FP64, ref64, elems=cut_list)
    ctx.sync()
    r64 = ref64.view(-1, nd * nd)
    rnorm = torch.linalg.vector_norm(r64, dim=1)
    nzr = rnorm > 0

    def max_rel_err(out):
        d = torch.linalg.vector_norm(out.view(-1, nd * nd).double() - r64, dim=1)
        return float((d[nzr] / rnorm[nzr]).max().item())

    paths = {}
    out_mfma = torch.empty(n_cut * 2 * nd * nd, dtype=odt, device=dev)
    ms_mfma, iso_mfma = timed(prec, out_mfma)
    launch_mfma = ms_mfma
    paths["plain_mfma"] = {"kernel": "k_pipeline<3,2,%s>" % ("float" if prec == 1 else "double"),
                           "ms": ms_mfma, "cut_elem_per_s": n_cut / (ms_mfma * 1e-3),
                           "max_rel_err_vs_fp64": max_rel_err(out_mfma),
                           "arith": "%s MFMA, B^T W B" % args.precision}
    out_mom = torch.empty(n_cut * 2 * nd * nd, dtype=odt, device=dev)
    ms_mom, iso_mom = timed(prec | _abi.XG_METHOD_MOMENTS, out_mom)
    launch_mom = ms_mom
    paths["moments"] = {"kernel": "k_moment_pipeline<3,2,%s>" % ("float" if prec == 1 else "double"),
                        "ms": ms_mom, "cut_elem_per_s": n_cut / (ms_mom * 1e-3),
                        "max_rel_err_vs_fp64": max_rel_err(out_mom),
                        "arith": "FP64 moments + FP64 identities, %s store" % args.precision}
    # the split-BF16 MFMA contraction (XG_PREC_BF16X3) beside FP32 MFMA, on
    # P2 (the workload's order 4) and on P3 (order 6, 20 dofs) of the same cut
    # elements, each against the FP64 MFMA result
    for pp, oo in ((2, 4), (3, 6)):
        ndp = {2: 10, 3: 20}[pp]
        r64p = torch.empty(n_cut * 2 * ndp * ndp, dtype=torch.float64, device=dev)
        ctx.cut_pipeline(D, pp, oo, d_el, d_x, d_l, dts, 3, _abi.XG_PREC_FP64, r64p, elems=cut_list)
        ctx.sync()
        rr = r64p.view(-1, ndp * ndp)
        rn = torch.linalg.vector_norm(rr, dim=1)
        nzp = rn > 0
        for name, pr in (("fp32", _abi.XG_PREC_FP32), ("bf16x3", _abi.XG_PREC_BF16X3)):
            o = torch.empty(n_cut * 2 * ndp * ndp, dtype=torch.float32, device=dev)
            step = ctx.prepare_cut_pipeline(D, pp, oo, d_el, d_x, d_l, dts, 3, pr, o, elems=cut_list)
            for _ in range(2):
                step()
            ctx.sync()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(5):
                step()
            e1.record(stream)
            ctx.sync()
            msp = e0.elapsed_time(e1) / 5
            dd = torch.linalg.vector_norm(o.view(-1, ndp * ndp).double() - rr, dim=1)
            paths["mfma_p%d_%s" % (pp, name)] = {
                "kernel": "k_pipeline<3,%d,float%s>" % (pp, ",split-BF16" if name == "bf16x3" else ""),
                "ms": msp, "cut_elem_per_s": n_cut / (msp * 1e-3),
                "max_rel_err_vs_fp64": float((dd[nzp] / rn[nzp]).max().item()),
                "arith": "FP32 MFMA" if name == "fp32" else "split-BF16 hi/mid/lo MFMA (6 products), FP32 sums",
                "note": "rule order %d, FP64 sums of 8-point groups" % oo}
    best = "moments" if args.method == "moments" else "plain_mfma"
    ms_per_step = paths[best]["ms"]
    value = n_cut / (ms_per_step * 1e-3)

    peak = PEAK_FP32_MFMA if prec == _abi.XG_PREC_FP32 else PEAK_FP64_MFMA
    roof_mfma = {"bound": "mfma", "kernel": paths["plain_mfma"]["kernel"],
                 "achieved": flops / (launch_mfma * 1e-3) / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                 "frac": flops / (launch_mfma * 1e-3) / peak, "traffic": read_traffic("pipeline_c2_pmc.txt"),
                 "algorithmic_flops_per_launch": flops, "avg_launch_ms": launch_mfma,
                 "isolated_launch_ms": iso_mfma}
    ach = bytes_alg / (launch_mom * 1e-3)
    roof_mom = {"bound": "hbm", "kernel": paths["moments"]["kernel"],
                "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": ach / PEAK_HBM,
                "traffic": read_traffic("moments_c2_pmc.txt"), "algorithmic_bytes_per_launch": bytes_alg,
                "avg_launch_ms": launch_mom, "isolated_launch_ms": iso_mom,
                "fp64_valu_frac": fp64_ops_moments(npts, n_cut) / (launch_mom * 1e-3) / PEAK_FP64_VALU,
                "fp64_valu_frac_measured": read_profile_value("fp64_valu_pmc.txt", "moment_pipeline_fp64_frac_measured")}
    roof = roof_mom if best == "moments" else roof_mfma

    # ---- CPU baseline: FP64 restatement of the reference algorithm ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = cpu_threads()
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
        "vs_baseline": None, "dtype": "fp64+%s" % args.precision if best == "moments" else args.precision,
        "data": "synthetic",
        "config": {"workload": "3D plain: unit cube %d^3 x 6 Kuhn tets (%d tets), P1 sphere r=0.35 + 0.01 "
                               "sin(7x)sin(7y)sin(7z), order-4 straight cut rule NEG+IF, P2 mass+stiffness "
                               "elmat (%s)" % (args.n, ne, paths[best]["arith"]),
                   "cut_elements": n_cut, "points_neg": npts[0], "points_if": npts[2]},
        "roofline": roof, "cpu_baseline": cpu, "paths": paths,
        "roofline_plain_mfma": roof_mfma,
        "marking": {"ms": mark_ms, "tets_per_s": ne / (mark_ms * 1e-3),
                    "hbm_frac_algorithmic": ne * 26.4 / (mark_ms * 1e-3) / PEAK_HBM,
                    "k_mark_traffic": read_traffic("marking_c2_pmc.txt")},
    }
    ctx.close()
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every element is independent, so the path "
              "shards by contiguous element ranges, one range per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.config == "c4":
        return bench_c4(args)
    if args.config == "c3":
        return bench_c3(args)
    if args.config == "c2":
        print(json.dumps(measure_c2(args)), flush=True)
        return
    if args.config == "extras":
        return bench_extras(args)
    # default: the large mixed workload (the largest configuration) as the
    # headline, the 3D plain workload beside it
    res = measure_c5(args)
    if res is None:
        return
    c2 = measure_c2(args)
    res["c2"] = {k: c2[k] for k in ("value", "ms_per_step", "config", "roofline", "cpu_baseline", "paths",
                                    "roofline_plain_mfma", "marking")}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
