"""Autotuner for the batched persistent pipeline (xg_batch).

For each H1 degree p the element-matrix method (MFMA contraction or moment
form) and the launch shape of the persistent degree-p kernel -- waves per
block, resident blocks per CU, cut elements per work grab -- are swept on the
batched workload; every candidate is timed with HIP events on
the context's stream (median of `reps` eager runs of that kernel alone) and
the fastest is written to tuning/<name>.json, which bench.py and
ngsxfem_amd.batch read.

    python -m ngsxfem_amd.autotune [--nmesh 256] [--out tuning/c3_gfx950.json]
"""
import argparse
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuning", "c3_gfx950.json")

CANDIDATES = [(w, b, g) for w in (1, 2) for b in (1, 2, 4, 8) for g in (1, 4)]
METHODS = {"mfma": 0, "moments": 16}


def load(path=DEFAULT_OUT):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def apply(batch, tuning):
    if not tuning:
        return
    for p in (1, 2, 3):
        t = tuning.get("p%d" % p)
        if t:
            batch.tune(p, t["waves"], t["blocks_per_cu"], t["grab"])
            batch.method(METHODS[t.get("method", "mfma")], p)


def tune(batch, out, ctx, reps=3, log=print):
    import torch
    stream = torch.cuda.current_stream()
    res = {}
    for p in (1, 2, 3):
        if not np.any(batch.mesh_p == p):
            continue
        best = None
        for mname, (w, b, g) in [(m, c) for m in METHODS for c in CANDIDATES]:
            batch.method(METHODS[mname], p)
            batch.tune(p, w, b, g)
            times = []
            for _ in range(reps + 1):
                batch.plan()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                batch.run(out, use_graph=False, only_p=p)
                e1.record(stream)
                ctx.sync()
                times.append(e0.elapsed_time(e1))
            t = float(np.median(times[1:]))
            if log:
                log("p=%d %s waves=%d blocks/CU=%d grab=%d: %.3f ms" % (p, mname, w, b, g, t))
            if best is None or t < best[0]:
                best = (t, w, b, g, mname)
        res["p%d" % p] = {"method": best[4], "waves": best[1], "blocks_per_cu": best[2], "grab": best[3],
                          "ms": best[0]}
        batch.method(METHODS[best[4]], p)
        batch.tune(p, best[1], best[2], best[3])
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nmesh", type=int, default=256)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    import torch

    from . import _abi
    from . import workloads as W
    coords, el2v, lset, off, ps, ns, axes = W.config3(args.nmesh)
    ctx = _abi.Context(0, stream=torch.cuda.current_stream())
    d_x = torch.from_numpy(coords).cuda()
    d_e = torch.from_numpy(el2v).cuda()
    d_l = torch.from_numpy(lset).cuda()
    batch = _abi.Batch(ctx, d_x, d_e, d_l, off, ps)
    cnt, oo = batch.plan()
    out = torch.empty(int(oo[-1]) + 1, dtype=torch.float32, device="cuda")
    res = tune(batch, out, ctx)
    res["workload"] = {"nmesh": args.nmesh, "tets": int(off[-1]), "cut": int(cnt.sum())}
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))
    batch.close()
    ctx.close()


if __name__ == "__main__":
    main()
