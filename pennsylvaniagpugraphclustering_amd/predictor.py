"""Training, threshold tuning and A/B evaluation of the active-set predictor.

The predictor (see include/pggc.h, "Active-set predictor") is trained on
move traces of the plain method: every plain sweep on a level of at least
`min_nnz` nonzeros yields the features of its active vertices and whether
each moved, and one Adam step is taken on it, on the GPU, by the library's
own forward and backward kernels.  The threshold is then chosen on separate
graphs by timing pruned runs against plain ones (the autotuner): the
fastest threshold whose modularity stays within `max_loss` of plain on every
tuning graph wins.

    python -m pennsylvaniagpugraphclustering_amd.predictor train  [--out weights.json]
    python -m pennsylvaniagpugraphclustering_amd.predictor ab --scale 24

The shipped weights (predictor_weights.json) were produced by `train` with
its defaults; the file records the graphs, seeds and settings.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import clustering as C
from . import graphs

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "predictor_weights.json")

# training graphs, seeds 1000..1063: R-MAT scale 17-19 (edge factor 16, a/b/c
# as the benchmark graphs) with a planted-partition SBM every fourth graph
# (online training drifts toward whatever family comes last, so the two are
# interleaved)
TRAIN_SPECS = [("sbm", 1 << (16 + (k // 4) % 3), 1000 + k) if k % 4 == 3 else ("rmat", 17 + k % 3, 1000 + k)
               for k in range(64)]
TUNE_SPECS = [("rmat", 20, 2000), ("rmat", 19, 2001), ("sbm", 1 << 18, 2002)]
TAUS = [-8.0, -7.0, -6.0, -5.0, -4.0, -3.0]


def make_graph(spec):
    """Device graph of a spec: ("rmat", scale, seed) or ("sbm", n, seed)."""
    kind, size, seed = spec
    if kind == "rmat":
        rp, col, _ = graphs.rmat_device(size, 16, seed=seed)
        return C.DeviceGraph(rp, col, None, None, objective="Mod")
    g = graphs.sbm(size, max(2, size // 128), 12, 4, seed=seed, permute_seed=seed)
    return C.DeviceGraph.from_csr(g, "Mod")


def save(params, tau, meta, path=WEIGHTS):
    with open(path, "w") as f:
        json.dump({"params": [float(x) for x in params], "tau": float(tau), **meta}, f, indent=1)


def load(path=WEIGHTS):
    with open(path) as f:
        d = json.load(f)
    p = np.asarray(d["params"], np.float32)
    if p.size != C.GNN_PARAMS:
        raise ValueError(f"{path}: expected {C.GNN_PARAMS} parameters, found {p.size}")
    return p, d


def enable(path=WEIGHTS, tau=None, min_nnz=None, full_every=None, ctx=None):
    """Loads trained weights and turns pruning on for later clustering calls."""
    p, d = load(path)
    C.gnn_set_params(p, ctx=ctx)
    C.gnn_configure("prune", tau=d["tau"] if tau is None else tau,
                    min_nnz=d.get("min_nnz", 1 << 20) if min_nnz is None else min_nnz,
                    full_every=d.get("full_every", 0) if full_every is None else full_every, ctx=ctx)
    return d


def train(specs=TRAIN_SPECS, epochs=1, lr=2e-3, pos_weight=4.0, min_nnz=1 << 16, algo="louvain", ex_iters=-1,
          ctx=None, log=print):
    """Online training on the plain method's sweeps; returns (params, history)."""
    C.gnn_configure("train", tau=0.0, min_nnz=min_nnz, lr=lr, pos_weight=pos_weight, ctx=ctx)
    hist = []
    for ep in range(epochs):
        for spec in specs:
            dg = make_graph(spec)
            C.gnn_reset_stats(ctx=ctx)
            t0 = time.time()
            q, _, info = C.cluster(dg, algo, "Mod", 1.0, ex_iters, seed=spec[2], ctx=ctx)
            torch.cuda.synchronize()
            st = C.gnn_stats(ctx=ctx)
            steps = max(st["train_steps"], 1)
            rec = {"epoch": ep, "graph": list(spec), "q": q, "steps": st["train_steps"],
                   "loss": st["loss_sum"] / steps, "recall": st["tp"] / max(st["tp"] + st["fn"], 1),
                   "kept": (st["tp"] + st["fp"]) / max(st["tp"] + st["fp"] + st["fn"] + st["tn"], 1),
                   "seconds": time.time() - t0}
            hist.append(rec)
            log(f"train ep {ep} {spec} steps {rec['steps']} loss {rec['loss']:.4f} recall@0 {rec['recall']:.3f} "
                f"kept@0 {rec['kept']:.3f} {rec['seconds']:.2f}s")
            del dg
    params = C.gnn_get_params(ctx=ctx)
    C.gnn_configure("off", ctx=ctx)
    return params, hist


def timed_cluster(dg, algo, ex_iters, seed, ctx=None):
    torch.cuda.synchronize()
    t0 = time.time()
    q, part, info = C.cluster(dg, algo, "Mod", 1.0, ex_iters, seed=seed, ctx=ctx)
    torch.cuda.synchronize()
    return time.time() - t0, q, part, info


def ab(dg, params, tau, min_nnz=1 << 16, full_every=0, algo="louvain", ex_iters=-1, seed=0, repeats=1, ctx=None):
    """Plain versus pruned run on one graph: times (best of `repeats`),
    modularities and the predictor's counters."""
    C.gnn_configure("off", ctx=ctx)
    tp = min(timed_cluster(dg, algo, ex_iters, seed, ctx)[0] for _ in range(repeats))
    _, qp, _, ip = timed_cluster(dg, algo, ex_iters, seed, ctx)
    C.gnn_set_params(params, ctx=ctx)
    C.gnn_configure("prune", tau=tau, min_nnz=min_nnz, full_every=full_every, ctx=ctx)
    times = []
    for _ in range(repeats):
        C.gnn_reset_stats(ctx=ctx)
        t, qg, _, ig = timed_cluster(dg, algo, ex_iters, seed, ctx)
        times.append(t)
    st = C.gnn_stats(ctx=ctx)
    C.gnn_configure("off", ctx=ctx)
    return {"plain_s": tp, "gnn_s": min(times), "speedup": tp / min(times), "plain_q": qp, "gnn_q": qg,
            "dq": qg - qp, "plain_sweeps": ip["sweeps"], "gnn_sweeps": ig["sweeps"],
            "pruned_frac": st["pruned"] / max(st["active"], 1), "restores": st["restores"]}


def tune(params, specs=TUNE_SPECS, taus=TAUS, min_nnz=1 << 16, max_loss=1e-3, algo="louvain", ex_iters=-1,
         ctx=None, log=print):
    """Fastest threshold whose modularity stays within max_loss of plain on
    every tuning graph (geometric-mean speedup)."""
    rows = []
    graphs_dev = [(s, make_graph(s)) for s in specs]
    best = (None, 0.0)
    for tau in taus:
        res = [ab(dg, params, tau, min_nnz, 0, algo, ex_iters, seed=s[2], ctx=ctx) for s, dg in graphs_dev]
        gm = float(np.exp(np.mean([np.log(r["speedup"]) for r in res])))
        ok = all(r["dq"] >= -max_loss for r in res)
        rows.append({"tau": tau, "geomean_speedup": gm, "ok": ok, "runs": res})
        log(f"tune tau {tau}: speedup {gm:.3f} ok {ok} dq {[round(r['dq'], 6) for r in res]} "
            f"pruned {[round(r['pruned_frac'], 3) for r in res]}")
        if ok and gm > best[1]:
            best = (tau, gm)
    return best[0], rows


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pennsylvaniagpugraphclustering_amd.predictor")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="train on plain traces, tune the threshold, write the weights")
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--epochs", type=int, default=1)
    t.add_argument("--graphs", type=int, default=len(TRAIN_SPECS), help="use the first k training graphs")
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--pos-weight", type=float, default=4.0)
    t.add_argument("--min-nnz", type=int, default=1 << 16)
    a = sub.add_parser("ab", help="plain versus pruned on an R-MAT graph")
    a.add_argument("--weights", default=WEIGHTS)
    a.add_argument("--scale", type=int, default=22)
    a.add_argument("--seed", type=int, default=2)
    a.add_argument("--algo", default="louvain")
    a.add_argument("--tau", type=float, default=None)
    a.add_argument("--repeats", type=int, default=1)
    args = ap.parse_args(argv)
    if args.cmd == "train":
        specs = TRAIN_SPECS[: args.graphs]
        params, hist = train(specs, args.epochs, args.lr, args.pos_weight, args.min_nnz)
        tau, rows = tune(params, min_nnz=args.min_nnz)
        if tau is None:
            tau = min(TAUS)
            print("no threshold kept modularity within bounds on every tuning graph; using the lowest", flush=True)
        meta = {"min_nnz": args.min_nnz, "full_every": 0,
                "training": {"graphs": [list(s) for s in specs], "epochs": args.epochs, "lr": args.lr,
                             "pos_weight": args.pos_weight, "algo": "louvain", "ex_iters": -1,
                             "final_loss": float(np.mean([h["loss"] for h in hist[-len(specs):]]))},
                "tuning": {"graphs": [list(s) for s in TUNE_SPECS], "taus": TAUS, "max_modularity_loss": 1e-3,
                           "rows": [{k: v for k, v in r.items() if k != "runs"} for r in rows]}}
        save(params, tau, meta, args.out)
        print(f"wrote {args.out} tau {tau}", flush=True)
    else:
        p, d = load(args.weights)
        tau = d["tau"] if args.tau is None else args.tau
        dg = make_graph(("rmat", args.scale, args.seed))
        r = ab(dg, p, tau, d.get("min_nnz", 1 << 16), d.get("full_every", 0), args.algo, -1, args.seed,
               args.repeats)
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    sys.exit(main())
