"""Benchmark: end-to-end pLouvain to stable (ex_iters = -1) on the R-MAT
scale-24 graph (edge factor 16, a/b/c/d = .57/.19/.19/.05, symmetrised,
deduplicated, no self loops, unit weights, int64 row offsets, int32 columns,
seed 2), one MI355X.

One step = one full clustering run (multilevel local moving, contraction and
the ex_iters loop) from the device-resident CSR to the device-resident
clustering.  value = undirected edges / seconds per step.

Prints one JSON line.  The cpu_baseline leg times the CPU oracle (a serial
restatement, 1 core) on a bounded sample (an R-MAT graph of smaller scale)
and reports its edges/s.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of a timed region, measured with rocprofv3 FETCH_SIZE
# and WRITE_SIZE counters scaled by a plain copy kernel of known size
# (profiles/rmat24_afterburner_traffic.md); keyed by (region, R-MAT scale)
MEASURED_TRAFFIC = {("afterburner", 24): 12.81e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--scale", type=int, default=24)
    p.add_argument("--algo", default="louvain", choices=["louvain", "leiden", "leiden+"])
    p.add_argument("--ex-iters", type=int, default=-1)
    p.add_argument("--cpu-scale", type=int, default=16)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-leiden", action="store_true", help="(default) skip the pLeiden+ report")
    p.add_argument("--no-gnn", action="store_true", help="skip the active-set predictor A/B step")
    p.add_argument("--batch", type=int, default=512,
                   help="batch leg: this many planted-partition graphs (n log-uniform in [2^10, 2^18]), pLouvain "
                        "ex_iters 0, one by one and as one batch (0 skips it)")
    p.add_argument("--batch-streams", type=int, default=8)
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 1 GiB read + 1 GiB write elementwise kernel 5 times (the reference point "
                        "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    p.add_argument("--knob", action="append", default=[], metavar="NAME=VALUE",
                   help="set a tuning knob of the HIP path before the run (hub_update_div, gnn_max_sweeps); "
                        "repeatable")
    p.add_argument("--leiden", type=int, default=0, metavar="ITERS",
                   help="also run pLeiden+ with ex_iters=ITERS (to stable it creeps up by ~1e-8 per outer "
                        "iteration for hundreds of iterations at scale 24)")
    return p.parse_args()


def cpu_baseline(scale, seed=2):
    """Serial CPU oracle on an R-MAT graph of the given scale (same
    construction, numpy stream): pLouvain to stable, edges/s."""
    from oracle import oracle as O
    from pennsylvaniagpugraphclustering_amd import graphs
    g = graphs.rmat(scale, 16, seed=seed)
    vw = g.vertex_weights("Mod")
    t0 = time.perf_counter()
    q, part, info = O.cluster(g.rp, g.col, None, vw, "Mod", 1.0, "louvain", -1, seed=0)
    dt = time.perf_counter() - t0
    return {"value": (g.nnz // 2) / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"R-MAT scale {scale} (n={g.n}, m={g.nnz // 2}), pLouvain ex_iters=-1, "
                      f"{dt:.2f} s, Q={q:.6f}"}


def roofline(kstats, scale):
    """Roofline object for the dominant timed kernel group that carries an
    algorithmic byte count (HIP events on the library stream)."""
    best = None
    for name, (n, ms, by) in kstats.items():
        if n and by > 0 and (best is None or ms > kstats[best][1]):
            best = name
    if best is None:
        return None
    n, ms, by = kstats[best]
    achieved = by / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": best, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get((best, scale)), "launches": n,
            "avg_launch_us": ms / n * 1000.0, "bytes_per_launch": by / n}


def batch_leg(a, ctx, C, graphs, torch):
    """Many small instances: `a.batch` planted-partition graphs with seeds
    100.., n = round(2^U(10,18)), blocks uniform in [n/256, n/32], expected
    intra degree 10 and inter degree 3 (graphs.sbm_torch), pLouvain ex_iters
    0.  Timed one by one on the context, then as one pggc_cluster_batch call
    on a.batch_streams streams; the objectives must agree."""
    import numpy as np
    rng = np.random.default_rng(100)
    bg = []
    for k in range(a.batch):
        n = int(round(2 ** rng.uniform(10, 18)))
        b = int(rng.integers(max(1, -(-n // 256)), max(1, n // 32) + 1))
        rp_, col_ = graphs.sbm_torch(n, b, 10, 3, seed=100 + k)
        bg.append(C.DeviceGraph(rp_, col_, None, None, objective="Mod", ctx=ctx))
    edges = sum(x.nnz // 2 for x in bg)
    ctx.set_verbose(0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    seq = [C.cluster(x, "louvain", "Mod", 1.0, 0, seed=k, ctx=ctx) for k, x in enumerate(bg)]
    torch.cuda.synchronize()
    ts = time.perf_counter() - t1
    t2 = time.perf_counter()
    res = C.cluster_batch(bg, "louvain", "Mod", 1.0, 0, seeds=list(range(len(bg))), streams=a.batch_streams, ctx=ctx)
    tb = time.perf_counter() - t2
    ctx.set_verbose(1)
    return {"graphs": len(bg), "vertices": sum(x.n for x in bg), "edges": edges, "one_by_one_s": ts,
            "batch_s": tb, "streams": a.batch_streams, "speedup": ts / tb, "batch_edges_per_s": edges / tb,
            "objectives_equal": all(r[0] == q[0] for r, q in zip(res, seq)),
            "mean_Q": float(np.mean([r[0] for r in res]))}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    if a.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"read_bytes": 1 << 30, "write_bytes": 1 << 30, "repeats": 5}}))
        return
    from pennsylvaniagpugraphclustering_amd import graphs
    from pennsylvaniagpugraphclustering_amd import clustering as C

    dev = torch.device("cuda")
    t0 = time.perf_counter()
    rp, col, _ = graphs.rmat_torch(a.scale, 16, seed=2, device=dev)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    ctx = C.default_context()
    ctx.set_verbose(1)
    for kv in a.knob:
        name, value = kv.split("=")
        ctx.set_knob(name, int(value))
        print(f"# knob {name} = {ctx.get_knob(name)}", file=sys.stderr, flush=True)
    g = C.DeviceGraph(rp, col, None, None, objective="Mod", ctx=ctx)
    m = g.nnz // 2
    print(f"# graph R-MAT s{a.scale}: n={g.n} nnz={g.nnz} m={m} (generated in {gen_s:.1f}s)", file=sys.stderr,
          flush=True)

    def step():
        return C.cluster(g, a.algo, "Mod", 1.0, a.ex_iters, seed=0, ctx=ctx)

    for w in range(a.warmup):
        q, part, info = step()
        print(f"# warmup {w}: Q={q:.6f} info={info}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for s in range(a.steps):
        q, part, info = step()
        print(f"# step {s}: Q={q:.6f}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    kstats = ctx.kernel_stats()
    ctx.set_timing(False)
    for name, (n, ms, by) in kstats.items():
        if n:
            print(f"# kernel group {name}: {n} launches, {ms / max(a.steps, 1):.1f} ms/step, "
                  f"{ms / n * 1000:.1f} us avg" + (f", {by / (ms * 1e-3) / 1e9:.0f} GB/s" if by else ""),
                  file=sys.stderr, flush=True)

    extra = {}
    if a.leiden and not a.no_leiden:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ql, _, infol = C.cluster(g, "leiden+", "Mod", 1.0, a.leiden, seed=0, ctx=ctx)
        torch.cuda.synchronize()
        extra["leiden_plus"] = {"ex_iters": a.leiden, "seconds": time.perf_counter() - t1, "Q": ql,
                                "labels": infol["labels"], "sweeps": infol["sweeps"]}
        print(f"# pLeiden+: {extra['leiden_plus']}", file=sys.stderr, flush=True)

    if not a.no_gnn:
        # A/B: one more step with the active-set predictor pruning (shipped
        # weights and threshold), against the plain steps timed above
        from pennsylvaniagpugraphclustering_amd import predictor
        d = predictor.enable(ctx=ctx)
        C.gnn_reset_stats(ctx=ctx)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        qg, _, ig = step()
        torch.cuda.synchronize()
        tg = time.perf_counter() - t1
        st = C.gnn_stats(ctx=ctx)
        C.gnn_configure("off", ctx=ctx)
        extra["gnn"] = {"tau": d["tau"], "min_nnz": d.get("min_nnz"), "seconds": tg, "plain_seconds": dt,
                        "speedup": dt / tg, "Q": qg, "dQ_vs_plain": qg - q, "sweeps": ig["sweeps"],
                        "outer_iterations": ig["outer"], "predicted_sweeps": st["sweeps"],
                        "pruned_fraction": st["pruned"] / max(st["active"], 1), "restores": st["restores"]}
        print(f"# predictor A/B: {extra['gnn']}", file=sys.stderr, flush=True)

    if a.batch > 0:
        extra["batch"] = batch_leg(a, ctx, C, graphs, torch)
        print(f"# batch: {extra['batch']}", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a.cpu_scale)
        print(f"# cpu baseline: {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": "end-to-end clustering edges/s (pLouvain to stable)",
        "value": m / dt,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1000.0,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+fp32",
        "data": "synthetic",
        "config": {"workload": f"R-MAT scale {a.scale} ef16 (.57,.19,.19,.05) seed 2, unit weights, "
                               f"{a.algo} ex_iters={a.ex_iters}, objective Mod",
                   "n": g.n, "nnz": g.nnz, "undirected_edges": m},
        "modularity": q,
        "labels": info["labels"],
        "sweeps": info["sweeps"],
        "levels": info["levels"],
        "outer_iterations": info["outer"],
        "edges_per_s_per_sweep": m * info["sweeps"] / dt,
        "roofline": roofline(kstats, a.scale),
        "kernel_groups_ms_per_step": {k: v[1] / max(a.steps, 1) for k, v in kstats.items() if v[0]},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
