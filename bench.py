"""Benchmark: end-to-end pLouvain to stable (ex_iters = -1) on the R-MAT
scale-24 graph (edge factor 16, a/b/c/d = .57/.19/.19/.05, seed 2,
symmetrised, deduplicated, no self loops, unit weights, int64 row offsets,
int32 columns), one MI355X.  The graph is built on the GPU by the library's
counter-based generator; the CPU oracle builds the identical graph.

One step = one full clustering run (multilevel local moving, contraction and
the ex_iters loop) from the device-resident CSR to the device-resident
clustering.  value = undirected edges / seconds per step.

Also reported on the same graph:
- edges_per_s_per_sweep: m / the time of the first full top-level sweep
  (the table sweep after the raw sweep of the first local-move call);
- cpu_baseline: the oracle (the OpenMP CPU restatement, all host threads) on
  the same graph copied from HBM, timed on the same full top-level sweep
  (its call stopped after two sweeps), value m / that time;
- roofline of the kernel group that takes the most time, from HIP events
  around its launches and the algorithmic bytes the library counts for it;
- the pLeiden+ leg (--leiden ITERS, a bounded ex_iters), the active-set
  predictor A/B leg and the batch leg.

Prints one JSON line.
"""
import argparse
import hashlib
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GOLDEN = os.path.join(ROOT, "tests", "golden", "rmat{scale}_seed2.json")
TRAFFIC = os.path.join(ROOT, "profiles", "traffic_rmat{scale}.json")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--scale", type=int, default=24)
    p.add_argument("--algo", default="louvain", choices=["louvain", "leiden", "leiden+"])
    p.add_argument("--ex-iters", type=int, default=-1)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: all host threads, at most 16)")
    p.add_argument("--no-gnn", action="store_true", help="skip the active-set predictor A/B step")
    p.add_argument("--batch", type=int, default=512,
                   help="batch leg: this many planted-partition graphs (n log-uniform in [2^10, 2^18]), pLouvain "
                        "ex_iters 0, one by one and as one batch (0 skips it)")
    p.add_argument("--batch-streams", type=int, default=8)
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 1 GiB read + 1 GiB write elementwise kernel 5 times (the reference point "
                        "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters) and exit")
    p.add_argument("--kernels-only", action="store_true",
                   help="only run one step with the kernel-group timers and byte counters on and print them (the "
                        "command the rocprofv3 kernel summaries and counter passes are taken from)")
    p.add_argument("--knob", action="append", default=[], metavar="NAME=VALUE",
                   help="set a tuning knob of the HIP path before the run (repeatable)")
    p.add_argument("--lfr-n", type=int, default=8000000,
                   help="LFR-style leg: n vertices (power-law degrees: exponent 2.5, mean 20, max 1000; community "
                        "sizes: exponent 1.5 in [20, 2000]; mixing 0.4; weights uniform in [1, 100]; seed 3), "
                        "pLeiden with weighted modularity, plain and with the active-set predictor (0 skips it)")
    p.add_argument("--leiden", type=int, default=2, metavar="ITERS",
                   help="pLeiden+ leg with ex_iters=ITERS (-1: to stable, which creeps up by ~1e-8 per outer "
                        "iteration for hundreds of iterations at scale 24; 0 or less than -1 skips it)")
    p.add_argument("--c5", action="store_true",
                   help="only run the largest single-GPU workload and exit: R-MAT scale 27, edge factor 16, weights "
                        "uniform in [1, 10], seed 4 (4.2e9 nonzeros) built on the GPU, one pLeiden+ pass and one "
                        "pLouvain pass (ex_iters 0, weighted modularity), each checked against the evaluator")
    p.add_argument("--c4b", action="store_true",
                   help="only run the seeded-trials workload and exit: 128 pLeiden+ trials (ex_iters 0, seeds "
                        "0..127) on one R-MAT scale-22 graph (edge factor 16, seed 2), one by one and as one "
                        "pggc_cluster_batch call, every trial checked against its single run")
    return p.parse_args()


def sha(t):
    return hashlib.sha256(t.detach().cpu().numpy().tobytes()).hexdigest()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def first_sweep_times(C, g, ctx):
    """GPU: the first local-move call from singletons on the top level, with
    the per-sweep trace; returns the sweep records."""
    recs = []
    ctx.set_sweep_trace(recs)
    try:
        st, vol = C.state_init(g, "Mod", 1.0, ctx=ctx)
        import torch
        part = torch.arange(g.n, dtype=torch.int32, device="cuda")
        C.local_move(g, part, st, True, ctx=ctx)
    finally:
        ctx.set_sweep_trace(None)
    return recs


def cpu_baseline(rp, col, threads):
    """The oracle (OpenMP restatement) on the same graph: the first local-move
    call from singletons on the top level, stopped after two sweeps (the raw
    sweep and the first full table sweep); value = m / the full sweep's time."""
    import numpy as np
    from oracle import oracle as O
    n = rp.numel() - 1
    hrp = rp.cpu().numpy()
    hcol = col.cpu().numpy()
    vw = np.diff(hrp).astype(np.int64)
    m = int(hrp[-1]) // 2
    nthreads = threads or min(os.cpu_count() or 1, 16)
    O.set_threads(nthreads)
    O.set_sweep_limit(2)
    try:
        lam = 1.0 / float(hrp[-1])
        obj0 = -lam * float(np.dot(vw.astype(np.float64), vw.astype(np.float64)))
        t0 = time.perf_counter()
        _, _, _, _, _, tr = O.local_move_trace(hrp, hcol, None, vw, np.arange(n, dtype=np.int32), vw, n, 0, lam,
                                               obj0, True)
        total = time.perf_counter() - t0
    finally:
        O.set_sweep_limit(0)
    t_full = float(tr[1, 4]) if tr.shape[0] > 1 else float("nan")
    return {"value": m / t_full, "unit": "edges/s per full top-level sweep", "cores": nthreads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"oracle (OpenMP CPU restatement) on the same R-MAT graph (n={n}, m={m}): first top-level "
                      f"local-move call of pLouvain stopped after 2 sweeps; raw sweep {tr[0, 4]:.2f} s, full "
                      f"table sweep {t_full:.2f} s ({int(tr[1, 1]) if tr.shape[0] > 1 else 0} candidates); "
                      f"{total:.1f} s with set-up"}


def roofline(kstats, scale):
    """Roofline object for the kernel group that takes the most time (HIP
    events on the library stream, algorithmic bytes counted by the library).
    traffic: HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes recorded in profiles/ for this group and workload, if present."""
    best = None
    for name, (n, ms, by) in kstats.items():
        if n and by > 0 and (best is None or ms > kstats[best][1]):
            best = name
    if best is None:
        return None
    n, ms, by = kstats[best]
    achieved = by / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": best, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "launches": n, "avg_launch_us": ms / n * 1000.0,
           "bytes_per_launch": by / n}
    path = TRAFFIC.format(scale=scale)
    if os.path.exists(path):
        t = json.load(open(path)).get(best)
        if t:
            out["traffic"] = t["hbm_bytes_per_launch"]
            out["traffic_source"] = {k: t[k] for k in ("commit", "profile") if k in t}
    return out


def instrumented_step(step, ctx, torch):
    """One step with the library's kernel-group timers (HIP events on its
    stream) and algorithmic byte counters on; returns the raw stats, the
    per-group summary and the step's wall time."""
    ctx.set_timing(True)
    t1 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t1
    kstats = ctx.kernel_stats()
    ctx.set_timing(False)
    groups = {}
    for name, (n, ms, by) in kstats.items():
        if n:
            groups[name] = {"ms_per_step": ms, "launches_per_step": n, "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None,
                            "GB_per_step": by / 1e9}
            print(f"# kernel group {name}: {n} launches, {ms:.1f} ms/step, {ms / n * 1000:.1f} us avg, "
                  f"{groups[name]['GBps']:.0f} GB/s", file=sys.stderr, flush=True)
    return kstats, groups, seconds


def lfr_leg(a, ctx, C, graphs, torch):
    """The LFR-style workload: pLeiden (ex_iters 0) with weighted modularity
    on a graph built on the GPU, timed plain (best of 2) and with the
    active-set predictor (shipped weights and threshold)."""
    t0 = time.perf_counter()
    rp, col, w = graphs.lfr_device(a.lfr_n, 2.5, 20.0, 1000, 1.5, 20, 2000, 0.4, weights=(1, 100), seed=3, ctx=ctx)
    g = C.DeviceGraph(rp, col, w, None, objective="WMod", ctx=ctx)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        q, _, info = C.cluster(g, "leiden", "WMod", 1.0, 0, seed=0, ctx=ctx)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
    out = {"n": g.n, "nnz": g.nnz, "undirected_edges": g.nnz // 2, "generate_s": gen_s, "seconds": min(times),
           "edges_per_s": (g.nnz // 2) / min(times), "Q": q, "labels": info["labels"], "sweeps": info["sweeps"],
           "levels": info["levels"]}
    if not a.no_gnn:
        from pennsylvaniagpugraphclustering_amd import predictor
        d = predictor.enable(ctx=ctx)
        C.gnn_reset_stats(ctx=ctx)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        qg, _, ig = C.cluster(g, "leiden", "WMod", 1.0, 0, seed=0, ctx=ctx)
        torch.cuda.synchronize()
        tg = time.perf_counter() - t1
        st = C.gnn_stats(ctx=ctx)
        C.gnn_configure("off", ctx=ctx)
        out["gnn"] = {"tau": d["tau"], "seconds": tg, "speedup": min(times) / tg, "Q": qg, "dQ_vs_plain": qg - q,
                      "sweeps": ig["sweeps"], "pruned_fraction": st["pruned"] / max(st["active"], 1)}
    del g, rp, col, w
    return out


def batch_leg(a, ctx, C, graphs, torch):
    """Many small instances: `a.batch` planted-partition graphs with seeds
    100.., n = round(2^U(10,18)), blocks uniform in [n/256, n/32], expected
    intra degree 10 and inter degree 3 (graphs.sbm_torch), pLouvain ex_iters
    0.  Timed one by one on the context, then as one pggc_cluster_batch call
    (one disjoint union graph); the objectives must agree."""
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
            "batch_s": tb, "mode": "one disjoint union graph: each kernel launched once over all instances",
            "speedup": ts / tb, "batch_edges_per_s": edges / tb,
            "objectives_equal": all(r[0] == q[0] for r, q in zip(res, seq)),
            "mean_Q": float(np.mean([r[0] for r in res]))}


def c5_leg(C, graphs, torch):
    """The largest single-GPU workload: R-MAT scale 27 with weights in
    [1, 10] (seed 4).  pLeiden+ to stable is not run (each outer iteration
    is a full pass of about half a minute here); one pass of pLeiden+ and of
    pLouvain, each with its objective recomputed by the evaluator."""
    ctx = C.default_context()
    t0 = time.perf_counter()
    rp, col, w = graphs.rmat_device(27, 16, seed=4, weights=(1, 10), ctx=ctx)
    torch.cuda.synchronize()
    out = {"workload": "R-MAT scale 27 ef16 (.57,.19,.19,.05) seed 4, weights uniform in [1, 10], WMod",
           "n": int(rp.numel() - 1), "nnz": int(col.numel()), "generate_s": time.perf_counter() - t0, "runs": {}}
    g = C.DeviceGraph(rp, col, w, None, objective="WMod", ctx=ctx)
    for algo in ("leiden+", "louvain"):
        print(f"# c5 {algo} ex_iters=0", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        q, part, info = C.cluster(g, algo, "WMod", 1.0, 0, seed=1, ctx=ctx)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t1
        qe = C.modularity(g, part, "WMod", ctx=ctx)
        out["runs"][algo] = {"seconds": sec, "edges_per_s": (g.nnz // 2) / sec, "Q": q, "evaluator_Q": qe,
                             "matches_evaluator": q == qe, **info}
        print(f"# c5 {algo}: {sec:.2f} s Q={q!r} info={info}", file=sys.stderr, flush=True)
        del part
    out["free_gib_after"] = torch.cuda.mem_get_info()[0] / 2 ** 30
    return out


def c4b_leg(C, graphs, torch, trials=128):
    """Seeded pLeiden+ trials on one graph (the reference's -trials loop,
    app/cluster_main.cpp:189-204): one by one, then as one batch call, whose
    workers share the context's stream (kernels of one trial at a time)."""
    ctx = C.default_context()
    rp, col, _ = graphs.rmat_device(22, 16, seed=2, ctx=ctx)
    g = C.DeviceGraph(rp, col, None, None, objective="Mod", ctx=ctx)
    seeds = list(range(trials))
    C.cluster(g, "leiden+", "Mod", 1.0, 0, seed=0, ctx=ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    single = [C.cluster(g, "leiden+", "Mod", 1.0, 0, seed=s, ctx=ctx)[0] for s in seeds]
    torch.cuda.synchronize()
    one_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = C.cluster_batch([g] * trials, "leiden+", "Mod", 1.0, 0, seeds=seeds, streams=8, ctx=ctx)
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t0
    return {"workload": f"{trials} pLeiden+ trials (ex_iters 0, seeds 0..{trials - 1}) on R-MAT scale 22 ef16 seed 2",
            "n": g.n, "nnz": g.nnz, "one_by_one_s": one_s, "batch_s": batch_s, "speedup": one_s / batch_s,
            "mode": "worker host threads sharing the context's stream",
            "objectives_equal": all(r[0] == q for r, q in zip(res, single)),
            "best_Q": max(single), "mean_Q": sum(single) / len(single)}


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
    if a.c5:
        print(json.dumps({"c5": c5_leg(C, graphs, torch)}))
        return
    if a.c4b:
        print(json.dumps({"c4b": c4b_leg(C, graphs, torch)}))
        return

    t0 = time.perf_counter()
    ctx = C.default_context()
    rp, col, _ = graphs.rmat_device(a.scale, 16, seed=2, ctx=ctx)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    ctx.set_verbose(1)
    for kv in a.knob:
        name, value = kv.split("=")
        ctx.set_knob(name, int(value))
        print(f"# knob {name} = {ctx.get_knob(name)}", file=sys.stderr, flush=True)
    g = C.DeviceGraph(rp, col, None, None, objective="Mod", ctx=ctx)
    m = g.nnz // 2
    print(f"# graph R-MAT s{a.scale}: n={g.n} nnz={g.nnz} m={m} (built on the GPU in {gen_s:.1f}s)",
          file=sys.stderr, flush=True)

    def step():
        return C.cluster(g, a.algo, "Mod", 1.0, a.ex_iters, seed=0, ctx=ctx)

    if a.kernels_only:
        kstats, groups, instrumented_s = instrumented_step(step, ctx, torch)
        print(json.dumps({"kernels_only": True, "workload": f"R-MAT scale {a.scale}, {a.algo} ex_iters={a.ex_iters}",
                          "instrumented_step_s": instrumented_s, "roofline": roofline(kstats, a.scale),
                          "kernel_groups": groups}))
        return
    for w in range(a.warmup):
        q, part, info = step()
        print(f"# warmup {w}: Q={q:.6f} info={info}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(a.steps):
        q, part, info = step()
        print(f"# step {s}: Q={q:.6f}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    # one more step with the kernel timers and byte counters on (they cost
    # time of their own, so the timed steps above run without them)
    kstats, groups, instrumented_s = instrumented_step(step, ctx, torch)

    extra = {}
    golden = GOLDEN.format(scale=a.scale)
    if os.path.exists(golden) and a.algo == "louvain" and a.ex_iters == -1:
        gr = json.load(open(golden)).get("runs", {}).get("louvain_ex-1")
        if gr:
            extra["matches_oracle"] = (sha(part) == gr["sha_part"] and float(q).hex() == gr["objective_hex"])
            extra["oracle_seconds"] = gr.get("oracle_seconds")
            extra["oracle_threads"] = gr.get("oracle_threads")

    recs = first_sweep_times(C, g, ctx)
    full = [r for r in recs if r["n"] == g.n and r["sweep"] == 1]
    sweep_s = full[0]["us"] * 1e-6 if full else None
    print(f"# first top-level sweeps: {recs[:3]}", file=sys.stderr, flush=True)

    if a.leiden > 0 or a.leiden == -1:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ql, partl, infol = C.cluster(g, "leiden+", "Mod", 1.0, a.leiden, seed=0, ctx=ctx)
        torch.cuda.synchronize()
        tl = time.perf_counter() - t1
        extra["leiden_plus"] = {"ex_iters": a.leiden, "seconds": tl, "Q": ql, "labels": infol["labels"],
                                "sweeps": infol["sweeps"], "outer_iterations": infol["outer"],
                                "time_vs_plouvain": tl / dt}
        if os.path.exists(golden):
            gr = json.load(open(golden)).get("runs", {}).get(f"leiden+_ex{a.leiden}")
            if gr:
                extra["leiden_plus"]["matches_oracle"] = (sha(partl) == gr["sha_part"] and
                                                          float(ql).hex() == gr["objective_hex"])
        # one more pLeiden+ run with the kernel timers and byte counters on:
        # the refinement's group (leiden) and the others of this algorithm
        _, lgroups, _ = instrumented_step(lambda: C.cluster(g, "leiden+", "Mod", 1.0, a.leiden, seed=0, ctx=ctx),
                                          ctx, torch)
        extra["leiden_plus"]["kernel_groups"] = lgroups
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

    if a.lfr_n > 0:
        extra["lfr"] = lfr_leg(a, ctx, C, graphs, torch)
        print(f"# LFR leg: {extra['lfr']}", file=sys.stderr, flush=True)

    if a.batch > 0:
        extra["batch"] = batch_leg(a, ctx, C, graphs, torch)
        print(f"# batch: {extra['batch']}", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(rp, col, a.cpu_threads)
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
        "config": {"workload": f"R-MAT scale {a.scale} ef16 (.57,.19,.19,.05) seed 2 (Philox4x32-10 draws), "
                               f"unit weights, {a.algo} ex_iters={a.ex_iters}, objective Mod",
                   "n": g.n, "nnz": g.nnz, "undirected_edges": m},
        "modularity": q,
        "labels": info["labels"],
        "sweeps": info["sweeps"],
        "levels": info["levels"],
        "outer_iterations": info["outer"],
        "time_to_stable_s": dt,
        "edges_per_s_per_sweep": m / sweep_s if sweep_s else None,
        "first_full_sweep_ms": sweep_s * 1000.0 if sweep_s else None,
        "roofline": roofline(kstats, a.scale),
        "kernel_groups": groups,
        "instrumented_step_s": instrumented_s,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
