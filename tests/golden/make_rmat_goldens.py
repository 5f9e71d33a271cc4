"""Writes the oracle's results on the counter-based R-MAT graphs as small
JSON fixtures (hashes and counts), for GPU tests at full size.

Run once in the build container (it takes tens of minutes at scale 24 on 8
cores):

    python tests/golden/make_rmat_goldens.py --scale 24 --seed 2 --louvain --leiden-plus 2

The graph is the R-MAT graph of the oracle's generator (Philox4x32-10 draws,
a/b/c/d = .57/.19/.19/.05, edge factor 16), symmetrised, deduplicated,
without self loops, unit weights.  Recorded:

- the graph: n, nnz and SHA-256 of row offsets (int64) and columns (int32);
- the contraction of the fixed 65,536-cluster labelling
  labels[i] = Philox(counter (i, 0, 0, 2), key seed)[0] mod 65536: coarse nnz
  and SHA-256 of its row offsets, columns and int64 weights, and the
  evaluator's uncut weight and exact sum of squared volumes;
- pLouvain with ex_iters -1 (and pLeiden+ with the given ex_iters): the
  objective (float.hex), the run counters and SHA-256 of the int32 labels.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--clusters", type=int, default=65536)
    ap.add_argument("--louvain", action="store_true")
    ap.add_argument("--leiden-plus", type=int, default=None, help="ex_iters of a pLeiden+ run")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    out = a.out or os.path.join(HERE, f"rmat{a.scale}_seed{a.seed}.json")
    res = json.load(open(out)) if os.path.exists(out) else {}
    t = time.time()
    rp, col, _ = O.rmat(a.scale, 16, 0.57, 0.19, 0.19, a.seed)
    n = len(rp) - 1
    res["graph"] = {"scale": a.scale, "edge_factor": 16, "abc": [0.57, 0.19, 0.19], "seed": a.seed, "n": n,
                    "nnz": int(rp[-1]), "sha_rp": sha(rp), "sha_col": sha(col)}
    print(f"graph n={n} nnz={rp[-1]} {time.time() - t:.1f}s", flush=True)
    vw = np.diff(rp).astype(np.int64)
    lab = O.random_labels(n, a.clusters, a.seed)
    crp, ccol, cw = O.contract(rp, col, None, lab, a.clusters)
    vol, uncut, sq = O.evaluate(rp, col, None, vw, lab, a.clusters)
    res["contract"] = {"clusters": a.clusters, "label_seed": a.seed, "nnz": int(crp[-1]), "sha_rp": sha(crp),
                       "sha_col": sha(ccol), "sha_w": sha(cw), "uncut": uncut, "sumsq": str(sq),
                       "sha_vol": sha(vol)}
    print(f"contraction nnz={crp[-1]} {time.time() - t:.1f}s", flush=True)
    del crp, ccol, cw
    runs = res.setdefault("runs", {})
    todo = []
    if a.louvain:
        todo.append(("louvain", -1))
    if a.leiden_plus is not None:
        todo.append(("leiden+", a.leiden_plus))
    for algo, ex in todo:
        t1 = time.time()
        q, part, info = O.cluster(rp, col, None, vw, "Mod", 1.0, algo, ex, seed=0)
        runs[f"{algo}_ex{ex}"] = {"algo": algo, "ex_iters": ex, "seed": 0, "objective_hex": float(q).hex(),
                                  "objective": q, "info": info, "sha_part": sha(part),
                                  "oracle_seconds": round(time.time() - t1, 1), "oracle_threads": O.get_threads()}
        print(f"{algo} ex_iters={ex}: Q={q} {info} {time.time() - t1:.1f}s", flush=True)
        json.dump(res, open(out, "w"), indent=1)
    json.dump(res, open(out, "w"), indent=1)


if __name__ == "__main__":
    main()
