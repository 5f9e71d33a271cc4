"""The top level's local mover on a copy of the graph renumbered by degree
(renumber knob 1; the default 2 uses it for runs of several outer
iterations) gives the same clustering, bit for bit, as on the caller's
graph: objective, partition, sweeps, moves and outer iterations, for
pLouvain, pLeiden and pLeiden+, the objectives that create new clusters
(CPM, LCC), weighted graphs, graphs whose rows are not sorted, and the
seam entry points pggc_louvain / pggc_leiden.  (The oracle parity tests run
with the default as well.)"""
import ctypes

import numpy as np
import pytest

from tests.graphcases import hub_graph
from pennsylvaniagpugraphclustering_amd import graphs

pytestmark = pytest.mark.gpu


def _shuffled_rows(g, seed):
    """The same graph with every row's entries in a random order."""
    col, w = g.col.copy(), None if g.w is None else g.w.copy()
    rng = np.random.default_rng(seed)
    for i in range(g.n):
        s, e = g.rp[i], g.rp[i + 1]
        p = rng.permutation(e - s)
        col[s:e] = col[s:e][p]
        if w is not None:
            w[s:e] = w[s:e][p]
    return graphs.CSR(g.n, g.rp, col, w)


CASES = {
    "rmat14": (lambda: graphs.rmat(14, 16, seed=2), "Mod"),
    "rmat13w": (lambda: graphs.rmat(13, 16, seed=3, weights=(1, 100)), "WMod"),
    "sbm8k_cpm": (lambda: graphs.sbm(8192, 64, 12, 4, seed=1, permute_seed=1), "CPM"),
    "er4k_lcc": (lambda: graphs.random_graph(4000, 12, seed=4, weights=(1, 10)), "LCC"),
    "hub30k": (lambda: hub_graph(), "Mod"),
    "rmat12_unsorted": (lambda: _shuffled_rows(graphs.rmat(12, 16, seed=5, weights=(1, 50)), 7), "WMod"),
}


def _run(gpu_ctx, dg, obj, algo, ex_iters, renumber):
    from pennsylvaniagpugraphclustering_amd import clustering as C
    gpu_ctx.set_knob("renumber", renumber)
    try:
        q, part, info = C.cluster(dg, algo, obj, 0.05 if obj == "CPM" else 1.0, ex_iters, seed=3)
    finally:
        gpu_ctx.set_knob("renumber", 2)
    return q, part.cpu().numpy(), info


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("algo,ex_iters", [("louvain", -1), ("leiden", 0), ("leiden+", 1)])
def test_renumbered_equals_caller_order(gpu_ctx, name, algo, ex_iters):
    from pennsylvaniagpugraphclustering_amd import clustering as C
    make, obj = CASES[name]
    g = make()
    print(f"renumber {name} {algo} ex_iters={ex_iters}", flush=True)
    dg = C.DeviceGraph.from_csr(g, obj)
    a = _run(gpu_ctx, dg, obj, algo, ex_iters, 1)
    b = _run(gpu_ctx, dg, obj, algo, ex_iters, 0)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    for k in ["labels", "uncut", "sweeps", "moves", "outer", "levels"]:
        assert a[2][k] == b[2][k], k


@pytest.mark.parametrize("entry", ["louvain", "louvain_constrained", "leiden", "leiden+_improve"])
def test_seam_entry_points(gpu_ctx, entry):
    """pggc_louvain / pggc_leiden (one louvain_part / leiden_part call)."""
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    g = graphs.rmat(13, 16, seed=8)
    dg = C.DeviceGraph.from_csr(g, "Mod")
    print(f"renumber seam {entry}", flush=True)
    outs = []
    for rn in (1, 0):
        gpu_ctx.set_knob("renumber", rn)
        # the refinement's random draws restart from the same seed
        gpu_ctx.lib.pggc_ctx_set_seed(gpu_ctx.h, ctypes.c_uint64(7), ctypes.c_uint64(0))
        try:
            st, vol = C.state_init(dg, "Mod", 1.0, ctx=gpu_ctx)
            if entry == "louvain":
                part = C.louvain_part(dg, st, ctx=gpu_ctx)
            elif entry == "louvain_constrained":
                cons = torch.arange(dg.n, dtype=torch.int32, device="cuda") // 64
                part = C.louvain_part(dg, st, constrained=True, constraint=cons, ctx=gpu_ctx)
            elif entry == "leiden":
                part = C.leiden_part(dg, st, plus=False, ctx=gpu_ctx)
            else:
                first = C.louvain_part(dg, st, ctx=gpu_ctx).clone()
                st, vol = C.state_init(dg, "Mod", 1.0, ctx=gpu_ctx)
                part = C.leiden_part(dg, st, plus=True, improve=True, input=first, ctx=gpu_ctx)
            outs.append((st.objective(), st.label_count, part.cpu().numpy()))
        finally:
            gpu_ctx.set_knob("renumber", 2)
    assert outs[0][:2] == outs[1][:2]
    assert np.array_equal(outs[0][2], outs[1][2])
