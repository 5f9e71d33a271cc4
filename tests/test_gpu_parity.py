"""GPU parity: every hot-path entry point of the HIP library, called through
the C ABI, against the CPU oracle on the same seeded inputs.  Integer and
index results must be bit-identical; objectives are compared as the exact
same double."""
import numpy as np
import pytest

from oracle import oracle as O
from graphcases import cases, hub_graph
from pennsylvaniagpugraphclustering_amd import graphs

pytestmark = pytest.mark.gpu

CASES = cases()


def C():
    from pennsylvaniagpugraphclustering_amd import clustering
    return clustering


def dgraph(g, objective):
    return C().DeviceGraph.from_csr(g, objective)


def weights_for(g, objective):
    return g.w if objective not in ("Mod", "CPM") else None


def npy(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", list(CASES))
def test_orderings(gpu_ctx, name):
    g, obj = CASES[name]
    print(f"orderings {name}", flush=True)
    o1, o2, offs = C().generate_orderings(dgraph(g, obj))
    r1, r2, roffs = O.orderings(g.rp)
    assert np.array_equal(npy(o1), r1)
    assert np.array_equal(npy(o2), r2)
    assert list(offs) == list(roffs)


@pytest.mark.parametrize("name", list(CASES))
def test_evaluator(gpu_ctx, name):
    g, obj = CASES[name]
    dg = dgraph(g, obj)
    w = weights_for(g, obj)
    vw = g.vertex_weights(obj)
    rng = np.random.default_rng(1)
    for L in [1, 5, max(1, g.n // 10), g.n]:
        print(f"evaluate {name} L={L}", flush=True)
        lab = rng.integers(0, L, g.n).astype(np.int32)
        vol, uncut, sq = C().evaluate(dg, lab, L)
        rv, ru, rs = O.evaluate(g.rp, g.col, w, vw, lab, L)
        assert np.array_equal(npy(vol), rv)
        assert uncut == ru and sq == rs


@pytest.mark.parametrize("name", list(CASES))
def test_own_weight(gpu_ctx, name):
    g, obj = CASES[name]
    lab = np.random.default_rng(2).integers(0, max(2, g.n // 50), g.n).astype(np.int32)
    print(f"own_weight {name}", flush=True)
    out = C().own_weight(dgraph(g, obj), lab)
    assert np.array_equal(npy(out), O.own_weight(g.rp, g.col, weights_for(g, obj), lab))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("frac", [0.0, 0.001, 0.05, 0.5])
def test_contraction(gpu_ctx, name, frac):
    """Coarse graph of a random labelling: rows, columns and summed weights
    bit-identical (rows sorted); frac=0 puts everything in 3 clusters so the
    coarse rows exceed the LDS tables."""
    g, obj = CASES[name]
    nc = 3 if frac == 0.0 else max(1, int(g.n * frac))
    rng = np.random.default_rng(int(frac * 1000) + 7)
    vc = rng.integers(0, nc, g.n).astype(np.int32)
    vc[: min(nc, g.n)] = np.arange(min(nc, g.n))
    print(f"contract {name} nc={nc}", flush=True)
    rp, col, w = C().build_coarse_graph(dgraph(g, obj), vc, nc)
    rrp, rcol, rw = O.contract(g.rp, g.col, weights_for(g, obj), vc, nc)
    assert np.array_equal(npy(rp), rrp)
    assert np.array_equal(npy(col), rcol)
    assert np.array_equal(npy(w), rw)


@pytest.mark.parametrize("extra", [0, 3000])
def test_contraction_with_runs_of_empty_rows(gpu_ctx, extra):
    """A perfect matching whose pair (2k, 2k+1) maps to clusters 3k and 3k+1
    leaves every third coarse row empty: a chunk of 4096 coarse entries spans
    about 6144 rows, more than the row starts the copy stages in LDS, so the
    copy searches the row offsets in global memory.  `extra` random edges
    make some rows longer.  Bit-identical to the oracle."""
    n = 200000
    rng = np.random.default_rng(5)
    src = np.concatenate([np.arange(0, n, 2), rng.integers(0, n, extra)])
    dst = np.concatenate([np.arange(1, n, 2), rng.integers(0, n, extra)])
    g = graphs.from_edges(n, src, dst)
    u = np.arange(n)
    vc = (3 * (u // 2) + (u % 2)).astype(np.int32)
    nc = 3 * (n // 2)
    print(f"contract matching n={n} extra={extra} nc={nc}", flush=True)
    rp, col, w = C().build_coarse_graph(dgraph(g, "Mod"), vc, nc)
    rrp, rcol, rw = O.contract(g.rp, g.col, None, vc, nc)
    assert int(np.diff(rrp).astype(bool).sum()) <= 2 * nc // 3 + 2 * extra
    assert np.array_equal(npy(rp), rrp)
    assert np.array_equal(npy(col), rcol)
    assert np.array_equal(npy(w), rw)


def test_contraction_of_coarse_graph(gpu_ctx):
    """Contraction input with int64 weights (a coarse level)."""
    g = graphs.random_graph(5000, 12, seed=3, weights=(1, 9))
    vc = np.random.default_rng(0).integers(0, 800, g.n).astype(np.int32)
    vc[:800] = np.arange(800)
    rrp, rcol, rw = O.contract(g.rp, g.col, g.w, vc, 800)
    vc2 = np.random.default_rng(1).integers(0, 40, 800).astype(np.int32)
    vc2[:40] = np.arange(40)
    dg = C().DeviceGraph(rrp, rcol, rw, np.ones(800, np.int64))
    rp, col, w = C().build_coarse_graph(dg, vc2, 40)
    erp, ecol, ew = O.contract(rrp, rcol, rw, vc2, 40)
    assert np.array_equal(npy(rp), erp) and np.array_equal(npy(col), ecol) and np.array_equal(npy(w), ew)


def sweep_state(g, obj, seed, L_frac):
    rng = np.random.default_rng(seed)
    vw = g.vertex_weights(obj)
    L = max(2, int(g.n * L_frac))
    part = rng.integers(0, L, g.n).astype(np.int32)
    vol, _, _ = O.evaluate(g.rp, g.col, weights_for(g, obj), vw, part, L)
    dest = np.where(rng.random(g.n) < 0.2, -3, -1).astype(np.int32)
    gain = rng.standard_normal(g.n).astype(np.float32)
    return vw, part, vol, dest, gain


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("ratio", [0.0, 0.75])
@pytest.mark.parametrize("L_frac", [0.01, 0.3])
def test_one_sweep_identical_moves(gpu_ctx, name, ratio, L_frac):
    """One local-move sweep from a fixed state: destinations, FP32 gains and
    the accepted move list (order1 order) are bit-identical."""
    g, obj = CASES[name]
    vw, part, vol, dest, gain = sweep_state(g, obj, 5, L_frac)
    W = g.nnz if obj == "Mod" else int(vw.sum())
    lam = 1.0 / W
    print(f"sweep {name} ratio={ratio} L_frac={L_frac}", flush=True)
    mv, d2, g2, nc = C().sweep(dgraph(g, obj), part, vol, lam, ratio, False, dest, gain)
    rmv, rd, rg, rnc = O.sweep(g.rp, g.col, weights_for(g, obj), vw, part, vol, lam, ratio, False, dest, gain)
    assert nc == rnc
    assert np.array_equal(npy(d2), rd)
    assert np.array_equal(npy(g2).view(np.int32), rg.view(np.int32))
    assert np.array_equal(npy(mv), rmv)


@pytest.mark.parametrize("name", list(CASES))
def test_first_sweep_raw(gpu_ctx, name):
    g, obj = CASES[name]
    vw = g.vertex_weights(obj)
    part = np.arange(g.n, dtype=np.int32)
    dest = np.full(g.n, -1, np.int32)
    gain = np.zeros(g.n, np.float32)
    W = g.nnz if obj == "Mod" else int(vw.sum())
    print(f"raw sweep {name}", flush=True)
    mv, d2, g2, nc = C().sweep(dgraph(g, obj), part, vw, 1.0 / W, 0.75, True, dest, gain)
    rmv, rd, rg, rnc = O.sweep(g.rp, g.col, weights_for(g, obj), vw, part, vw, 1.0 / W, 0.75, True, dest, gain)
    assert nc == rnc
    assert np.array_equal(npy(d2), rd)
    assert np.array_equal(npy(g2).view(np.int32), rg.view(np.int32))
    assert np.array_equal(npy(mv), rmv)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("mode", ["initial", "improve", "strict"])
def test_local_move(gpu_ctx, name, mode):
    g, obj = CASES[name]
    vw = g.vertex_weights(obj)
    w = weights_for(g, obj)
    W = g.nnz if obj == "Mod" else int(vw.sum())
    lam = 1.0 / W
    if mode == "initial":
        part = np.arange(g.n, dtype=np.int32)
        L = g.n
    else:
        L = max(2, g.n // 20)
        part = np.random.default_rng(9).integers(0, L, g.n).astype(np.int32)
        part[:L] = np.arange(L)
    vol, uncut, sq = O.evaluate(g.rp, g.col, w, vw, part, L)
    obj0 = float(uncut) - lam * float(sq)
    print(f"local_move {name} {mode}", flush=True)
    rp_, rv, rL, ru, ro = O.local_move(g.rp, g.col, w, vw, part, vol, L, uncut, lam, obj0, mode == "initial",
                                        strict=(mode == "strict"))
    st, volt = C().make_state(vol, L, uncut, lam, obj0, cap=g.n)
    out = C().local_move(dgraph(g, obj), part, st, mode == "initial", strict=(mode == "strict"))
    assert np.array_equal(npy(out), rp_)
    assert st.label_count == rL and st.uncut == ru and st.obj == ro
    assert np.array_equal(npy(volt)[:rL], rv)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("seed", [1, 2])
def test_leiden_refine(gpu_ctx, name, seed):
    g, obj = CASES[name]
    vw = g.vertex_weights(obj)
    w = weights_for(g, obj)
    _, louv, _ = O.cluster(g.rp, g.col, w, vw, obj, 1.0, "louvain", 0)
    L = int(louv.max()) + 1
    vol, _, _ = O.evaluate(g.rp, g.col, w, vw, louv, L)
    W = g.nnz if obj == "Mod" else int(vw.sum())
    uni = w is None
    print(f"leiden_refine {name} seed={seed}", flush=True)
    vc, nc = C().coarsen_leidenR(dgraph(g, obj), louv, vol, 1.0 / W, uni, seed)
    rvc, rnc = O.leiden_refine(g.rp, g.col, w, vw, louv, vol, 1.0 / W, uni, seed)
    assert nc == rnc
    assert np.array_equal(npy(vc), rvc)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("algo", ["louvain", "leiden", "leiden+"])
@pytest.mark.parametrize("ex_iters", [0, -1])
def test_full_clustering(gpu_ctx, name, algo, ex_iters):
    """Whole runs (multilevel + ex_iters loop): identical clustering and the
    exact same objective value as the oracle."""
    g, obj = CASES[name]
    vw = g.vertex_weights(obj)
    w = weights_for(g, obj)
    print(f"cluster {name} {algo} ex_iters={ex_iters}", flush=True)
    q, part, info = C().cluster(dgraph(g, obj), algo, obj, 1.0, ex_iters, seed=7)
    rq, rpart, rinfo = O.cluster(g.rp, g.col, w, vw, obj, 1.0, algo, ex_iters, seed=7)
    assert np.array_equal(npy(part), rpart)
    assert q == rq
    for k in ["labels", "uncut", "sweeps", "moves", "outer"]:
        assert info[k] == rinfo[k], k


@pytest.mark.parametrize("algo", ["louvain", "leiden+"])
def test_full_clustering_rmat16(gpu_ctx, algo):
    """R-MAT scale 16 (65,536 vertices, 1.8M nonzeros, hubs up to degree
    ~10^4): one full run is bit-identical to the oracle."""
    g = graphs.rmat(16, 16, seed=2)
    vw = g.vertex_weights("Mod")
    print(f"cluster rmat16 {algo}", flush=True)
    q, part, info = C().cluster(dgraph(g, "Mod"), algo, "Mod", 1.0, 0, seed=3)
    rq, rpart, rinfo = O.cluster(g.rp, g.col, None, vw, "Mod", 1.0, algo, 0, seed=3)
    assert np.array_equal(npy(part), rpart)
    assert q == rq


@pytest.mark.parametrize("hub_build", [0, 1, 2])
@pytest.mark.parametrize("algo,ex_iters", [("louvain", -1), ("leiden+", 0)])
def test_wide_hub_tables(gpu_ctx, hub_build, algo, ex_iters):
    """Edge weights of 2^20..2^21 on the hub graph: the hub's row weight
    (about 3e10) does not fit the packed 32-bit table slots, so the hub
    tables hold separate keys and 64-bit weights; a degree-20,000 hub has a
    table of two LDS slices.  Every build method (LDS slices, edge items with
    HBM atomics, a mix of both) gives the oracle's clustering."""
    g = hub_graph(seed=11, weights=(1 << 20, 1 << 21))
    vw = g.vertex_weights("WMod")
    print(f"cluster hub30k wide {algo} hub_build={hub_build}", flush=True)
    try:
        gpu_ctx.set_knob("hub_build", hub_build)
        q, part, info = C().cluster(dgraph(g, "WMod"), algo, "WMod", 1.0, ex_iters, seed=4)
    finally:
        gpu_ctx.set_knob("hub_build", 0)
    rq, rpart, rinfo = O.cluster(g.rp, g.col, g.w, vw, "WMod", 1.0, algo, ex_iters, seed=4)
    assert np.array_equal(npy(part), rpart)
    assert q == rq
    assert info["sweeps"] == rinfo["sweeps"]


def test_determinism_rmat20(gpu_ctx):
    """Size-independent property at scale 20 (1M vertices, 33M nonzeros):
    two runs give the same clustering and objective, and the objective equals
    a fresh evaluation of the returned clustering."""
    import torch
    rp, col, _ = graphs.rmat_device(20, 16, seed=2)
    dg = C().DeviceGraph(rp, col, None, None, objective="Mod")
    print("determinism rmat20", flush=True)
    q1, p1, i1 = C().cluster(dg, "louvain", "Mod", 1.0, 0, seed=1)
    q2, p2, i2 = C().cluster(dg, "louvain", "Mod", 1.0, 0, seed=1)
    assert torch.equal(p1, p2) and q1 == q2
    q3 = C().modularity(dg, p1, "Mod")
    assert abs(q3 - q1) <= 1e-12
    assert i1["labels"] == int(p1.max().item()) + 1


@pytest.mark.parametrize("name", ["rmat12w", "er3k_w"])
@pytest.mark.parametrize("objective", ["CPM", "NLCC", "LCC"])
@pytest.mark.parametrize("algo", ["louvain", "leiden+"])
def test_lambdacc_objectives(gpu_ctx, name, objective, algo):
    """CPM, NLCC and LCC (interface/cluster_data.hpp:135-182): whole runs
    bit-identical to the oracle, lambda multiplier 0.5."""
    g, _ = CASES[name]
    vw = g.vertex_weights(objective)
    w = None if objective == "CPM" else g.w
    print(f"cluster {name} {objective} {algo}", flush=True)
    q, part, info = C().cluster(dgraph(g, objective), algo, objective, 0.5, 0, seed=5)
    rq, rpart, rinfo = O.cluster(g.rp, g.col, w, vw, objective, 0.5, algo, 0, seed=5)
    assert np.array_equal(npy(part), rpart)
    assert q == rq


def test_sbm_131072_planted_partition(gpu_ctx):
    """Planted-partition SBM: 131,072 vertices in 1,024 communities of 128,
    expected intra degree 12 and inter degree 4, ids permuted with seed 1;
    plain pLouvain (modularity, ex_iters 0).  Clustering, objective and run
    counters are bit-identical to the oracle."""
    g = graphs.sbm(131072, 1024, 12, 4, seed=1, permute_seed=1)
    vw = g.vertex_weights("Mod")
    print(f"cluster sbm131072 n={g.n} nnz={g.nnz}", flush=True)
    q, part, info = C().cluster(dgraph(g, "Mod"), "louvain", "Mod", 1.0, 0, seed=0)
    rq, rpart, rinfo = O.cluster(g.rp, g.col, None, vw, "Mod", 1.0, "louvain", 0, seed=0)
    assert np.array_equal(npy(part), rpart)
    assert q == rq
    for k in ["labels", "uncut", "sweeps", "moves", "outer"]:
        assert info[k] == rinfo[k], k


@pytest.mark.parametrize("batch", [1, 4096, 1 << 16])
@pytest.mark.parametrize("name,nc", [("rmat12w", 3), ("hub30k", 40), ("rmat12", 17)])
def test_contraction_in_batches(gpu_ctx, name, nc, batch):
    """The big coarse rows (bound > 2048) merged and sorted in batches of at
    most `batch` table slots (a quarter as many sort entries; 1 puts every
    row in a batch of its own): bit-identical to the oracle, as with the
    default single batch."""
    g, obj = CASES[name]
    rng = np.random.default_rng(nc + batch)
    vc = rng.integers(0, nc, g.n).astype(np.int32)
    vc[:nc] = np.arange(nc)
    print(f"contract {name} nc={nc} batch={batch}", flush=True)
    gpu_ctx.set_knob("contract_batch", batch)
    try:
        rp, col, w = C().build_coarse_graph(dgraph(g, obj), vc, nc)
    finally:
        gpu_ctx.set_knob("contract_batch", 1 << 30)
    rrp, rcol, rw = O.contract(g.rp, g.col, weights_for(g, obj), vc, nc)
    assert np.array_equal(npy(rp), rrp)
    assert np.array_equal(npy(col), rcol)
    assert np.array_equal(npy(w), rw)
