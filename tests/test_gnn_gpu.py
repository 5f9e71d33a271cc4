"""GPU parity of the active-set predictor, called through the C ABI, against
the CPU oracle (oracle/gnn.py), and the invariants of its use in the local
mover.

Tolerances: the bf16 input rows are compared bit for bit.  Logits, loss and
gradient depend on the MFMA's fp32 accumulation order and on the bf16
rounding of the hidden layer that follows it, so a hidden unit can round one
bf16 step differently: logits within 2e-2 * (1 + |logit|) everywhere and
within 1e-4 * (1 + |logit|) for at least 95% of vertices; loss within 1e-4
relative; gradient within 2e-2 in relative L2 norm."""
import numpy as np
import pytest

from oracle import gnn as N
from graphcases import cases, hub_graph
from pennsylvaniagpugraphclustering_amd import graphs

pytestmark = pytest.mark.gpu

CASES = cases()
FLT_MAX = float(np.finfo(np.float32).max)


def C():
    from pennsylvaniagpugraphclustering_amd import clustering
    return clustering


def npy(t):
    return t.detach().cpu().numpy()


def state(g, seed, first=False):
    rng = np.random.default_rng(seed)
    n = g.n
    L = max(2, n // 20)
    part = rng.integers(0, L, n).astype(np.int32)
    dest = np.where(rng.random(n) < 0.3, -3, rng.integers(-1, L, n)).astype(np.int32)
    gain = (rng.standard_normal(n) * 10.0 ** rng.integers(-2, 3, n)).astype(np.float32)
    gain[rng.random(n) < 0.1] = -FLT_MAX
    stamp = rng.integers(0, 4, n).astype(np.int32)
    vol = np.bincount(part, weights=g.vertex_weights("Mod"), minlength=L).astype(np.int64)
    return part, dest, gain, stamp, 3, first, vol


def params_for(seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(N.GNN_P) * 0.4).astype(np.float32)
    p[N.GP_B2] = -1.0
    return p


GNN_CASES = ["karate", "sbm4k", "rmat12", "rmat12w", "hub30k"]


@pytest.mark.parametrize("name", GNN_CASES)
@pytest.mark.parametrize("first", [False, True])
def test_rows_logits_grad_match_oracle(gpu_ctx, name, first):
    g, obj = CASES[name]
    w = g.w if obj == "WMod" else None
    vw = g.vertex_weights(obj)
    part, dest, gain, stamp, cur, first, vol = state(g, 11, first)
    p = params_for(3)
    C().gnn_set_params(p)
    y = (np.random.default_rng(5).random(g.n) < 0.2).astype(np.int32)
    dg = C().DeviceGraph.from_csr(g, obj)
    print(f"gnn eval {name} first={first}", flush=True)
    rows, lg, loss, grad = C().gnn_eval(dg, part, dest, gain, stamp, cur, first, vol, labels=y)
    erows = N.rows(g.rp, g.col, w, vw, part, dest, gain, stamp, cur, first, vol)
    assert np.array_equal(rows, erows)
    act = dest != -3
    eloss, egrad, elog = N.loss_and_grad(g.rp, g.col, erows, p, act, y, 4.0)
    d = np.abs(lg[act] - elog[act])
    scale = 1.0 + np.abs(elog[act])
    assert np.all(d <= 2e-2 * scale)
    assert np.mean(d <= 1e-4 * scale) >= 0.95
    assert np.all(lg[~act] == 0)
    assert abs(loss - eloss) <= 1e-4 * abs(eloss)
    assert np.linalg.norm(grad - egrad) <= 2e-2 * np.linalg.norm(egrad)


def test_mfma_layer_exact_on_representable_case(gpu_ctx):
    """Small integer and dyadic weights keep the products of bf16 inputs and
    weights exact, so the hidden layer rounds like the oracle's: logits agree
    to 1e-5 on every active vertex, which fixes the MFMA operand layout (a
    transposed or shifted operand moves them by far more)."""
    g = graphs.random_graph(500, 8, seed=7)
    part, dest, gain, stamp, cur, first, vol = state(g, 2)
    rng = np.random.default_rng(9)
    p = np.zeros(N.GNN_P, np.float32)
    p[N.GP_W1:N.GP_B1] = rng.integers(-3, 4, 512).astype(np.float32)
    p[N.GP_B1:N.GP_W2S] = rng.integers(-2, 3, 32).astype(np.float32) / 4
    p[N.GP_W2S:N.GP_W2N] = rng.integers(-3, 4, 32).astype(np.float32) / 8
    p[N.GP_W2N:N.GP_B2] = rng.integers(-3, 4, 32).astype(np.float32) / 8
    C().gnn_set_params(p)
    dg = C().DeviceGraph.from_csr(g, "Mod")
    rows, lg, _, _ = C().gnn_eval(dg, part, dest, gain, stamp, cur, False, vol)
    erows = N.rows(g.rp, g.col, None, g.vertex_weights("Mod"), part, dest, gain, stamp, cur, False, vol)
    assert np.array_equal(rows, erows)
    elog = N.forward(g.rp, g.col, erows, p)[5]
    act = dest != -3
    np.testing.assert_allclose(lg[act], elog[act], rtol=1e-5, atol=1e-5)


def test_train_mode_leaves_clustering_unchanged(gpu_ctx):
    """Training observes the plain sweeps: the clustering is bit-identical
    to a run without the predictor, and the mean loss falls over epochs."""
    g = graphs.rmat(14, 16, seed=21)
    dg = C().DeviceGraph.from_csr(g, "Mod")
    C().gnn_configure("off")
    q0, p0, i0 = C().cluster(dg, "louvain", "Mod", 1.0, -1, seed=1)
    p0 = npy(p0)
    C().gnn_set_params(params_for(4))
    C().gnn_configure("train", min_nnz=0, lr=3e-3)
    losses = []
    for ep in range(6):
        C().gnn_reset_stats()
        q, p, info = C().cluster(dg, "louvain", "Mod", 1.0, -1, seed=1)
        st = C().gnn_stats()
        print(f"epoch {ep} steps {st['train_steps']} loss {st['loss_sum'] / max(st['train_steps'], 1):.4f}", flush=True)
        assert q == q0 and np.array_equal(npy(p), p0) and info == i0
        assert st["train_steps"] > 0 and st["pruned"] == 0
        losses.append(st["loss_sum"] / st["train_steps"])
    C().gnn_configure("off")
    assert losses[-1] < losses[0]


@pytest.mark.parametrize("algo", ["louvain", "leiden+"])
def test_prune_extremes_equal_plain(gpu_ctx, algo):
    """tau = -inf prunes nothing; tau = +inf prunes every active vertex, so
    each sweep first finds nothing and then scores the restored vertices --
    the plain sweep.  Both give the plain clustering bit for bit."""
    g = hub_graph()
    dg = C().DeviceGraph.from_csr(g, "Mod")
    C().gnn_configure("off")
    q0, p0, i0 = C().cluster(dg, algo, "Mod", 1.0, 0, seed=2)
    p0 = npy(p0)
    C().gnn_set_params(params_for(6))
    for tau in (-1e30, 1e30):
        C().gnn_configure("prune", tau=tau, min_nnz=0)
        C().gnn_reset_stats()
        print(f"prune {algo} tau={tau}", flush=True)
        q, p, info = C().cluster(dg, algo, "Mod", 1.0, 0, seed=2)
        st = C().gnn_stats()
        assert q == q0 and np.array_equal(npy(p), p0)
        assert st["sweeps"] > 0
        if tau > 0:
            assert st["pruned"] == st["active"] and st["restores"] > 0
        else:
            assert st["pruned"] == 0 and st["restores"] == 0
    C().gnn_configure("off")


def test_prune_mode_modularity_close(gpu_ctx):
    """A pruning run with a middle threshold ends every local-move call under
    the plain condition; its modularity stays within 1e-3 of plain here."""
    g = graphs.sbm(65536, 512, 12, 4, seed=3, permute_seed=3)
    dg = C().DeviceGraph.from_csr(g, "Mod")
    C().gnn_configure("off")
    q0, _, _ = C().cluster(dg, "louvain", "Mod", 1.0, -1, seed=1)
    C().gnn_set_params(params_for(8))
    C().gnn_configure("prune", tau=-1.0, min_nnz=0)
    C().gnn_reset_stats()
    q, p, info = C().cluster(dg, "louvain", "Mod", 1.0, -1, seed=1)
    st = C().gnn_stats()
    C().gnn_configure("off")
    print(f"plain {q0} pruned {q} stats {st}", flush=True)
    assert q >= q0 - 1e-3
    assert abs(C().modularity(dg, p, "Mod") - q) <= 1e-12
