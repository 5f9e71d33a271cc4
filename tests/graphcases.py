"""Small deterministic graphs shared by the tests."""
import numpy as np

from pennsylvaniagpugraphclustering_amd import graphs


def hub_graph(n=30000, hub_deg=20000, seed=5, weights=None):
    """SBM background plus one hub adjacent to hub_deg random vertices, and a
    second hub of degree 3000 (exercises every degree class)."""
    base = graphs.sbm(n, n // 128, 8, 3, seed=seed)
    rng = np.random.default_rng(seed + 1)
    rows = np.repeat(np.arange(n), np.diff(base.rp))
    src = [rows, np.zeros(hub_deg, np.int64), np.full(3000, 1, np.int64)]
    dst = [base.col.astype(np.int64), rng.choice(np.arange(2, n), hub_deg, replace=False),
           rng.choice(np.arange(2, n), 3000, replace=False)]
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    w = None
    if weights is not None:
        w = rng.integers(weights[0], weights[1] + 1, len(src))
    return graphs.from_edges(n, src, dst, w)


def cases():
    return {
        "karate": (graphs.karate(), "Mod"),
        "sbm4k": (graphs.sbm(4096, 32, 12, 4, seed=1, permute_seed=1), "Mod"),
        "rmat12": (graphs.rmat(12, 16, seed=2), "Mod"),
        "rmat12w": (graphs.rmat(12, 16, seed=3, weights=(1, 100)), "WMod"),
        "er3k_w": (graphs.random_graph(3000, 10, seed=4, weights=(1, 10)), "WMod"),
        "hub30k": (hub_graph(), "Mod"),
    }
