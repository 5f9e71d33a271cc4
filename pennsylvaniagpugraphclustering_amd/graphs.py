"""Synthetic graph generators and CSR helpers.

Graphs are undirected, stored as symmetric CSR with no self loops and sorted
rows: ``rp`` int64[n+1], ``col`` int32[nnz], ``w`` int32[nnz] or None (unit
weights).  This is the layout the reference loads (app/io/io_mtx.cpp:95-201
symmetrises, deduplicates and sorts; app/io/io_metis.cpp reads it as given).

Small graphs are generated with numpy (deterministic across machines); the
large R-MAT workloads are built on the GPU by the library (rmat_device), from
counter-based draws that the CPU oracle reproduces.
"""
import numpy as np


class CSR:
    __slots__ = ("n", "rp", "col", "w")

    def __init__(self, n, rp, col, w=None):
        self.n = int(n)
        self.rp = np.ascontiguousarray(rp, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.w = None if w is None else np.ascontiguousarray(w, dtype=np.int32)

    @property
    def nnz(self):
        return int(self.rp[-1])

    def degrees(self):
        return np.diff(self.rp)

    def vertex_weights(self, objective="Mod"):
        """app/io/vertex_weighting.cpp:27-44: degree for Mod, weighted degree
        for WMod, unit otherwise."""
        if objective == "Mod":
            return self.degrees().astype(np.int64)
        if objective == "WMod":
            return _wdeg(self)
        return np.ones(self.n, np.int64)


def _wdeg(g):
    out = np.zeros(g.n, np.int64)
    if g.nnz == 0:
        return out
    w = np.ones(g.nnz, np.int64) if g.w is None else g.w.astype(np.int64)
    rows = np.repeat(np.arange(g.n), np.diff(g.rp))
    np.add.at(out, rows, w)
    return out


def from_edges(n, src, dst, w=None, combine="first"):
    """Symmetrise, drop self loops, deduplicate and sort an edge list into
    CSR.  Duplicate (u,v) pairs keep the weight of the first occurrence in
    (u,v)-sorted order, so the result is symmetric."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    if w is not None:
        w = np.asarray(w, dtype=np.int64)[keep]
    # canonical undirected key (min, max) so both directions share a weight
    lo = np.minimum(src, dst)
    hi = np.maximum(src, dst)
    key = lo * n + hi
    order = np.argsort(key, kind="stable")
    key = key[order]
    first = np.ones(len(key), bool)
    first[1:] = key[1:] != key[:-1]
    key = key[first]
    if w is not None:
        w = w[order][first]
    lo = key // n
    hi = key % n
    s = np.concatenate([lo, hi])
    d = np.concatenate([hi, lo])
    ww = None if w is None else np.concatenate([w, w])
    k2 = s * n + d
    o2 = np.argsort(k2, kind="stable")
    s, d = s[o2], d[o2]
    if ww is not None:
        ww = ww[o2]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, s + 1, 1)
    rp = np.cumsum(rp)
    return CSR(n, rp, d.astype(np.int32), None if ww is None else ww.astype(np.int32))


def permute(g, perm):
    """Relabel vertex v as perm[v] (sorted rows are kept)."""
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(g.rp))
    return from_edges(n, perm[rows], perm[g.col], g.w)


def sbm(n, blocks, d_in, d_out, seed, permute_seed=None, weights=None):
    """Planted-partition SBM: `blocks` near-equal communities, expected intra
    degree d_in (p_in = d_in/(b-1)) and inter degree d_out
    (p_out = d_out/(n-b)).  Intra pairs are Bernoulli-sampled exactly; inter
    pairs are drawn uniformly (count ~ Binomial) and deduplicated."""
    rng = np.random.default_rng(seed)
    sizes = np.full(blocks, n // blocks, np.int64)
    sizes[: n % blocks] += 1
    starts = np.concatenate([[0], np.cumsum(sizes)])
    src, dst = [], []
    for b in range(blocks):
        s, m = starts[b], sizes[b]
        if m < 2:
            continue
        p_in = min(1.0, d_in / (m - 1))
        iu, ju = np.triu_indices(m, 1)
        sel = rng.random(len(iu)) < p_in
        src.append(iu[sel] + s)
        dst.append(ju[sel] + s)
    block_of = np.repeat(np.arange(blocks), sizes)
    mean_b = n / blocks
    inter_pairs = n * (n - mean_b) / 2.0
    p_out = d_out / max(n - mean_b, 1)
    m_out = rng.binomial(int(inter_pairs), min(1.0, p_out)) if inter_pairs > 0 else 0
    u = rng.integers(0, n, int(m_out * 1.1) + 16)
    v = rng.integers(0, n, len(u))
    ok = block_of[u] != block_of[v]
    u, v = u[ok][:m_out], v[ok][:m_out]
    src.append(u)
    dst.append(v)
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    w = None
    if weights is not None:
        w = rng.integers(weights[0], weights[1] + 1, len(src))
    if permute_seed is not None:
        perm = np.random.default_rng(permute_seed).permutation(n)
        src, dst = perm[src], perm[dst]
    return from_edges(n, src, dst, w)


def rmat_edges_numpy(scale, edge_factor, a, b, c, seed):
    """R-MAT edge draws (Chakrabarti et al. 2004), no noise."""
    rng = np.random.default_rng(seed)
    m = edge_factor << scale
    src = np.zeros(m, np.int64)
    dst = np.zeros(m, np.int64)
    for lvl in range(scale):
        r = rng.random(m)
        bit_s = r >= a + b            # quadrants c, d
        bit_d = ((r >= a) & (r < a + b)) | (r >= a + b + c)  # quadrants b, d
        src |= bit_s.astype(np.int64) << (scale - 1 - lvl)
        dst |= bit_d.astype(np.int64) << (scale - 1 - lvl)
    return src, dst


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=2, weights=None):
    n = 1 << scale
    src, dst = rmat_edges_numpy(scale, edge_factor, a, b, c, seed)
    w = None
    if weights is not None:
        w = np.random.default_rng(seed + 7919).integers(weights[0], weights[1] + 1, len(src))
    return from_edges(n, src, dst, w)


def _devgraph_tensors(h, ctx):
    """Copy a library-owned device graph into torch tensors and free it."""
    import ctypes
    import torch
    from . import _lib
    lib = ctx.lib
    try:
        v = _lib.Graph()
        _lib.check(lib.pggc_devgraph_view(h, ctypes.byref(v)))
        dev = torch.device("cuda")
        rp = torch.empty(v.n + 1, dtype=torch.int64, device=dev)
        col = torch.empty(max(v.nnz, 1), dtype=torch.int32, device=dev)
        w = torch.empty(max(v.nnz, 1), dtype=torch.int32, device=dev) if v.w else None
        _lib.check(lib.pggc_devgraph_copy(ctx.h, h, ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(col.data_ptr()),
                                          None if w is None else ctypes.c_void_p(w.data_ptr())))
        torch.cuda.current_stream().synchronize()
    finally:
        lib.pggc_devgraph_free(h)
    nnz = int(v.nnz)
    return rp, col[:nnz], None if w is None else w[:nnz]


def rmat_device(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=2, weights=None, ctx=None):
    """R-MAT graph built on the GPU by the library (pggc_rmat): Philox4x32-10
    draws, symmetrised, deduplicated, no self loops, sorted rows; weights
    uniform in [lo, hi] per undirected pair when `weights` = (lo, hi).  The
    oracle's generator builds the identical graph on the CPU.  Returns device
    tensors (rp int64, col int32, w int32 or None)."""
    import ctypes
    from . import _lib
    from .clustering import default_context
    ctx = ctx or default_context()
    wmin, wmax = (0, 0) if weights is None else weights
    h = ctypes.c_void_p()
    _lib.check(ctx.lib.pggc_rmat(ctx.h, scale, edge_factor, a, b, c, seed, wmin, wmax, ctypes.byref(h)))
    return _devgraph_tensors(h, ctx)


def lfr_device(n, tau1=2.5, kmean=20.0, kmax=1000, tau2=1.5, cmin=20, cmax=2000, mu=0.4, weights=None, seed=3,
               ground_truth=False, ctx=None):
    """LFR-style benchmark graph built on the GPU by the library (pggc_lfr):
    power-law degrees and community sizes, mixing mu, stub matching inside
    and across communities, permuted ids, per-pair weights uniform in
    `weights` = (lo, hi).  The oracle's generator builds the identical graph
    on the CPU.  Returns device tensors (rp, col, w or None) and, with
    ground_truth, the planted community of every vertex."""
    import ctypes
    import torch
    from . import _lib
    from .clustering import default_context
    ctx = ctx or default_context()
    wmin, wmax = (0, 0) if weights is None else weights
    gt = torch.empty(max(n, 1), dtype=torch.int32, device="cuda") if ground_truth else None
    h = ctypes.c_void_p()
    _lib.check(ctx.lib.pggc_lfr(ctx.h, n, tau1, kmean, kmax, tau2, cmin, cmax, mu, wmin, wmax, seed,
                                ctypes.c_void_p(gt.data_ptr()) if gt is not None else None, ctypes.byref(h)))
    rp, col, w = _devgraph_tensors(h, ctx)
    return (rp, col, w, gt[:n]) if ground_truth else (rp, col, w)


def from_edges_device(n, src, dst, w=None, ctx=None):
    """construct_from_edgelist on the GPU (pggc_graph_from_edges): device
    int32 src/dst (and int32 weights) -> (rp, col, w) device tensors."""
    import ctypes
    import torch
    from . import _lib
    from .clustering import default_context
    ctx = ctx or default_context()
    src = src.to(torch.int32).contiguous()
    dst = dst.to(torch.int32).contiguous()
    w = None if w is None else w.to(torch.int32).contiguous()
    h = ctypes.c_void_p()
    _lib.check(ctx.lib.pggc_graph_from_edges(ctx.h, n, src.numel(), ctypes.c_void_p(src.data_ptr()),
                                             ctypes.c_void_p(dst.data_ptr()),
                                             None if w is None else ctypes.c_void_p(w.data_ptr()),
                                             ctypes.byref(h)))
    return _devgraph_tensors(h, ctx)


def random_labels_device(n, L, seed, ctx=None):
    """labels[i] = Philox4x32-10(counter (i, 0, 0, 2), key seed)[0] mod L on
    the GPU (pggc_random_labels)."""
    import ctypes
    import torch
    from . import _lib
    from .clustering import default_context
    ctx = ctx or default_context()
    out = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    _lib.check(ctx.lib.pggc_random_labels(ctx.h, n, L, seed, ctypes.c_void_p(out.data_ptr())))
    return out[:n]


def _csr_from_pairs_torch(lo, hi, n, device):
    """Symmetric deduplicated CSR (rp int64, col int32) from undirected pairs lo < hi."""
    import torch
    key = torch.unique(lo * n + hi)
    lo, hi = key // n, key % n
    k2, _ = torch.sort(torch.cat([lo * n + hi, hi * n + lo]))
    col = (k2 % n).to(torch.int32)
    cnt = torch.bincount(k2 // n, minlength=n)
    rp = torch.zeros(n + 1, dtype=torch.int64, device=device)
    rp[1:] = torch.cumsum(cnt, 0)
    return rp, col


def sbm_torch(n, blocks, d_in, d_out, seed, device="cuda"):
    """Planted-partition graph built with torch on `device`: vertex i is in
    block i*blocks//n; each vertex draws Poisson(d_in/2) partners uniformly
    from its own block and Poisson(d_out/2) uniformly from the whole graph
    (partners in its own block dropped), so expected intra and inter degrees
    are about d_in and d_out after symmetrising.  Self loops and repeated
    pairs are removed.  A fast stand-in for :func:`sbm` (whose per-pair
    Bernoulli draws run on the host) for batches of many graphs; the two give
    different graphs.  Returns (rp int64, col int32) on the device."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    v = torch.arange(n, dtype=torch.int64, device=device)
    blk = v * blocks // n
    start = (blk * n + blocks - 1) // blocks  # first vertex of each vertex's block
    end = ((blk + 1) * n + blocks - 1) // blocks
    k_in = torch.poisson(torch.full((n,), d_in / 2.0, device=device), generator=gen).to(torch.int64)
    src = torch.repeat_interleave(v, k_in)
    span = (end - start)[src]
    dst = start[src] + (torch.rand(src.numel(), generator=gen, device=device) * span).to(torch.int64).clamp_max(span - 1)
    k_out = torch.poisson(torch.full((n,), d_out / 2.0, device=device), generator=gen).to(torch.int64)
    s2 = torch.repeat_interleave(v, k_out)
    d2 = torch.randint(0, n, (s2.numel(),), generator=gen, device=device)
    keep2 = blk[d2] != blk[s2]
    src = torch.cat([src, s2[keep2]])
    dst = torch.cat([dst, d2[keep2]])
    keep = src != dst
    src, dst = src[keep], dst[keep]
    return _csr_from_pairs_torch(torch.minimum(src, dst), torch.maximum(src, dst), n, device)


def karate():
    """Zachary's karate club (34 vertices, 78 edges)."""
    edges = [
        (0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (0, 6), (0, 7), (0, 8), (0, 10), (0, 11), (0, 12), (0, 13),
        (0, 17), (0, 19), (0, 21), (0, 31), (1, 2), (1, 3), (1, 7), (1, 13), (1, 17), (1, 19), (1, 21), (1, 30),
        (2, 3), (2, 7), (2, 8), (2, 9), (2, 13), (2, 27), (2, 28), (2, 32), (3, 7), (3, 12), (3, 13), (4, 6),
        (4, 10), (5, 6), (5, 10), (5, 16), (6, 16), (8, 30), (8, 32), (8, 33), (9, 33), (13, 33), (14, 32),
        (14, 33), (15, 32), (15, 33), (18, 32), (18, 33), (19, 33), (20, 32), (20, 33), (22, 32), (22, 33),
        (23, 25), (23, 27), (23, 29), (23, 32), (23, 33), (24, 25), (24, 27), (24, 31), (25, 31), (26, 29),
        (26, 33), (27, 33), (28, 31), (28, 33), (29, 32), (29, 33), (30, 32), (30, 33), (31, 32), (31, 33),
        (32, 33),
    ]
    e = np.array(edges)
    return from_edges(34, e[:, 0], e[:, 1])


def random_graph(n, avg_deg, seed, weights=None):
    """Erdos-Renyi style multigraph draw, deduplicated."""
    rng = np.random.default_rng(seed)
    m = int(n * avg_deg / 2)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    w = None if weights is None else rng.integers(weights[0], weights[1] + 1, m)
    return from_edges(n, u, v, w)


def read_graph(path, objective="Mod"):
    """load_graph (app/io/io_mtx.cpp:337-346) through the library's reader:
    ``.mtx`` files are Matrix Market, anything else METIS; edge weights are
    dropped for the objectives that treat edges as unit (Mod, CPM).  Host
    only; returns a CSR."""
    import ctypes
    from . import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    weighted = ctypes.c_int()
    unit = 1 if objective in ("Mod", "CPM") else 0
    _lib.check(lib.pggc_read_graph(str(path).encode(), unit, ctypes.byref(h), ctypes.byref(weighted)))
    try:
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.pggc_host_graph_info(h, ctypes.byref(n), ctypes.byref(nnz)))
        rp = np.empty(n.value + 1, np.int64)
        col = np.empty(nnz.value, np.int32)
        w = np.empty(nnz.value, np.int32) if weighted.value else None
        _lib.check(lib.pggc_host_graph_copy(h, rp.ctypes.data, col.ctypes.data, None if w is None else w.ctypes.data))
    finally:
        lib.pggc_host_graph_free(h)
    return CSR(n.value, rp, col, w)


def write_metis(g, path):
    """METIS text: header 'n m [1]', one line per vertex with 1-based
    neighbours (and weights when the graph has them)."""
    with open(path, "w") as f:
        f.write(f"{g.n} {g.nnz // 2}" + (" 1" if g.w is not None else "") + "\n")
        for i in range(g.n):
            a, b = g.rp[i], g.rp[i + 1]
            if g.w is None:
                f.write(" ".join(str(v + 1) for v in g.col[a:b]) + "\n")
            else:
                f.write(" ".join(f"{v + 1} {x}" for v, x in zip(g.col[a:b], g.w[a:b])) + "\n")


def write_mtx(g, path, symmetric=True):
    """Matrix Market coordinate file: the lower triangle when symmetric,
    every directed entry otherwise (pattern when unweighted)."""
    field = "pattern" if g.w is None or not symmetric else "integer"
    rows = np.repeat(np.arange(g.n), np.diff(g.rp))
    keep = (rows > g.col) if symmetric else np.ones(len(rows), bool)
    r, c = rows[keep], g.col[keep]
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {'symmetric' if symmetric else 'general'}\n")
        f.write("% written by graphs.write_mtx\n")
        f.write(f"{g.n} {g.n} {len(r)}\n")
        if field == "integer":
            for u, v, x in zip(r, c, g.w[keep]):
                f.write(f"{u + 1} {v + 1} {x}\n")
        else:
            for u, v in zip(r, c):
                f.write(f"{u + 1} {v + 1}\n")
