"""Config C5 readiness: a graph of more than 2^31 nonzeros, built on the
device (R-MAT scale 26, edge factor 20, seed 5: 1.34e9 draws, about 2.6e9
nonzeros after symmetrising), and a random labelling into 2^22 clusters.

Size-independent properties, checked against an independent torch
computation over edge chunks (row of an edge by a binary search of the row
offsets):
- evaluator: volumes are the label-wise degree sums, the uncut weight is the
  number of edges inside clusters, and the allocation-free evaluator gives
  the same volumes, uncut weight and exact sum of squares;
- contraction: the coarse graph has L rows, sorted strictly increasing
  columns without self loops, positive weights summing to nnz minus the
  uncut weight, and row sums equal to vol[a] minus the inside weight of
  cluster a;
- a whole pLouvain run and a pLeiden+ pass: the objective each reports is
  the evaluator's objective of the partition it returns."""
import ctypes

import pytest

from pennsylvaniagpugraphclustering_amd import _lib, graphs

pytestmark = pytest.mark.gpu

SCALE, EF, SEED, L = 26, 20, 5, 1 << 22
CHUNK = 1 << 28


@pytest.fixture(scope="module")
def big(gpu_ctx):
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    print(f"build rmat s{SCALE} ef{EF}", flush=True)
    rp, col, _ = graphs.rmat_device(SCALE, EF, seed=SEED)
    dg = C.DeviceGraph(rp, col, None, None, objective="Mod")
    assert dg.nnz > (1 << 31)
    lab = graphs.random_labels_device(dg.n, L, SEED + 1)
    # independent reference values, chunk by chunk over the edges
    print(f"torch reference over nnz={dg.nnz}", flush=True)
    inside = 0
    intra = torch.zeros(L, dtype=torch.int64, device="cuda")
    for j0 in range(0, dg.nnz, CHUNK):
        j = torch.arange(j0, min(j0 + CHUNK, dg.nnz), dtype=torch.int64, device="cuda")
        row = torch.searchsorted(rp, j, right=True) - 1
        lr = lab[row]
        same = lr == lab[col[j].to(torch.int64)]
        inside += int(same.sum())
        intra.index_add_(0, lr[same].to(torch.int64), torch.ones(int(same.sum()), dtype=torch.int64,
                                                                 device="cuda"))
        del j, row, lr, same
    deg = (rp[1:] - rp[:-1]).to(torch.int64)
    vol = torch.zeros(L, dtype=torch.int64, device="cuda").index_add_(0, lab.to(torch.int64), deg)
    yield dg, rp, col, lab, vol, inside, intra
    del dg, rp, col, lab
    torch.cuda.empty_cache()


def test_evaluator_beyond_2_31_nonzeros(big, gpu_ctx):
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    dg, _, _, lab, vol_ref, inside, _ = big
    print("evaluate", flush=True)
    vol, uncut, sq = C.evaluate(dg, lab, L)
    assert torch.equal(vol, vol_ref)
    assert uncut == inside
    v = vol_ref.cpu().numpy().astype(object)
    assert sq == int((v * v).sum())
    print("evaluate_noalloc", flush=True)
    vol2 = torch.full((L,), -7, dtype=torch.int64, device="cuda")
    uncut2 = ctypes.c_int64(-1)
    sq2 = (ctypes.c_uint64 * 2)()
    _lib.check(gpu_ctx.lib.pggc_evaluate_noalloc(gpu_ctx.h, ctypes.byref(dg.c()), ctypes.c_void_p(lab.data_ptr()),
                                                 L, ctypes.c_void_p(vol2.data_ptr()), ctypes.byref(uncut2), sq2))
    assert torch.equal(vol2, vol)
    assert uncut2.value == uncut
    assert ((int(sq2[0]) << 64) | int(sq2[1])) == sq


def test_contraction_beyond_2_31_nonzeros(big, gpu_ctx):
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    dg, _, _, lab, vol_ref, inside, intra = big
    print("contract", flush=True)
    crp, ccol, cw = C.build_coarse_graph(dg, lab, L)
    nnzc = ccol.numel()
    print(f"coarse nnz={nnzc}", flush=True)
    assert crp.numel() == L + 1 and int(crp[0]) == 0 and int(crp[-1]) == nnzc
    assert bool((crp[1:] >= crp[:-1]).all())
    rowsum = torch.zeros(L, dtype=torch.int64, device="cuda")
    total = 0
    for j0 in range(0, nnzc, CHUNK):
        j1 = min(j0 + CHUNK, nnzc)
        j = torch.arange(j0, j1, dtype=torch.int64, device="cuda")
        row = torch.searchsorted(crp, j, right=True) - 1
        c = ccol[j0:j1].to(torch.int64)
        w = cw[j0:j1]
        assert bool((w > 0).all())
        assert not bool((c == row).any())                     # no self loops
        assert bool(((c >= 0) & (c < L)).all())
        # strictly increasing columns inside a row (the first edge of a chunk
        # is compared with the last of the previous one)
        if j1 - j0 > 1:
            same_row = row[1:] == row[:-1]
            assert bool((c[1:][same_row] > c[:-1][same_row]).all())
        if j0 > 0 and int(row[0]) == int(torch.searchsorted(crp, torch.tensor([j0 - 1], device="cuda"),
                                                           right=True)[0]) - 1:
            assert int(ccol[j0 - 1]) < int(c[0])
        rowsum.index_add_(0, row, w)
        total += int(w.sum())
        del j, row, c, w
    assert total == dg.nnz - inside
    assert torch.equal(rowsum, vol_ref - intra)


def test_louvain_beyond_2_31_nonzeros(big, gpu_ctx):
    """One pLouvain run (ex_iters 0) on the graph: the objective it reports
    is the evaluator's objective of the partition it returns, and the
    labels are a partition into `labels` clusters."""
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    dg = big[0]
    print("cluster", flush=True)
    q, part, info = C.cluster(dg, "louvain", "Mod", 1.0, 0, seed=1)
    print(f"n={dg.n} nnz={dg.nnz} Q={q} info={info}", flush=True)
    assert part.numel() == dg.n
    k = info["labels"]
    assert int(part.min()) >= 0 and int(part.max()) < k
    assert int(torch.unique(part).numel()) == k
    assert q == C.modularity(dg, part)
    assert 0.0 < q < 1.0


def test_leiden_plus_beyond_2_31_nonzeros(big, gpu_ctx):
    """One pLeiden+ pass (ex_iters 0) on the graph: the reported objective is
    the evaluator's, and the labels are a partition into `labels` clusters."""
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    dg = big[0]
    print("cluster leiden+", flush=True)
    q, part, info = C.cluster(dg, "leiden+", "Mod", 1.0, 0, seed=1)
    print(f"Q={q} info={info}", flush=True)
    k = info["labels"]
    assert int(part.min()) >= 0 and int(part.max()) < k
    assert int(torch.unique(part).numel()) == k
    assert q == C.modularity(dg, part)
    assert 0.0 < q < 1.0
