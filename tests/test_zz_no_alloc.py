"""pggc_evaluate_noalloc (the evaluator variant that allocates nothing; its
scratch is a static device array) against pggc_evaluate, through the C ABI:
volumes, uncut weight and the exact sum of squared volumes must be equal, on
two calls in a row (the second on a smaller graph, so stale scratch from the
first call would show), plus the size limit, which is checked on the host
before anything is launched."""
import ctypes

import numpy as np
import pytest

from pennsylvaniagpugraphclustering_amd import _lib, graphs

pytestmark = pytest.mark.gpu


def _call(ctx, entry, dg, lab, L):
    import torch
    vol = torch.full((max(L, 1),), -7, dtype=torch.int64, device="cuda")
    uncut = ctypes.c_int64(-1)
    sq = (ctypes.c_uint64 * 2)()
    rc = getattr(ctx.lib, entry)(ctx.h, ctypes.byref(dg.c()), ctypes.c_void_p(lab.data_ptr()), L,
                                 ctypes.c_void_p(vol.data_ptr()), ctypes.byref(uncut), sq)
    _lib.check(rc)
    return vol[:L].cpu().numpy(), uncut.value, (int(sq[0]) << 64) | int(sq[1])


def _case(scale, weights, L, seed):
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    rp, col, w = graphs.rmat_device(scale, 16, seed=seed, weights=weights)
    dg = C.DeviceGraph(rp, col, w, None, objective="WMod" if weights else "Mod")
    lab = graphs.random_labels_device(dg.n, L, seed + 11)
    return dg, lab


@pytest.mark.parametrize("first,second", [((15, (1, 100), 4096, 3), (11, None, 50, 5)),
                                          ((14, None, 1, 7), (9, (1, 10), 300, 8))])
def test_noalloc_matches_evaluate_twice(gpu_ctx, first, second):
    import torch
    for scale, weights, L, seed in (first, second):
        print(f"noalloc evaluate rmat s{scale} w={weights} L={L} seed={seed}", flush=True)
        dg, lab = _case(scale, weights, L, seed)
        ref = _call(gpu_ctx, "pggc_evaluate", dg, lab, L)
        got = _call(gpu_ctx, "pggc_evaluate_noalloc", dg, lab, L)
        torch.cuda.synchronize()
        assert np.array_equal(got[0], ref[0])
        assert got[1] == ref[1] and got[2] == ref[2]
        # the sum of squares is the exact integer sum
        assert got[2] == sum(int(v) * int(v) for v in ref[0])


def test_noalloc_empty_graph(gpu_ctx):
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C
    rp = torch.zeros(1, dtype=torch.int64, device="cuda")
    col = torch.zeros(0, dtype=torch.int32, device="cuda")
    dg = C.DeviceGraph(rp, col, None, None, objective="Mod")
    lab = torch.zeros(0, dtype=torch.int32, device="cuda")
    vol, uncut, sq = _call(gpu_ctx, "pggc_evaluate_noalloc", dg, lab, 0)
    assert uncut == 0 and sq == 0 and vol.size == 0


def test_noalloc_rejects_graphs_beyond_the_scratch(gpu_ctx):
    max_nnz = ctypes.c_int64()
    nbytes = ctypes.c_int64()
    _lib.check(gpu_ctx.lib.pggc_evaluate_noalloc_limits(ctypes.byref(max_nnz), ctypes.byref(nbytes)))
    assert max_nnz.value == 1 << 32 and 0 < nbytes.value <= 16 << 20
    # rejected on the host before any launch: the pointers are never used
    g = _lib.Graph(2, max_nnz.value + 1, 8, 8, None, None, 8)
    uncut = ctypes.c_int64()
    sq = (ctypes.c_uint64 * 2)()
    rc = gpu_ctx.lib.pggc_evaluate_noalloc(gpu_ctx.h, ctypes.byref(g), ctypes.c_void_p(8), 2, ctypes.c_void_p(8),
                                           ctypes.byref(uncut), sq)
    assert rc == _lib.ECAPACITY
    assert b"scratch limit" in gpu_ctx.lib.pggc_last_error()
