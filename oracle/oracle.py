"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Every function
takes numpy arrays and returns numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PGGC_ORACLE_LIB selects another build of the oracle (the tests of the
# pinning tests load variants with a planted mistake this way)
_LIB_PATH = os.environ.get("PGGC_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
vp = ctypes.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c = ctypes
        L.or_murmur3_i32.restype = c.c_uint32
        L.or_murmur3_i32.argtypes = [c.c_int32]
        L.or_xorshift32.restype = c.c_uint32
        L.or_xorshift32.argtypes = [c.c_uint32]
        L.or_seed_draw.restype = c.c_uint32
        L.or_seed_draw.argtypes = [c.c_uint64, c.c_uint64]
        L.or_orderings.restype = None
        L.or_orderings.argtypes = [c.c_int64, i64p, i32p, i32p, i64p]
        L.or_evaluate.restype = None
        L.or_evaluate.argtypes = [c.c_int64, i64p, i32p, vp, i64p, i32p, c.c_int64, i64p, i64p, u64p]
        L.or_contract.restype = c.c_int64
        L.or_contract.argtypes = [c.c_int64, i64p, i32p, vp, vp, i32p, c.c_int64, i64p, i32p, i64p, c.c_int64]
        L.or_sweep.restype = c.c_int64
        L.or_sweep.argtypes = [c.c_int64, i64p, i32p, vp, vp, i64p, i32p, i64p, c.c_int64, c.c_double,
                               c.c_float, c.c_int, vp, i32p, f32p, i32p, i64p]
        L.or_own_weight.restype = None
        L.or_own_weight.argtypes = [c.c_int64, i64p, i32p, vp, vp, i32p, i64p]
        L.or_cluster.restype = c.c_double
        L.or_cluster.argtypes = [c.c_int64, i64p, i32p, vp, i64p, c.c_int, c.c_double, c.c_int, c.c_int,
                                 c.c_uint64, i32p, i64p]
        L.or_local_move.restype = c.c_int64
        L.or_local_move.argtypes = [c.c_int64, i64p, i32p, vp, vp, i64p, i32p, i64p, c.c_int64, i64p,
                                    c.c_double, c.c_int, c.c_int, c.c_int, ctypes.POINTER(c.c_double)]
        L.or_set_sweep_limit.restype = None
        L.or_set_sweep_limit.argtypes = [c.c_int]
        L.or_set_threads.restype = None
        L.or_set_threads.argtypes = [c.c_int]
        L.or_get_threads.restype = c.c_int
        L.or_get_threads.argtypes = []
        u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        L.or_philox.restype = None
        L.or_philox.argtypes = [u32p, u32p, u32p]
        L.or_rmat.restype = c.c_void_p
        L.or_rmat.argtypes = [c.c_int, c.c_int, c.c_double, c.c_double, c.c_double, c.c_uint64, c.c_int32,
                              c.c_int32]
        L.or_graph_info.restype = None
        L.or_graph_info.argtypes = [c.c_void_p, c.POINTER(c.c_int64), c.POINTER(c.c_int64), c.POINTER(c.c_int)]
        L.or_graph_copy.restype = None
        L.or_graph_copy.argtypes = [c.c_void_p, i64p, i32p, vp]
        L.or_graph_free.restype = None
        L.or_graph_free.argtypes = [c.c_void_p]
        L.or_lfr.restype = c.c_void_p
        L.or_lfr.argtypes = [c.c_int64, c.c_double, c.c_double, c.c_int32, c.c_double, c.c_int32, c.c_int32,
                             c.c_double, c.c_int32, c.c_int32, c.c_uint64, vp, vp, vp, c.c_int64,
                             c.POINTER(c.c_int64)]
        L.or_power_law_table.restype = c.c_int64
        L.or_power_law_table.argtypes = [c.c_double, c.c_int32, c.c_int32, c.c_double, c.POINTER(c.c_int32), u64p,
                                         c.c_int64]
        L.or_random_labels.restype = None
        L.or_random_labels.argtypes = [c.c_int64, c.c_int64, c.c_uint64, i32p]
        L.or_sweep_detail.restype = c.c_int64
        L.or_sweep_detail.argtypes = L.or_sweep.argtypes + [i32p, f32p]
        L.or_local_move_trace.restype = c.c_int64
        L.or_local_move_trace.argtypes = L.or_local_move.argtypes + [
            np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"), c.c_int64, i64p]
        L.or_leiden_refine.restype = c.c_int64
        L.or_leiden_refine.argtypes = [c.c_int64, i64p, i32p, vp, vp, i64p, i32p, i64p, c.c_int64, c.c_double,
                                       c.c_int, c.c_uint64, c.c_uint64, i32p]
        L.or_leiden_refine_detail.restype = c.c_int64
        L.or_leiden_refine_detail.argtypes = L.or_leiden_refine.argtypes + [i32p, i32p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _w(w):
    """(w32, w64) pointer pair for a weight array (None = unit weights)."""
    if w is None:
        return None, None
    if w.dtype == np.int32:
        return _ptr(w), None
    assert w.dtype == np.int64
    return None, _ptr(w)


def murmur3(key):
    return lib().or_murmur3_i32(int(np.int32(key)))


def xorshift32(x):
    return lib().or_xorshift32(int(x) & 0xFFFFFFFF)


def seed_draw(seed, k):
    return lib().or_seed_draw(seed, k)


def orderings(rp):
    n = len(rp) - 1
    o1 = np.empty(n, np.int32)
    o2 = np.empty(n, np.int32)
    offs = np.empty(4, np.int64)
    lib().or_orderings(n, rp, o1, o2, offs)
    return o1, o2, offs


def evaluate(rp, col, w, vw, labels, L):
    """Returns (vol[L], uncut, sum of squared volumes as a Python int)."""
    n = len(rp) - 1
    vol = np.empty(max(L, 1), np.int64)
    uncut = np.zeros(1, np.int64)
    sq = np.zeros(2, np.uint64)
    assert w is None or w.dtype == np.int32
    lib().or_evaluate(n, rp, col, _ptr(w), vw, labels, L, vol, uncut, sq)
    return vol[:L], int(uncut[0]), (int(sq[0]) << 64) | int(sq[1])


def contract(rp, col, w, vcmap, nc):
    n = len(rp) - 1
    cap = max(int(rp[-1]), 1)
    orp = np.empty(nc + 1, np.int64)
    ocol = np.empty(cap, np.int32)
    ow = np.empty(cap, np.int64)
    w32, w64 = _w(w)
    m = lib().or_contract(n, rp, col, w32, w64, vcmap, nc, orp, ocol, ow, cap)
    assert m >= 0
    return orp, ocol[:m].copy(), ow[:m].copy()


def sweep(rp, col, w, vw, part, vol, lam, ratio, raw, dest, gain, constraint=None):
    n = len(rp) - 1
    dest = dest.copy()
    gain = gain.copy()
    moves = np.empty(max(n, 1), np.int32)
    nc = np.zeros(1, np.int64)
    w32, w64 = _w(w)
    k = lib().or_sweep(n, rp, col, w32, w64, vw, part, vol, len(vol), lam, ratio, int(raw), _ptr(constraint),
                       dest, gain, moves, nc)
    return moves[:k].copy(), dest, gain, int(nc[0])


def own_weight(rp, col, w, labels):
    n = len(rp) - 1
    out = np.empty(n, np.int64)
    w32, w64 = _w(w)
    lib().or_own_weight(n, rp, col, w32, w64, labels, out)
    return out


ALGOS = {"louvain": 0, "leiden": 1, "leiden+": 2}
OBJECTIVES = {"Mod": 0, "WMod": 1, "CPM": 2, "NLCC": 3, "LCC": 4}


def cluster(rp, col, w, vw, objective="Mod", mult=1.0, algo="louvain", ex_iters=0, seed=0):
    """Returns (objective value, part[n], info dict)."""
    n = len(rp) - 1
    part = np.empty(n, np.int32)
    info = np.zeros(6, np.int64)
    assert w is None or w.dtype == np.int32
    q = lib().or_cluster(n, rp, col, _ptr(w), vw, OBJECTIVES[objective], mult, ALGOS[algo], ex_iters, seed,
                         part, info)
    keys = ["labels", "uncut", "levels", "sweeps", "moves", "outer"]
    return q, part, dict(zip(keys, (int(x) for x in info)))


def local_move(rp, col, w, vw, part, vol, L, uncut, lam, obj, is_initial, strict=False, SA=True):
    n = len(rp) - 1
    part = part.copy()
    v = np.zeros(n, np.int64)
    v[:L] = vol[:L]
    u = np.array([uncut], np.int64)
    o = ctypes.c_double(obj)
    w32, w64 = _w(w)
    L2 = lib().or_local_move(n, rp, col, w32, w64, vw, part, v, L, u, lam, int(is_initial), int(strict),
                             int(SA), ctypes.byref(o))
    return part, v[:L2].copy(), int(L2), int(u[0]), o.value


def leiden_refine(rp, col, w, vw, louv, vol, lam, uniform, seed, draw0=0):
    n = len(rp) - 1
    out = np.empty(n, np.int32)
    w32, w64 = _w(w)
    nc = lib().or_leiden_refine(n, rp, col, w32, w64, vw, louv, vol, len(vol), lam, int(uniform), seed, draw0,
                                out)
    return out, int(nc)


def set_threads(k):
    """OpenMP threads of the oracle's per-vertex loops (results do not depend
    on it)."""
    lib().or_set_threads(int(k))


def set_sweep_limit(k):
    """Stop every local_move call after k sweeps (0 = the reference's
    schedule).  Only for bounded CPU timing samples."""
    lib().or_set_sweep_limit(int(k))


def get_threads():
    return int(lib().or_get_threads())


def philox(ctr, key):
    """Philox4x32-10 of a 4-word counter and 2-word key."""
    out = np.empty(4, np.uint32)
    lib().or_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=2, weights=None):
    """Counter-based R-MAT graph (Philox4x32-10 draws), symmetrised,
    deduplicated, without self loops, rows sorted: (rp, col, w or None)."""
    wmin, wmax = (0, 0) if weights is None else weights
    L = lib()
    h = L.or_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax)
    try:
        n, nnz, wd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        L.or_graph_info(h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(wd))
        rp = np.empty(n.value + 1, np.int64)
        col = np.empty(nnz.value, np.int32)
        w = np.empty(nnz.value, np.int32) if wd.value else None
        L.or_graph_copy(h, rp, col, _ptr(w))
    finally:
        L.or_graph_free(h)
    return rp, col, w


def _take_graph(L, h):
    try:
        n, nnz, wd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        L.or_graph_info(h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(wd))
        rp = np.empty(n.value + 1, np.int64)
        col = np.empty(nnz.value, np.int32)
        w = np.empty(nnz.value, np.int32) if wd.value else None
        L.or_graph_copy(h, rp, col, _ptr(w))
    finally:
        L.or_graph_free(h)
    return rp, col, w


def lfr(n, tau1=2.5, kmean=20.0, kmax=1000, tau2=1.5, cmin=20, cmax=2000, mu=0.4, weights=None, seed=3):
    """LFR-style benchmark graph (the oracle's generator; Philox4x32-10
    draws): (rp, col, w or None, planted community per vertex, drawn degree
    per generated vertex, community sizes)."""
    wmin, wmax = (0, 0) if weights is None else weights
    L = lib()
    gt = np.empty(n, np.int32)
    deg = np.empty(n, np.int32)
    cap = n + 1
    sizes = np.empty(cap, np.int64)
    nc = ctypes.c_int64()
    h = L.or_lfr(n, tau1, kmean, kmax, tau2, cmin, cmax, mu, wmin, wmax, seed, gt.ctypes.data, deg.ctypes.data,
                 sizes.ctypes.data, cap, ctypes.byref(nc))
    rp, col, w = _take_graph(L, h)
    return rp, col, w, gt, deg, sizes[: nc.value].copy()


def power_law_table(tau, lo, top, mean=0.0):
    """(k0, thresholds) of the integer power law the generators draw from."""
    thr = np.empty(top - lo + 2, np.uint64)
    k0 = ctypes.c_int32()
    k = lib().or_power_law_table(tau, lo, top, mean, ctypes.byref(k0), thr, len(thr))
    return k0.value, thr[:k].copy()


def random_labels(n, L, seed):
    """labels[i] = Philox4x32-10(counter (i, 0, 0, 2), key seed)[0] mod L."""
    out = np.empty(n, np.int32)
    lib().or_random_labels(n, L, seed, out)
    return out


def sweep_detail(rp, col, w, vw, part, vol, lam, ratio, raw, dest, gain, constraint=None):
    """sweep() plus the candidate list and each candidate's afterburner change."""
    n = len(rp) - 1
    dest = dest.copy()
    gain = gain.copy()
    moves = np.empty(max(n, 1), np.int32)
    cands = np.empty(max(n, 1), np.int32)
    change = np.empty(max(n, 1), np.float32)
    nc = np.zeros(1, np.int64)
    w32, w64 = _w(w)
    k = lib().or_sweep_detail(n, rp, col, w32, w64, vw, part, vol, len(vol), lam, ratio, int(raw),
                              _ptr(constraint), dest, gain, moves, nc, cands, change)
    m = int(nc[0])
    return moves[:k].copy(), dest, gain, cands[:m].copy(), change[:m].copy()


def local_move_trace(rp, col, w, vw, part, vol, L, uncut, lam, obj, is_initial, strict=False, SA=True):
    """local_move() plus its per-sweep records (ratio, candidates, moves,
    objective after the sweep, wall seconds) as an array of shape (sweeps, 5)."""
    n = len(rp) - 1
    part = part.copy()
    v = np.zeros(n, np.int64)
    v[:L] = vol[:L]
    u = np.array([uncut], np.int64)
    o = ctypes.c_double(obj)
    tr = np.zeros(5 * 64, np.float64)
    nt = np.zeros(1, np.int64)
    w32, w64 = _w(w)
    L2 = lib().or_local_move_trace(n, rp, col, w32, w64, vw, part, v, L, u, lam, int(is_initial), int(strict),
                                   int(SA), ctypes.byref(o), tr, 64, nt)
    k = min(int(nt[0]), 64)
    return part, v[:L2].copy(), int(L2), int(u[0]), o.value, tr[:5 * k].reshape(k, 5)


def leiden_refine_detail(rp, col, w, vw, louv, vol, lam, uniform, seed, draw0=0):
    """leiden_refine() plus the forest before the gamma breaks (root per
    vertex) and the order its trees are sorted by."""
    n = len(rp) - 1
    out = np.empty(n, np.int32)
    trees = np.empty(n, np.int32)
    order = np.empty(n, np.int32)
    w32, w64 = _w(w)
    nc = lib().or_leiden_refine_detail(n, rp, col, w32, w64, vw, louv, vol, len(vol), lam, int(uniform), seed,
                                       draw0, out, trees, order)
    return out, int(nc), trees, order
