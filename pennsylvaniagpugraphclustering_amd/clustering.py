"""Host-side mirror of the reference's clustering interface over libpggc.so.

Names follow the reference: ``louvain_part`` / ``leiden_part``
(interface/clustering_methods.h:20-24), ``generate_orderings``
(src/header/ordering.h:21), ``build_coarse_graph`` (src/header/contract.h:20),
``coarsen_leidenR`` (src/header/leidenR.h:26), ``local_move`` /
``local_move_strict`` (src/header/local_mover.h:21-25) and the evaluator of
app/verify_objective.cpp.  Device memory and streams come from torch; all
compute runs in the HIP library.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import Graph, State, check, OBJECTIVES, ALGOS


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("pggc needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda")


def _t(x, dtype):
    """numpy/torch -> contiguous device tensor of dtype."""
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        return x.to(device=_dev(), dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x)).to(device=_dev(), dtype=dtype)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Context:
    """pggc_ctx bound to torch's current stream."""

    def __init__(self, stream=None):
        self.lib = _lib.load()
        _dev()
        s = stream if stream is not None else torch.cuda.current_stream()
        self.stream = s
        h = ctypes.c_void_p()
        check(self.lib.pggc_ctx_create(ctypes.byref(h), ctypes.c_void_p(s.cuda_stream)))
        self.h = h

    def close(self):
        if self.h:
            check(self.lib.pggc_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_verbose(self, level=1):
        check(self.lib.pggc_ctx_set_verbose(self.h, int(level)))

    def memory(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.pggc_ctx_memory(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    KERNEL_SLOTS = ("score_group", "score_table", "score_hub", "afterburner", "moves", "contract_agg",
                    "contract_big", "contract_copy", "eval", "leiden", "order")

    def set_timing(self, on=True):
        """Record HIP events on the context stream around the timed kernel
        groups (resets the totals)."""
        check(self.lib.pggc_ctx_set_timing(self.h, 1 if on else 0))

    def kernel_stats(self):
        """{slot: (launches, milliseconds, algorithmic bytes)} since set_timing."""
        out = {}
        for i, name in enumerate(self.KERNEL_SLOTS):
            n, ms, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
            check(self.lib.pggc_ctx_kernel_stats(self.h, i, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)))
            out[name] = (n.value, ms.value, by.value)
        return out

    KNOBS = {"hub_update_div": 0, "gnn_max_sweeps": 1, "table_threads": 2, "table_slots_per_edge": 3,
             "hub_lds_max_degree": 4, "hub_edges_per_item": 5, "hub_slots_per_item": 6, "raw_sweeps": 7,
             "hub_build": 8, "renumber": 9, "contract_batch": 10}

    # values the execution-shape knobs accept (results are identical for all)
    KNOB_VALUES = {"table_threads": (64, 128, 256), "table_slots_per_edge": (2, 4),
                   "hub_lds_max_degree": (2048, 4096), "hub_edges_per_item": (512, 1024, 2048),
                   "hub_slots_per_item": (2048, 4096, 8192), "hub_update_div": (16, 32, 64)}

    def set_knob(self, name, value):
        """Tuning knob of the HIP path (gnn_max_sweeps: predicted sweeps per
        local-move call; every other knob changes time only)."""
        check(self.lib.pggc_ctx_set_knob(self.h, self.KNOBS[name], int(value)))

    def get_knob(self, name):
        v = ctypes.c_int64()
        check(self.lib.pggc_ctx_get_knob(self.h, self.KNOBS[name], ctypes.byref(v)))
        return v.value

    def autotune(self, g, objective="Mod", reps=2):
        """pggc_autotune: coordinate descent over the execution-shape knobs,
        timing one pLouvain clustering of g per value; the fastest values stay
        set.  Returns (chosen {knob: value}, [(knob, value, us), ...])."""
        names = {v: k for k, v in self.KNOBS.items()}
        cap = 64
        rep = (ctypes.c_int64 * (3 * cap))()
        rows = ctypes.c_int()
        check(self.lib.pggc_autotune(self.h, ctypes.byref(g.c()), OBJECTIVES[objective], int(reps), rep, cap,
                                     ctypes.byref(rows)))
        table = [(names[rep[3 * i]], int(rep[3 * i + 1]), int(rep[3 * i + 2])) for i in range(min(rows.value, cap))]
        return {k: self.get_knob(k) for k in self.KNOB_VALUES}, table

    SWEEP_FIELDS = ("n", "nnz", "phase", "sweep", "active", "cands", "moves", "us")
    _SweepFn = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int)

    def set_sweep_trace(self, records):
        """Append one dict per local-move sweep (SWEEP_FIELDS) to the list
        `records`; None turns the trace off.  Synchronises every sweep."""
        if records is None:
            self._sweep_cb = None
            check(self.lib.pggc_ctx_set_sweep_trace(self.h, None, None))
            return

        def cb(_user, rec, nrec):
            records.append({k: int(rec[i]) for i, k in enumerate(self.SWEEP_FIELDS[:nrec])})

        self._sweep_cb = self._SweepFn(cb)
        check(self.lib.pggc_ctx_set_sweep_trace(self.h, ctypes.cast(self._sweep_cb, ctypes.c_void_p), None))


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context()
    return _default


class DeviceGraph:
    """weighted_graph (interface/weighted_graph.h:10-19) resident in HBM."""

    def __init__(self, rp, col, w=None, vtx_w=None, n=None, objective=None, ctx=None):
        self.rp = _t(rp, torch.int64)
        self.col = _t(col, torch.int32)
        self.w = None
        self.w64 = None
        if w is not None:
            wt = w if isinstance(w, torch.Tensor) else torch.from_numpy(np.asarray(w))
            if wt.dtype == torch.int64:
                self.w64 = _t(wt, torch.int64)
            else:
                self.w = _t(wt, torch.int32)
        self.n = int(self.rp.numel() - 1) if n is None else int(n)
        self.nnz = int(self.col.numel())
        self.vtx_w = _t(vtx_w, torch.int64) if vtx_w is not None else None
        if self.vtx_w is None and objective is not None:
            self.vtx_w = vertex_weights(self, objective, ctx)

    @classmethod
    def from_csr(cls, g, objective="Mod", ctx=None):
        w = None if objective in ("Mod", "CPM") else g.w
        return cls(g.rp, g.col, w, g.vertex_weights(objective), n=g.n)

    def c(self):
        return Graph(self.n, self.nnz, self.rp.data_ptr(), self.col.data_ptr() if self.nnz else None,
                     self.w.data_ptr() if self.w is not None else None,
                     self.w64.data_ptr() if self.w64 is not None else None,
                     self.vtx_w.data_ptr() if self.vtx_w is not None else None)


def _ctx(ctx):
    return ctx if ctx is not None else default_context()


def vertex_weights(g, objective="Mod", ctx=None):
    ctx = _ctx(ctx)
    out = torch.empty(max(g.n, 1), dtype=torch.int64, device=_dev())
    check(ctx.lib.pggc_vertex_weights(ctx.h, ctypes.byref(g.c()), OBJECTIVES[objective], _p(out)))
    return out[: g.n]


def generate_orderings(g, ctx=None):
    ctx = _ctx(ctx)
    o1 = torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev())
    o2 = torch.empty_like(o1)
    offs = (ctypes.c_int64 * 4)()
    check(ctx.lib.pggc_orderings(ctx.h, ctypes.byref(g.c()), _p(o1), _p(o2), offs))
    return o1[: g.n], o2[: g.n], list(offs)


def evaluate(g, labels, L, ctx=None):
    """Returns (vol[L] device tensor, uncut, exact sum of squared volumes)."""
    ctx = _ctx(ctx)
    lab = _t(labels, torch.int32)
    vol = torch.empty(max(L, 1), dtype=torch.int64, device=_dev())
    uncut = ctypes.c_int64()
    sq = (ctypes.c_uint64 * 2)()
    check(ctx.lib.pggc_evaluate(ctx.h, ctypes.byref(g.c()), _p(lab), L, _p(vol), ctypes.byref(uncut), sq))
    return vol[:L], uncut.value, (int(sq[0]) << 64) | int(sq[1])


def modularity(g, labels, objective="Mod", mult=1.0, ctx=None):
    """verify_objective (app/verify_objective.cpp:64-89): objective value of a
    clustering (labels must be < n)."""
    st, vol = state_init(g, objective, mult, ctx)
    _, uncut, sq = evaluate(g, labels, g.n, ctx)
    obj = float(uncut) - st.lambda_ * float(sq)
    return obj * st.scale + st.offset


def own_weight(g, labels, ctx=None):
    ctx = _ctx(ctx)
    out = torch.empty(max(g.n, 1), dtype=torch.int64, device=_dev())
    check(ctx.lib.pggc_own_weight(ctx.h, ctypes.byref(g.c()), _p(_t(labels, torch.int32)), _p(out)))
    return out[: g.n]


def build_coarse_graph(g, vcmap, nc, ctx=None):
    """Returns (rp, col, w) device tensors of the contracted graph."""
    ctx = _ctx(ctx)
    m = _t(vcmap, torch.int32)
    rp = torch.empty(nc + 1, dtype=torch.int64, device=_dev())
    cap = max(g.nnz, 1)
    col = torch.empty(cap, dtype=torch.int32, device=_dev())
    w = torch.empty(cap, dtype=torch.int64, device=_dev())
    nnz = ctypes.c_int64()
    check(ctx.lib.pggc_contract(ctx.h, ctypes.byref(g.c()), _p(m), nc, _p(rp), _p(col), _p(w), cap,
                                ctypes.byref(nnz)))
    return rp, col[: nnz.value], w[: nnz.value]


def sweep(g, part, vol, lam, ratio, raw, dest, gain, constraint=None, ctx=None):
    """One candidates + afterburner sweep; returns (moves, dest, gain, ncand)."""
    ctx = _ctx(ctx)
    part = _t(part, torch.int32)
    vol = _t(vol, torch.int64)
    dest = _t(dest, torch.int32).clone()
    gain = _t(gain, torch.float32).clone()
    cons = _t(constraint, torch.int32)
    moves = torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev())
    nm, nc = ctypes.c_int64(), ctypes.c_int64()
    check(ctx.lib.pggc_sweep(ctx.h, ctypes.byref(g.c()), _p(part), _p(vol), vol.numel(), lam, ratio, int(raw),
                             _p(cons), _p(dest), _p(gain), _p(moves), ctypes.byref(nm), ctypes.byref(nc)))
    return moves[: nm.value], dest, gain, nc.value


def state_init(g, objective="Mod", mult=1.0, ctx=None):
    """get_objective (app/objective_helpers.hpp:9-22)."""
    ctx = _ctx(ctx)
    vol = torch.empty(max(g.n, 1) + 1, dtype=torch.int64, device=_dev())
    st = State()
    st.vol = vol.data_ptr()
    st.vol_cap = g.n
    check(ctx.lib.pggc_state_init(ctx.h, ctypes.byref(g.c()), OBJECTIVES[objective], mult, ctypes.byref(st)))
    return st, vol


def make_state(vol_values, L, uncut, lam, obj, top_nnz=0, cap=None):
    vals = _t(vol_values, torch.int64)
    cap = max(cap or vals.numel(), L)
    vol = torch.zeros(cap + 1, dtype=torch.int64, device=_dev())
    vol[: vals.numel()] = vals
    st = State()
    st.vol = vol.data_ptr()
    st.vol_cap = cap
    st.uncut = uncut
    st.top_nnz = top_nnz
    st.obj = obj
    st.lambda_ = lam
    st.scale = 1.0
    st.offset = 0.0
    st.label_count = L
    return st, vol


def local_move(g, part, st, is_initial, strict=False, simulated_annealing=True, ctx=None):
    ctx = _ctx(ctx)
    part = _t(part, torch.int32).clone()
    check(ctx.lib.pggc_local_move(ctx.h, ctypes.byref(g.c()), _p(part), ctypes.byref(st), int(is_initial),
                                  int(strict), int(simulated_annealing)))
    return part


def local_move_strict(g, part, st, is_initial, ctx=None):
    return local_move(g, part, st, is_initial, strict=True, ctx=ctx)


def coarsen_leidenR(g, louv, vol, lam, uniform, seed, draw0=0, ctx=None):
    ctx = _ctx(ctx)
    lv = _t(louv, torch.int32)
    vv = _t(vol, torch.int64)
    out = torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev())
    nc = ctypes.c_int64()
    check(ctx.lib.pggc_leiden_refine(ctx.h, ctypes.byref(g.c()), _p(lv), _p(vv), vv.numel(), lam, int(uniform), seed,
                                     draw0, _p(out), ctypes.byref(nc)))
    return out[: g.n], nc.value


def louvain_part(g, st, constrained=False, constraint=None, ctx=None):
    """louvain_part<constrained> (interface/clustering_methods.h:23-24)."""
    ctx = _ctx(ctx)
    out = torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev())
    cons = _t(constraint, torch.int32)
    check(ctx.lib.pggc_louvain(ctx.h, ctypes.byref(g.c()), ctypes.byref(st), int(constrained), _p(cons), _p(out)))
    return out[: g.n]


def leiden_part(g, st, plus=False, improve=False, input=None, ctx=None):
    """leiden_part<plus, improve> (interface/clustering_methods.h:20-21)."""
    ctx = _ctx(ctx)
    out = torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev())
    inp = _t(input, torch.int32)
    check(ctx.lib.pggc_leiden(ctx.h, ctypes.byref(g.c()), ctypes.byref(st), int(plus), int(improve), _p(inp),
                              _p(out)))
    return out[: g.n]


INFO_KEYS = ["labels", "uncut", "levels", "sweeps", "moves", "outer", "cands", "total_nnz"]


def cluster_batch(gs, algo="louvain", objective="Mod", mult=1.0, ex_iters=0, seeds=None, streams=8, ctx=None):
    """pggc_cluster_batch: independent runs on the graphs `gs` (or seeded
    trials of one graph repeated in `gs`).  pLouvain with ex_iters 0 (and the
    predictor off) runs the whole batch as one disjoint union graph, every
    kernel once over all instances; other runs go to `streams` worker host
    threads that share the context's stream.  Returns a list of (objective,
    part device tensor, info dict), each equal to the instance run alone."""
    ctx = _ctx(ctx)
    k = len(gs)
    garr = (Graph * max(k, 1))(*[g.c() for g in gs])
    outs = [torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev()) for g in gs]
    parr = (ctypes.c_void_p * max(k, 1))(*[o.data_ptr() for o in outs])
    sarr = (ctypes.c_uint64 * max(k, 1))(*(seeds if seeds is not None else range(k)))
    qs = (ctypes.c_double * max(k, 1))()
    info = (ctypes.c_int64 * (8 * max(k, 1)))()
    check(ctx.lib.pggc_cluster_batch(ctx.h, k, garr, ALGOS[algo], OBJECTIVES[objective], mult, ex_iters, sarr, parr,
                                     qs, info, int(streams)))
    torch.cuda.synchronize()
    return [(qs[x], outs[x][: gs[x].n], dict(zip(INFO_KEYS, (int(v) for v in info[8 * x: 8 * x + 8]))))
            for x in range(k)]


def cluster(g, algo="louvain", objective="Mod", mult=1.0, ex_iters=0, seed=0, ctx=None):
    """run_clustering (app/cluster_main.cpp:116-165) for one trial.
    Returns (objective value, part device tensor, info dict)."""
    ctx = _ctx(ctx)
    out = torch.empty(max(g.n, 1), dtype=torch.int32, device=_dev())
    q = ctypes.c_double()
    info = (ctypes.c_int64 * 8)()
    check(ctx.lib.pggc_cluster(ctx.h, ctypes.byref(g.c()), ALGOS[algo], OBJECTIVES[objective], mult, ex_iters, seed,
                               _p(out), ctypes.byref(q), info))
    return q.value, out[: g.n], dict(zip(INFO_KEYS, (int(x) for x in info)))


# ------------------------------------------------------ active-set predictor
GNN_PARAMS = 609
GNN_MODES = {"off": 0, "train": 1, "prune": 2}
GNN_STAT_KEYS = ("sweeps", "active", "pruned", "restores", "train_steps", "tp", "fp", "fn", "tn")


def gnn_configure(mode, tau=0.0, min_nnz=1 << 20, full_every=0, lr=1e-3, pos_weight=4.0, ctx=None):
    """Predictor mode for the local mover of every later clustering call on
    the context: "off", "train" (plain sweeps, one Adam step after each) or
    "prune" (active vertices with logit < tau are skipped)."""
    ctx = _ctx(ctx)
    check(ctx.lib.pggc_gnn_configure(ctx.h, GNN_MODES[mode], float(tau), int(min_nnz), int(full_every), float(lr),
                                     float(pos_weight)))


def gnn_set_params(params, ctx=None):
    ctx = _ctx(ctx)
    p = np.ascontiguousarray(params, dtype=np.float32)
    check(ctx.lib.pggc_gnn_set_params(ctx.h, p.ctypes.data, int(p.size)))


def gnn_get_params(ctx=None):
    ctx = _ctx(ctx)
    p = np.empty(GNN_PARAMS, np.float32)
    check(ctx.lib.pggc_gnn_get_params(ctx.h, p.ctypes.data, GNN_PARAMS))
    return p


def gnn_stats(ctx=None):
    """Counters since the last reset (GNN_STAT_KEYS) plus the summed loss."""
    ctx = _ctx(ctx)
    c = (ctypes.c_int64 * 9)()
    loss = ctypes.c_double()
    check(ctx.lib.pggc_gnn_stats(ctx.h, c, ctypes.byref(loss)))
    out = dict(zip(GNN_STAT_KEYS, (int(x) for x in c)))
    out["loss_sum"] = loss.value
    return out


def gnn_reset_stats(ctx=None):
    ctx = _ctx(ctx)
    check(ctx.lib.pggc_gnn_reset_stats(ctx.h))


def gnn_eval(g, part, dest, gain, stamp, cur, first, vol, labels=None, ctx=None):
    """One predictor pass from a given state.  Returns (input rows (n, 16)
    as uint16 bf16 bits, logits (0 for inactive vertices), loss, gradient);
    loss and gradient are None without labels."""
    ctx = _ctx(ctx)
    n = g.n
    part_t, dest_t = _t(part, torch.int32), _t(dest, torch.int32)
    gain_t, stamp_t = _t(gain, torch.float32), _t(stamp, torch.int32)
    vol_t = _t(vol, torch.int64)
    lab_t = _t(labels, torch.int32) if labels is not None else None
    rows = torch.empty(max(n, 1) * 16, dtype=torch.int16, device=_dev())
    logit = torch.empty(max(n, 1), dtype=torch.float32, device=_dev())
    loss = ctypes.c_double()
    grad = np.zeros(GNN_PARAMS, np.float32)
    check(ctx.lib.pggc_gnn_eval(ctx.h, ctypes.byref(g.c()), _p(part_t), _p(dest_t), _p(gain_t), _p(stamp_t), int(cur),
                                1 if first else 0, _p(vol_t), _p(lab_t) if lab_t is not None else None, _p(rows),
                                _p(logit), ctypes.byref(loss) if lab_t is not None else None,
                                grad.ctypes.data if lab_t is not None else None))
    rows_np = rows[: n * 16].cpu().numpy().view(np.uint16).reshape(n, 16)
    lg = logit[:n].cpu().numpy()
    if lab_t is None:
        return rows_np, lg, None, None
    return rows_np, lg, loss.value, grad
