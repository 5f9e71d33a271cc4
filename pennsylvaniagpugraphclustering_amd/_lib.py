"""ctypes binding of libpggc.so (the HIP library behind include/pggc.h).

The library is loaded from the package directory; a missing or stale build
raises immediately -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpggc.so")

OK = 0
EINVAL, ENOMEM, EHIP, ECAPACITY, EINTERNAL = -1, -2, -3, -4, -5

OBJECTIVES = {"Mod": 0, "WMod": 1, "CPM": 2, "NLCC": 3, "LCC": 4}
ALGOS = {"louvain": 0, "leiden": 1, "leiden+": 2}


class PggcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pggc error {code}: {msg}")
        self.code = code


class Graph(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("row_ptr", ctypes.c_void_p),
        ("col", ctypes.c_void_p),
        ("w", ctypes.c_void_p),
        ("w64", ctypes.c_void_p),
        ("vtx_w", ctypes.c_void_p),
    ]


class State(ctypes.Structure):
    _fields_ = [
        ("vol", ctypes.c_void_p),
        ("vol_cap", ctypes.c_int64),
        ("uncut", ctypes.c_int64),
        ("top_nnz", ctypes.c_int64),
        ("obj", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("scale", ctypes.c_double),
        ("offset", ctypes.c_double),
        ("label_count", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def objective(self):
        return self.obj * self.scale + self.offset


# name -> (restype, argtypes); mirrors include/pggc.h
c = ctypes
P = c.c_void_p
I64P = c.POINTER(c.c_int64)
SIGNATURES = {
    "pggc_version": (c.c_int, []),
    "pggc_last_error": (c.c_char_p, []),
    "pggc_ctx_create": (c.c_int, [c.POINTER(c.c_void_p), P]),
    "pggc_ctx_destroy": (c.c_int, [P]),
    "pggc_ctx_set_seed": (c.c_int, [P, c.c_uint64, c.c_uint64]),
    "pggc_ctx_get_draws": (c.c_int, [P, c.POINTER(c.c_uint64)]),
    "pggc_ctx_set_verbose": (c.c_int, [P, c.c_int]),
    "pggc_ctx_memory": (c.c_int, [P, I64P, I64P]),
    "pggc_ctx_set_progress": (c.c_int, [P, P, P]),
    "pggc_ctx_set_sweep_trace": (c.c_int, [P, P, P]),
    "pggc_ctx_set_knob": (c.c_int, [P, c.c_int, c.c_int64]),
    "pggc_ctx_get_knob": (c.c_int, [P, c.c_int, P]),
    "pggc_autotune": (c.c_int, [P, c.POINTER(Graph), c.c_int, c.c_int, P, c.c_int, P]),
    "pggc_gnn_configure": (c.c_int, [P, c.c_int, c.c_double, c.c_int64, c.c_int, c.c_double, c.c_double]),
    "pggc_gnn_set_params": (c.c_int, [P, P, c.c_int]),
    "pggc_gnn_get_params": (c.c_int, [P, P, c.c_int]),
    "pggc_gnn_stats": (c.c_int, [P, P, P]),
    "pggc_gnn_reset_stats": (c.c_int, [P]),
    "pggc_gnn_eval": (c.c_int, [P, P, P, P, P, P, c.c_int32, c.c_int, P, P, P, P, P, P]),
    "pggc_ctx_last_times": (c.c_int, [P, c.POINTER(c.c_double), c.POINTER(c.c_double)]),
    "pggc_ctx_set_timing": (c.c_int, [P, c.c_int]),
    "pggc_ctx_kernel_stats": (c.c_int, [P, c.c_int, I64P, c.POINTER(c.c_double), c.POINTER(c.c_double)]),
    "pggc_vertex_weights": (c.c_int, [P, c.POINTER(Graph), c.c_int, P]),
    "pggc_cut_signature": (c.c_int, [P, c.POINTER(Graph), P, P]),
    "pggc_signature_diff": (c.c_int, [P, P, P, c.c_int64, I64P]),
    "pggc_intersect": (c.c_int, [P, c.c_int64, P, P, c.c_int64, P, I64P]),
    # pggc_io.h (host only)
    "pggc_read_graph": (c.c_int, [c.c_char_p, c.c_int, c.POINTER(c.c_void_p), c.POINTER(c.c_int)]),
    "pggc_host_graph_info": (c.c_int, [P, I64P, I64P]),
    "pggc_host_graph_copy": (c.c_int, [P, P, P, P]),
    "pggc_host_graph_free": (c.c_int, [P]),
    "pggc_write_part": (c.c_int, [c.c_char_p, P, c.c_int64]),
    "pggc_read_ints": (c.c_int, [c.c_char_p, c.c_int64, P]),
    "pggc_state_init": (c.c_int, [P, c.POINTER(Graph), c.c_int, c.c_double, c.POINTER(State)]),
    "pggc_orderings": (c.c_int, [P, c.POINTER(Graph), P, P, I64P]),
    "pggc_evaluate": (c.c_int, [P, c.POINTER(Graph), P, c.c_int64, P, I64P, c.POINTER(c.c_uint64)]),
    "pggc_evaluate_noalloc": (c.c_int, [P, c.POINTER(Graph), P, c.c_int64, P, I64P, c.POINTER(c.c_uint64)]),
    "pggc_evaluate_noalloc_limits": (c.c_int, [I64P, I64P]),
    "pggc_own_weight": (c.c_int, [P, c.POINTER(Graph), P, P]),
    "pggc_contract": (c.c_int, [P, c.POINTER(Graph), P, c.c_int64, P, P, P, c.c_int64, I64P]),
    "pggc_sweep": (c.c_int, [P, c.POINTER(Graph), P, P, c.c_int64, c.c_double, c.c_float, c.c_int, P, P, P, P,
                             I64P, I64P]),
    "pggc_local_move": (c.c_int, [P, c.POINTER(Graph), P, c.POINTER(State), c.c_int, c.c_int, c.c_int]),
    "pggc_leiden_refine": (c.c_int, [P, c.POINTER(Graph), P, P, c.c_int64, c.c_double, c.c_int, c.c_uint64,
                                     c.c_uint64, P, I64P]),
    "pggc_louvain": (c.c_int, [P, c.POINTER(Graph), c.POINTER(State), c.c_int, P, P]),
    "pggc_leiden": (c.c_int, [P, c.POINTER(Graph), c.POINTER(State), c.c_int, c.c_int, P, P]),
    "pggc_cluster": (c.c_int, [P, c.POINTER(Graph), c.c_int, c.c_int, c.c_double, c.c_int, c.c_uint64, P,
                               c.POINTER(c.c_double), I64P]),
    "pggc_cluster_batch": (c.c_int, [P, c.c_int, P, c.c_int, c.c_int, c.c_double, c.c_int, P, P, P, P, c.c_int]),
    "pggc_graph_from_edges": (c.c_int, [P, c.c_int64, c.c_int64, P, P, P, c.POINTER(c.c_void_p)]),
    "pggc_rmat": (c.c_int, [P, c.c_int, c.c_int, c.c_double, c.c_double, c.c_double, c.c_uint64, c.c_int32,
                            c.c_int32, c.POINTER(c.c_void_p)]),
    "pggc_lfr": (c.c_int, [P, c.c_int64, c.c_double, c.c_double, c.c_int32, c.c_double, c.c_int32, c.c_int32,
                           c.c_double, c.c_int32, c.c_int32, c.c_uint64, P, c.POINTER(c.c_void_p)]),
    "pggc_devgraph_view": (c.c_int, [P, c.POINTER(Graph)]),
    "pggc_devgraph_copy": (c.c_int, [P, P, P, P, P]),
    "pggc_devgraph_free": (c.c_int, [P]),
    "pggc_random_labels": (c.c_int, [P, c.c_int64, c.c_int64, c.c_uint64, P]),
}

_lib = None


def load():
    """Load libpggc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build the HIP library first (__graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != OK:
        msg = load().pggc_last_error()
        raise PggcError(rc, msg.decode() if msg else "")
    return rc
