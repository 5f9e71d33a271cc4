"""The execution-shape knobs (threads per vertex and slots per edge of the
LDS-table scoring, the LDS/HBM cut-off for hubs, hub build and scan item
sizes, the hub-table update divisor, the hub-table build method, the
top-level renumbering, the contraction's batch size) change time only: every value gives
the same clustering, bit for bit, as the defaults.  The autotuner times each
candidate and keeps the fastest, so it cannot change results either."""
import numpy as np
import pytest

from tests.graphcases import hub_graph
from pennsylvaniagpugraphclustering_amd import graphs

pytestmark = pytest.mark.gpu

DEFAULTS = {"table_threads": 64, "table_slots_per_edge": 2, "hub_lds_max_degree": 2048, "hub_edges_per_item": 1024,
            "hub_slots_per_item": 4096, "hub_update_div": 32, "hub_build": 0, "renumber": 2,
            "contract_batch": 1 << 30}


@pytest.fixture(scope="module")
def graphs_under_test():
    from pennsylvaniagpugraphclustering_amd import clustering as C
    out = {}
    for name, (g, obj) in {"hub30k": (hub_graph(), "Mod"),
                           "hub30k_w": (hub_graph(seed=9, weights=(1, 10)), "WMod"),
                           "rmat16": (graphs.rmat(16, 16, seed=2), "Mod")}.items():
        out[name] = (C.DeviceGraph.from_csr(g, obj), obj)
    return out


def _run(dg, obj, algo):
    from pennsylvaniagpugraphclustering_amd import clustering as C
    q, part, info = C.cluster(dg, algo, obj, 1.0, -1 if algo == "louvain" else 1, seed=5)
    return q, part.cpu().numpy(), info["sweeps"]


def _set_all(ctx, knobs):
    for k, v in knobs.items():
        ctx.set_knob(k, v)


def test_defaults(gpu_ctx):
    for k, v in DEFAULTS.items():
        assert gpu_ctx.get_knob(k) == v


@pytest.mark.parametrize("algo", ["louvain", "leiden+"])
def test_every_knob_value_gives_identical_results(gpu_ctx, graphs_under_test, algo):
    try:
        for name, (dg, obj) in graphs_under_test.items():
            ref = _run(dg, obj, algo)
            variants = [{k: v} for k, vals in gpu_ctx.KNOB_VALUES.items() for v in vals if v != DEFAULTS[k]]
            variants.append({"table_threads": 256, "table_slots_per_edge": 4, "hub_lds_max_degree": 4096,
                             "hub_edges_per_item": 512, "hub_slots_per_item": 2048, "hub_update_div": 0})
            # hub tables built by edge items with HBM atomics, for every hub
            # and for every other hub (the rest from LDS slices)
            variants += [{"hub_build": 1}, {"hub_build": 2}, {"hub_build": 2, "hub_update_div": 0}]
            variants += [{"renumber": 0}, {"renumber": 1}, {"contract_batch": 4096}]
            for var in variants:
                print(f"{name} {algo} knobs {var}", flush=True)
                _set_all(gpu_ctx, DEFAULTS)
                _set_all(gpu_ctx, var)
                got = _run(dg, obj, algo)
                assert got[0] == ref[0], (name, var)
                assert np.array_equal(got[1], ref[1]), (name, var)
                assert got[2] == ref[2], (name, var)
    finally:
        _set_all(gpu_ctx, DEFAULTS)


def test_knob_values_are_checked(gpu_ctx):
    from pennsylvaniagpugraphclustering_amd import _lib
    with pytest.raises(_lib.PggcError):
        gpu_ctx.set_knob("table_threads", 96)
    with pytest.raises(_lib.PggcError):
        gpu_ctx.set_knob("hub_lds_max_degree", 3000)
    assert gpu_ctx.get_knob("table_threads") == 64


def test_autotune_keeps_results(gpu_ctx, graphs_under_test):
    dg, obj = graphs_under_test["rmat16"]
    ref = _run(dg, obj, "louvain")
    try:
        chosen, table = gpu_ctx.autotune(dg, obj, reps=1)
        print(f"autotune rmat16: {chosen}", flush=True)
        assert len(table) == sum(len(v) for v in gpu_ctx.KNOB_VALUES.values())
        assert all(us > 0 for _, _, us in table)
        for k, v in chosen.items():
            assert v in gpu_ctx.KNOB_VALUES[k]
            # the chosen value is the fastest of its knob's rows
            rows = [r for r in table if r[0] == k]
            assert min(rows, key=lambda r: r[2])[1] == v
        got = _run(dg, obj, "louvain")
        assert got[0] == ref[0] and np.array_equal(got[1], ref[1])
    finally:
        _set_all(gpu_ctx, DEFAULTS)
