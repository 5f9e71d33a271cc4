"""The command-line tools end to end on the GPU (app/cluster_main.cpp,
app/verify_objective.cpp): graph file in, the reference's stdout lines,
partition file and -metrics JSON out, checked against the oracle run on the
same graph (trial t uses seed t)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pennsylvaniagpugraphclustering_amd import graphs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pennsylvaniagpugraphclustering_amd", "bin")


def run(tool, *args):
    print(f"run {tool} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BIN, tool), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def best_objective(out):
    return float(re.search(r"Best objective found: (\S+)", out).group(1))


def read_part(path):
    return np.array([int(x) for x in open(path).read().split()], dtype=np.int32)


@pytest.mark.parametrize("tool,algo", [("plouvain", "louvain"), ("pleiden", "leiden"), ("pleiden+", "leiden+")])
def test_cluster_tool_matches_oracle(tmp_path, tool, algo):
    g = graphs.sbm(3000, 30, 12, 4, seed=2, permute_seed=2)
    gp = tmp_path / "sbm.graph"
    graphs.write_metis(g, gp)
    out_p = tmp_path / "part.txt"
    out = run(tool, "-i", str(gp), "-o", str(out_p), "-ex_iters", "2")
    q, part, info = O.cluster(g.rp, g.col, None, g.vertex_weights("Mod"), "Mod", 1.0, algo, 2, seed=0)
    assert np.array_equal(read_part(out_p), part)
    assert f"{best_objective(out):.9g}" == f"{q:.9g}"
    lines = re.findall(r"Cluster time: \S+; Cut: (\d+); Modularity: (\S+); Labels: (\d+)", out)
    assert len(lines) == info["outer"]
    assert int(lines[-1][2]) == info["labels"]
    assert f"Vertex Count: {g.n}; Undirected Edge Count: {g.nnz // 2}" in out


def test_trials_metrics_and_verify(tmp_path):
    """Three trials (seeds 0, 1, 2) of pleiden+ on a weighted Matrix Market
    graph with WMod: best trial written, metrics JSON well formed, and
    verify_obj reproduces the best objective from the written file."""
    g = graphs.random_graph(2500, 10, seed=9, weights=(1, 20))
    gp = tmp_path / "w.mtx"
    graphs.write_mtx(g, gp)
    out_p, met_p = tmp_path / "part.txt", tmp_path / "metrics.json"
    out = run("pleiden+", "-i", str(gp), "-objective", "WMod", "-trials", "3", "-o", str(out_p), "-metrics",
              str(met_p))
    vw = g.vertex_weights("WMod")
    runs = [O.cluster(g.rp, g.col, g.w, vw, "WMod", 1.0, "leiden+", 0, seed=s) for s in range(3)]
    best = max(range(3), key=lambda s: runs[s][0])
    assert np.array_equal(read_part(out_p), runs[best][1])
    assert f"{best_objective(out):.9g}" == f"{runs[best][0]:.9g}"
    m = json.load(open(met_p))
    assert len(m) == 3
    for rec, (q, part, info) in zip(m, runs):
        assert abs(rec["modularity"] - q) < 1e-6
        assert rec["edge-cut"] == (g.nnz - info["uncut"]) // 2
        assert rec["total-duration-seconds"] > 0
    v = run("verify_obj", "-i", str(gp), "-objective", "WMod", "-clusters", str(out_p))
    assert f"Modularity: {runs[best][0]:.9g}" in v


@pytest.mark.parametrize("objective", ["CPM", "NLCC", "LCC"])
def test_lambdacc_tools(tmp_path, objective):
    g = graphs.random_graph(1500, 8, seed=13, weights=(1, 5))
    gp = tmp_path / "g.graph"
    graphs.write_metis(g, gp)
    out_p = tmp_path / "part.txt"
    out = run("plouvain", "-i", str(gp), "-objective", objective, "-lambda", "0.5", "-o", str(out_p))
    w = None if objective == "CPM" else g.w
    q, part, _ = O.cluster(g.rp, g.col, w, g.vertex_weights(objective), objective, 0.5, "louvain", 0, seed=0)
    assert np.array_equal(read_part(out_p), part)
    label = {"CPM": "Constant-Potts", "NLCC": "Normalized LambdaCC", "LCC": "Objective"}[objective]
    # the cut is (nnz - weighted uncut) / 2 as in interface/cluster_data.hpp:98-99, negative with weights
    assert re.search(rf"Cut: -?\d+; {label}: ", out)
    v = run("verify_obj", "-i", str(gp), "-objective", objective, "-lambda", "0.5", "-clusters", str(out_p))
    assert f"{label}: {q:.9g}" in v


def test_custom_vertex_weights(tmp_path):
    """-vtx_weights with NLCC (Objectives.md, app/io/vertex_weighting.cpp:27-29)."""
    g = graphs.random_graph(1200, 8, seed=17)
    vw = np.random.default_rng(3).integers(1, 6, g.n).astype(np.int64)
    gp, wp, out_p = tmp_path / "g.graph", tmp_path / "vw.txt", tmp_path / "part.txt"
    graphs.write_metis(g, gp)
    wp.write_text("\n".join(str(x) for x in vw) + "\n")
    out = run("plouvain", "-i", str(gp), "-objective", "NLCC", "-vtx_weights", str(wp), "-o", str(out_p))
    q, part, _ = O.cluster(g.rp, g.col, None, vw, "NLCC", 1.0, "louvain", 0, seed=0)
    assert np.array_equal(read_part(out_p), part)
    assert f"{best_objective(out):.9g}" == f"{q:.9g}"


def test_meme_primitives_match_numpy(gpu_ctx):
    """gen_signature, get_cut_diff and intersection_cluster
    (app/memetic_clustering.cpp:27-116) against numpy on the same inputs."""
    import ctypes
    import torch
    from pennsylvaniagpugraphclustering_amd import _lib
    from pennsylvaniagpugraphclustering_amd import clustering as C
    g = graphs.random_graph(3000, 9, seed=21)
    dg = C.DeviceGraph.from_csr(g, "Mod")
    rng = np.random.default_rng(4)
    a = rng.integers(0, 50, g.n).astype(np.int32)
    b = rng.integers(0, 7, g.n).astype(np.int32)
    lib = _lib.load()
    words = (g.nnz + 31) // 32
    ta, tb = torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda")
    sa = torch.zeros(words, dtype=torch.int32, device="cuda")
    sb = torch.zeros(words, dtype=torch.int32, device="cuda")
    _lib.check(lib.pggc_cut_signature(gpu_ctx.h, ctypes.byref(dg.c()), ta.data_ptr(), sa.data_ptr()))
    _lib.check(lib.pggc_cut_signature(gpu_ctx.h, ctypes.byref(dg.c()), tb.data_ptr(), sb.data_ptr()))
    rows = np.repeat(np.arange(g.n), np.diff(g.rp))
    cut_a = a[rows] != a[g.col]
    cut_b = b[rows] != b[g.col]
    bits = np.zeros(words * 32, bool)
    bits[: g.nnz] = cut_a
    ref = np.packbits(bits.reshape(-1, 32)[:, ::-1], axis=1).view(">u4").ravel().astype(np.uint32)
    assert np.array_equal(sa.cpu().numpy().view(np.uint32), ref)
    diff = ctypes.c_int64()
    _lib.check(lib.pggc_signature_diff(gpu_ctx.h, sa.data_ptr(), sb.data_ptr(), words, ctypes.byref(diff)))
    assert diff.value == int((cut_a != cut_b).sum())
    out = torch.empty(g.n, dtype=torch.int32, device="cuda")
    L = ctypes.c_int64()
    _lib.check(lib.pggc_intersect(gpu_ctx.h, g.n, ta.data_ptr(), tb.data_ptr(), 7, out.data_ptr(), ctypes.byref(L)))
    pairs = a.astype(np.int64) * 7 + b
    uniq, inv = np.unique(pairs, return_inverse=True)
    assert L.value == len(uniq)
    assert np.array_equal(out.cpu().numpy(), inv.astype(np.int32))


def test_meme_tool(tmp_path):
    """Ten pLeiden+ members evolved for the 10 s minimum: the final objective
    is the best member, at least the best initial member, and equals the
    oracle's evaluation of the written clustering."""
    g = graphs.sbm(2000, 25, 10, 4, seed=6, permute_seed=6)
    gp, out_p = tmp_path / "g.graph", tmp_path / "best.txt"
    graphs.write_metis(g, gp)
    out = run("meme", "-i", str(gp), "-pop_size", "4", "-time_limit", "1", "-o", str(out_p))
    assert "Setting population size to 10" in out and "Setting time limit to 10s" in out
    init = [float(x) for x in re.findall(r"Adding clustering with obj: (\S+)", out)]
    assert len(init) == 10
    assert re.search(r"Parent 1 obj: \S+; Parent 2 obj: \S+; Offspring obj: \S+; Post leiden obj: \S+; Min diff: \d+",
                     out)
    final = float(re.search(r"Final objective: (\S+)", out).group(1))
    assert final >= max(init) - 1e-9
    part = read_part(out_p)
    vol, uncut, sq = O.evaluate(g.rp, g.col, None, g.vertex_weights("Mod"), part, int(part.max()) + 1)
    q = (uncut - sq / g.nnz) / g.nnz
    assert abs(q - final) < 1e-8


@pytest.mark.parametrize("spec,weights,objective", [("rmat:14:16:2", None, "Mod"),
                                                    ("rmat:13:16:3:1:10", (1, 10), "WMod")])
def test_generated_graph_spec(tmp_path, spec, weights, objective):
    """-i rmat:SCALE:EDGE_FACTOR:SEED[:WMIN:WMAX] builds the graph on the GPU
    (the library's R-MAT generator, bit-identical to the oracle's): the tool's
    partition equals the oracle's run on the oracle-generated graph."""
    f = [int(x) for x in spec.split(":")[1:4]]
    rp, col, w = O.rmat(f[0], f[1], seed=f[2], weights=weights)
    out_p = tmp_path / "part.txt"
    out = run("pleiden+", "-i", spec, "-objective", objective, "-o", str(out_p), "-ex_iters", "1")
    vw = np.diff(rp).astype(np.int64) if objective == "Mod" else np.bincount(
        np.repeat(np.arange(len(rp) - 1), np.diff(rp)), weights=w.astype(np.float64), minlength=len(rp) - 1
    ).astype(np.int64)
    q, part, info = O.cluster(rp, col, None if objective == "Mod" else w, vw, objective, 1.0, "leiden+", 1, seed=0)
    assert f"Vertex Count: {len(rp) - 1}; Undirected Edge Count: {int(rp[-1]) // 2}" in out
    assert np.array_equal(read_part(out_p), part)
    assert f"{best_objective(out):.9g}" == f"{q:.9g}"
