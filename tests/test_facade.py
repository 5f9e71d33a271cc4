"""The C++ clustering_methods facade (include/pggc_clustering_methods.hpp,
the drop-in for interface/clustering_methods.h:20-24).

CPU: the header compiles on its own with the host compiler, with every
template instantiated.  GPU: bin/facade_check, a C++ caller that holds its
graph as plain device pointers (unit, int32 and int64 weights), runs
louvain_part<false> and leiden_part<true,false> through the facade; the
objectives (exact) and partitions equal the Python mirror's calls of the
same entry points on the same generated graph."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TOOL = os.path.join(ROOT, "pennsylvaniagpugraphclustering_amd", "bin", "facade_check")


def test_facade_header_compiles(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "pggc_clustering_methods.hpp"\n'
                   "namespace cm = jet_community::clustering_methods;\n"
                   "void use(pggc_ctx* c, cm::Graph& g, pggc_state& s, int32_t* p, int64_t* a, int64_t* b) {\n"
                   "    s = cm::make_state(c, g, PGGC_OBJ_MOD, 1.0, a, b, g.n);\n"
                   "    cm::louvain_part<false>(c, g, s, nullptr, p);\n"
                   "    cm::louvain_part<true>(c, g, s, p, p);\n"
                   "    cm::leiden_part<true, false>(c, g, s, nullptr, p);\n"
                   "    cm::leiden_part<false, true>(c, g, s, p, p);\n"
                   "}\n")
    subprocess.check_call([cxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                           str(src)])


def _fnv(a):
    h = 1469598103934665603
    for byte in np.ascontiguousarray(a.astype(np.int32)).tobytes():
        h ^= byte
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.gpu
@pytest.mark.parametrize("weights", ["none", "int32", "int64"])
def test_facade_matches_python_mirror(gpu_ctx, weights):
    import ctypes
    import torch
    from pennsylvaniagpugraphclustering_amd import clustering as C, graphs
    if not os.path.exists(TOOL):
        pytest.fail("bin/facade_check is not built")
    scale, ef, seed = 12, 8, 3
    print(f"facade_check {weights}", flush=True)
    out = subprocess.run([TOOL, str(scale), str(ef), str(seed), weights], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    rec = {r["call"]: r for r in map(json.loads, out.stdout.strip().splitlines())}
    rp, col, w = graphs.rmat_device(scale, ef, seed=seed, weights=None if weights == "none" else (1, 100))
    if weights == "int64":
        w = w.to(torch.int64)
    obj = "Mod" if weights == "none" else "WMod"
    dg = C.DeviceGraph(rp, col, w, None, objective=obj)
    st, _vol = C.state_init(dg, obj)
    part = C.louvain_part(dg, st)
    r = rec["louvain_part"]
    assert float.fromhex(r["objective"]) == st.obj * st.scale + st.offset
    assert r["labels"] == st.label_count
    assert r["fnv"] == _fnv(part.cpu().numpy())
    gpu_ctx.lib.pggc_ctx_set_seed(gpu_ctx.h, ctypes.c_uint64(7), ctypes.c_uint64(0))
    st, _vol = C.state_init(dg, obj)
    part = C.leiden_part(dg, st, plus=True)
    r = rec["leiden_part_plus"]
    assert float.fromhex(r["objective"]) == st.obj * st.scale + st.offset
    assert r["labels"] == st.label_count
    assert r["fnv"] == _fnv(part.cpu().numpy())
