"""CPU tests that pin three oracle functions no library routine covers:

- the afterburner's keep rule (src/local_mover.cpp:280-334): each candidate's
  gain plus its change equals, to FP32 rounding, half the exact change of
  uncut - lambda * sum(vol^2) when the candidate moves after exactly its
  priority neighbours have moved (computed with the integer evaluator), and
  hand-worked swaps keep exactly the higher-priority vertex;
- local_move's schedule and best-state tracking (src/local_mover.cpp:1382-1413),
  read from the oracle's per-sweep records;
- leidenR's gamma breaks (src/leidenR.cpp:140-186), recomputed by brute force
  from vertex sets on the oracle's own forest.

test_planted_mistakes_are_caught builds oracle variants that each carry one
planted one-line mistake and checks that the matching test here fails on it.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pennsylvaniagpugraphclustering_amd import graphs

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EPS = np.float32(0.1)
NEW_PART = -4
NO_MOVE = -3


def xorshift32(x):
    """Marsaglia's 13/17/5 xorshift on 32 bits (pinned by test_oracle.py)."""
    x &= 0xFFFFFFFF
    x ^= (x << 13) & 0xFFFFFFFF
    x ^= x >> 17
    x ^= (x << 5) & 0xFFFFFFFF
    return x


def signed32(x):
    return x - (1 << 32) if x >= (1 << 31) else x


def has_priority(gv, gi, v, i):
    """Neighbour v goes before candidate i (src/local_mover.cpp:319): FP32
    gain lead of at least 0.1, or a tie band with the lower signed hash."""
    d = np.float32(gv) - np.float32(gi)
    if d >= EPS:
        return True
    return abs(d) < EPS and signed32(xorshift32(v)) < signed32(xorshift32(i))


def exact_obj2(rp, col, w, vw, lab, lam):
    """uncut - lam * sum(vol^2) from the integer evaluator (float64 result)."""
    L = int(lab.max()) + 1
    _, u, s = O.evaluate(rp, col, w, vw, lab.astype(np.int32), L)
    return u - lam * s


# ---------------------------------------------------------------- afterburner
def afterburner_cases():
    out = []
    for seed, (n, deg, wts) in enumerate([(160, 6, None), (220, 9, (1, 7)), (300, 5, None), (120, 12, (1, 3))]):
        g = graphs.random_graph(n, deg, seed=50 + seed, weights=wts)
        obj = "WMod" if wts else "Mod"
        vw = g.vertex_weights(obj).astype(np.int64)
        W = int(vw.sum())
        rng = np.random.default_rng(seed)
        for L, ratio in [(n // 6, 0.0), (n // 3, 0.75), (n, 0.25)]:
            part = rng.integers(0, L, g.n).astype(np.int32) if L < n else np.arange(n, dtype=np.int32)
            out.append((g, vw, 1.0 / W, part, ratio))
    return out


def test_afterburner_change_is_exact_delta():
    """For every candidate i with best cluster b and current cluster p:
    gain_i + change_i = 1/2 [obj(i -> b after S_i moved) - obj(S_i moved)],
    where S_i is the set of i's neighbours with priority over i, each moved
    to its own destination (a fresh singleton for NEW_PART).  The keep
    decision is gain + change >= 0."""
    checked = decided = 0
    for g, vw, lam, part, ratio in afterburner_cases():
        w = g.w
        L = int(part.max()) + 1
        vol, _, _ = O.evaluate(g.rp, g.col, w, vw, part, L)
        dest0 = np.full(g.n, -1, np.int32)
        gain0 = np.zeros(g.n, np.float32)
        moves, dest, gain, cands, change = O.sweep_detail(g.rp, g.col, w, vw, part, vol, lam, ratio, False,
                                                          dest0, gain0)
        kept = set(moves.tolist())
        fresh = L + 1
        for x, i in enumerate(cands.tolist()):
            lab = part.astype(np.int64).copy()
            nb = g.col[g.rp[i]:g.rp[i + 1]]
            for v in nb.tolist():
                if has_priority(gain[v], gain[i], v, i):
                    if dest[v] == NEW_PART:
                        lab[v] = fresh
                        fresh += 1
                    else:
                        lab[v] = dest[v]
            base = exact_obj2(g.rp, g.col, w, vw, lab, lam)
            lab2 = lab.copy()
            lab2[i] = fresh if dest[i] == NEW_PART else dest[i]
            fresh += 1
            exact = (exact_obj2(g.rp, g.col, w, vw, lab2, lam) - base) / 2.0
            got = float(np.float32(gain[i]) + np.float32(change[x]))
            scale = 1.0 + float(np.abs(w[g.rp[i]:g.rp[i + 1]]).sum() if w is not None else g.rp[i + 1] - g.rp[i])
            assert abs(got - exact) <= 2e-5 * scale, (i, got, exact)
            checked += 1
            if abs(exact) > 2e-5 * scale:
                assert (i in kept) == (exact >= 0), (i, exact)
                decided += 1
    assert checked > 300 and decided > 300


def hand_graph(n, edges, vw):
    src = [a for a, b, _ in edges]
    dst = [b for a, b, _ in edges]
    wt = [c for _, _, c in edges]
    g = graphs.from_edges(n, np.array(src), np.array(dst), np.array(wt))
    return g, np.asarray(vw, np.int64)


def one_sweep(g, vw, part, lam):
    L = int(part.max()) + 1
    vol = np.zeros(L, np.int64)
    np.add.at(vol, part, vw)
    moves, dest, gain, _ = O.sweep(g.rp, g.col, g.w, vw, part.astype(np.int32), vol, lam, 0.0, False,
                                   np.full(g.n, -1, np.int32), np.zeros(g.n, np.float32))
    return moves, dest, gain


def first_id(pred, start=2):
    v = start
    while not pred(xorshift32(v)):
        v += 1
    return v


@pytest.mark.parametrize("signs", ["mixed", "positive"])
def test_afterburner_hand_swap_keeps_the_lower_signed_hash(signs):
    """Two singletons x, y joined by one unit edge, vertex weight = degree,
    lambda = 1/nnz = 1/2.  Both want to join the other: gain = 1 - 1/2 =
    1/2 each.  The one with priority (tie band, lower signed xorshift32)
    sees the other's move: change = -2 * (1 - 1/2) = -1, 1/2 - 1 < 0, so it
    is dropped and exactly the higher-priority vertex is kept."""
    neg = first_id(lambda h: h >= 1 << 31)      # negative as int32
    pos = first_id(lambda h: h < 1 << 31, neg + 1)
    if signs == "positive":
        neg = first_id(lambda h: h < 1 << 31, pos + 1)
    x, y = neg, pos
    n = max(x, y) + 1
    g, _ = hand_graph(n, [(x, y, 1)], np.zeros(n))
    g.w = None
    vw = g.vertex_weights("Mod").astype(np.int64)
    moves, dest, gain = one_sweep(g, vw, np.arange(n), 0.5)
    assert gain[x] == np.float32(0.5) and gain[y] == np.float32(0.5)
    assert dest[x] == y and dest[y] == x
    winner = x if signed32(xorshift32(x)) < signed32(xorshift32(y)) else y
    if signs == "mixed":
        assert winner == x  # the int32 comparison, not the unsigned one
    assert moves.tolist() == [winner]


def test_afterburner_hand_gain_band():
    """x, z share cluster P (edge x-z weight 1); y is alone in Q (edge x-y
    weight 3); unit vertex weights, lambda = 0.1.  x: p_conn = 1 - 0.1 = 0.9,
    j_conn(Q) = 3 - 0.1 = 2.9, gain 2.0.  y: p_conn = 0, j_conn(P) = 3 - 0.2
    = 2.8, gain 2.8.  y leads by 0.8 >= 0.1, so x sees y's move into P:
    change = -2.9 (y joins p) - 2.9 (y leaves b) = -5.8 and x is dropped;
    y sees no priority neighbour and is kept, whatever the hashes."""
    x, y, z = 0, 1, 2
    g, vw = hand_graph(3, [(x, y, 3), (x, z, 1)], [1, 1, 1])
    part = np.array([0, 1, 0])
    moves, dest, gain = one_sweep(g, vw, part, 0.1)
    assert abs(gain[x] - 2.0) < 1e-6 and abs(gain[y] - 2.8) < 1e-6
    assert dest[x] == 1 and dest[y] == 0 and dest[z] == NO_MOVE
    assert moves.tolist() == [y]


# ---------------------------------------------------------------- local_move
def lm_cases():
    out = []
    for seed in range(6):
        g = graphs.random_graph(400, 4 + seed, seed=70 + seed, weights=(1, 5) if seed % 2 else None)
        out.append(g)
    out.append(graphs.sbm(600, 12, 8, 3, seed=3))
    out.append(graphs.rmat(10, 8, seed=5))
    return out


def check_schedule(tr, SA):
    ratios = [float(r) for r in tr[:, 0]]
    hi, lo = (0.75, 0.25) if SA else (0.0, 0.0)
    pos = 0
    for ratio, limit in ((hi, 4), (lo, 2)):
        # a phase runs `limit` sweeps or stops after a sweep that moved nothing
        start = pos
        while pos < len(ratios) and pos - start < limit:
            pos += 1
            if tr[pos - 1, 2] == 0:
                break
        assert pos > start, tr
        assert all(r == ratio for r in ratios[start:pos]), tr
    assert pos == len(ratios), tr
    # a sweep that moved nothing leaves the objective unchanged
    for k in range(len(tr)):
        if tr[k, 2] == 0:
            prev = tr[k - 1, 3] if k > 0 else None
            if prev is not None:
                assert tr[k, 3] == prev


def test_local_move_schedule_and_best_state():
    """The per-sweep ratios follow 0.75 x <=4 then 0.25 x <=2 (all 0 without
    annealing), each phase ending early only at a sweep that moved nothing;
    the returned state is the best objective on the trajectory (including
    the starting state), the returned labels evaluate to exactly that
    objective, and they are contiguous."""
    not_last = 0
    for g in lm_cases():
        w = g.w
        obj_name = "WMod" if w is not None else "Mod"
        vw = g.vertex_weights(obj_name).astype(np.int64)
        lam = 1.0 / int(vw.sum())
        for SA in (True, False):
            part = np.arange(g.n, dtype=np.int32)
            vol = vw.copy()
            obj0 = -lam * float(sum(int(v) * int(v) for v in vw))
            L, uncut, obj = g.n, 0, obj0
            for call in range(3):
                p2, v2, L2, u2, o2, tr = O.local_move_trace(g.rp, g.col, w, vw, part, vol, L, uncut, lam, obj,
                                                            call == 0, SA=SA)
                check_schedule(tr, SA)
                best = max([obj] + [float(t) for t in tr[:, 3]])
                assert o2 == best
                if tr.shape[0] and float(tr[-1, 3]) < best:
                    not_last += 1
                assert sorted(set(p2.tolist())) == list(range(L2))
                vv, uu, ss = O.evaluate(g.rp, g.col, w, vw, p2, L2)
                assert uu == u2 and np.array_equal(vv, v2)
                assert uu - lam * ss == pytest.approx(o2, rel=1e-12, abs=1e-12)
                part, vol, L, uncut, obj = p2, v2, L2, u2, o2
    # the trajectory's last state was worse than its best in some calls
    assert not_last >= 3, not_last


# ---------------------------------------------------------------- leidenR
def refine_cases():
    out = []
    for k, (gen, mult) in enumerate([
        (lambda: graphs.sbm(3000, 30, 9, 4, seed=4), 1.0),
        (lambda: graphs.sbm(3000, 30, 9, 4, seed=4), 16.0),
        (lambda: graphs.random_graph(2500, 6, seed=8, weights=(1, 30)), 1.0),
        (lambda: graphs.random_graph(2500, 6, seed=8, weights=(1, 30)), 16.0),
        (lambda: graphs.rmat(12, 12, seed=6), 1.0),
        (lambda: graphs.rmat(12, 12, seed=6), 4.0),
        (lambda: graphs.rmat(12, 16, seed=9), 1.0),
    ]):
        g = gen()
        weighted = g.w is not None
        obj = "WMod" if weighted else "Mod"
        vw = g.vertex_weights(obj).astype(np.int64)
        w = g.w if weighted else None
        q, louv, info = O.cluster(g.rp, g.col, w, vw, obj, 1.0, "louvain", 0)
        L = int(louv.max()) + 1
        vol, _, _ = O.evaluate(g.rp, g.col, w, vw, louv, L)
        out.append((g, w, vw, louv, vol, mult / int(vw.sum()), not weighted, k))
    return out


def test_leiden_gamma_breaks_by_brute_force():
    """On the oracle's own forest (trees before the breaks), each tree's
    members are taken in (order, id) order.  Member x breaks the tree when
    its weight to the members before it is below gamma * w_x * (their
    total vertex weight); the prefix ending at x breaks the tree after x
    when its cut weight to the rest of its Louvain cluster is below gamma *
    |prefix| * (vol(cluster) - |prefix|) (src/leidenR.cpp:140-178).  Members
    from the first break on become singletons (:180-186).  Recomputed here
    from vertex sets and the adjacency, the resulting partition must equal
    the oracle's."""
    breaks_inner = breaks_outer = trees_checked = 0
    for g, w, vw, louv, vol, gamma, uniform, k in refine_cases():
        vc, nc, trees, order = O.leiden_refine_detail(g.rp, g.col, w, vw, louv, vol, gamma, uniform, seed=11 + k)
        gam = float(np.float32(gamma))
        expect = trees.copy()
        members = {}
        for i in range(g.n):
            members.setdefault(int(trees[i]), []).append(i)
        adj = [dict() for _ in range(g.n)]
        for i in range(g.n):
            for j in range(g.rp[i], g.rp[i + 1]):
                adj[i][int(g.col[j])] = 1 if w is None else int(w[j])
        ambiguous = set()
        for root, mem in members.items():
            if len(mem) < 2:
                continue
            mem.sort(key=lambda v: (int(order[v]), v))
            if len({int(order[v]) for v in mem}) < len(mem):
                ambiguous.add(root)  # equal orders: the inner test counts strictly earlier orders only
                continue
            cl = int(louv[mem[0]])
            assert all(int(louv[v]) == cl for v in mem)
            cluster = set(np.nonzero(louv == cl)[0].tolist())
            brk = len(mem)
            prefix = set()
            prev = 0
            for x, i in enumerate(mem):
                inner = sum(wt for v, wt in adj[i].items() if v in prefix)
                rhs = gam * float(vw[i]) * prev
                if rhs != 0 and abs(inner - rhs) <= 1e-5 * max(1.0, rhs):
                    ambiguous.add(root)  # within FP32 rounding of the threshold
                if inner < rhs and x < brk:
                    brk = x
                    kind = "inner"
                prefix.add(i)
                prev += int(vw[i])
                outw = sum(wt for u in prefix for v, wt in adj[u].items() if v in cluster and v not in prefix)
                rhs2 = gam * prev * float(int(vol[cl]) - prev)
                if rhs2 != 0 and abs(outw - rhs2) <= 1e-5 * max(1.0, abs(rhs2)):
                    ambiguous.add(root)
                if outw < rhs2 and x + 1 < brk:
                    brk = x + 1
                    kind = "outer"
            if root in ambiguous:
                continue
            if brk < len(mem):
                if kind == "inner":
                    breaks_inner += 1
                else:
                    breaks_outer += 1
            for i in mem[brk:]:
                expect[i] = i
            trees_checked += 1
        # compare as partitions over the unambiguous trees
        for root, mem in members.items():
            if root in ambiguous:
                continue
            got = {int(vc[i]) for i in mem}
            want = {}
            for i in mem:
                want.setdefault(int(expect[i]), []).append(i)
            assert len(got) == len(want), (k, root)
            for grp in want.values():
                assert len({int(vc[i]) for i in grp}) == 1, (k, root, grp)
        assert len(ambiguous) <= 0.02 * max(1, len(members))
    assert trees_checked > 1000
    assert breaks_inner >= 3 and breaks_outer >= 5, (breaks_inner, breaks_outer)


# ---------------------------------------------------------------- planted mistakes
MUTANTS = [
    (1, "test_afterburner_hand_swap_keeps_the_lower_signed_hash"),
    (2, "test_afterburner_change_is_exact_delta"),
    (3, "test_local_move_schedule_and_best_state"),
    (4, "test_local_move_schedule_and_best_state"),
    (5, "test_leiden_gamma_breaks_by_brute_force"),
    (6, "test_leiden_gamma_breaks_by_brute_force"),
]


def test_planted_mistakes_are_caught(tmp_path):
    """Each oracle variant carries one planted one-line mistake (unsigned
    hash comparison; a part/dest term swapped; a wrong second filter ratio;
    last state kept instead of the best; the vertex weight dropped from the
    inner gamma test; the outer break one position early).  The test that
    pins the function must fail on it."""
    if os.environ.get("PGGC_ORACLE_LIB"):
        pytest.skip("running inside a planted-mistake check")
    libs = {}
    procs = []
    for k, _ in MUTANTS:
        out = str(tmp_path / f"liboracle_m{k}.so")
        libs[k] = out
        procs.append(subprocess.Popen(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mutant", f"MUTANT={k}",
                                       f"OUT={out}"]))
    for p in procs:
        assert p.wait() == 0
    for k, test in MUTANTS:
        env = dict(os.environ, PGGC_ORACLE_LIB=libs[k])
        r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                            f"{__file__}::{test}"], cwd=ROOT, env=env, capture_output=True, text=True)
        assert r.returncode != 0, f"planted mistake {k} not caught by {test}:\n{r.stdout[-2000:]}"
