// PGGC CPU oracle: a restatement of the reference's Louvain/Leiden hot path,
// used ONLY as test infrastructure (parity checker and CPU baseline).
// Nothing in the product library links or calls this file.
//
// Threads: or_set_threads(k) runs the per-vertex loops (scoring, afterburner,
// uncut delta, contraction rows, evaluator, leidenR neighbour choice) on k
// OpenMP threads.  Every parallel loop writes per-vertex results or reduces
// integers, and every order-dependent step (candidate lists, label
// assignment, FP32 folds) stays sequential, so results are bit-identical for
// every thread count.
//
// Every function cites the reference file:line whose behaviour it restates.
// Where the reference is nondeterministic (CAS-built hash tables, team
// reductions in FP32, random_device seeds, unordered coarse columns) this file
// fixes ONE deterministic choice; the same choice is implemented by the HIP
// library, and the choice is listed at each site ("canonical rule").
//
// Integer widening: offsets, volumes, uncut and coarse edge weights are int64
// here (the reference uses int32 everywhere, interface/core_types.h:9-14);
// values agree whenever the reference does not overflow.
//
// Floating point: compiled with -ffp-contract=off; the four expressions that
// the CUDA compiler contracts into an FMA in the reference (p_conn, j_conn,
// the afterburner term q and the weighted leidenR score) use fmaf explicitly.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <vector>

#include <omp.h>
#include <parallel/algorithm>

#ifndef ORC_MUTANT
#define ORC_MUTANT 0  // tests build variants with one planted mistake each
#endif

namespace orc {

using i32 = int32_t;
using i64 = int64_t;
using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;

// src/local_mover.cpp:74-80
constexpr float OBJ_MIN = -FLT_MAX;
constexpr i32 NULL_PART = -1;
constexpr i32 NO_MOVE = -3;
constexpr i32 NEW_PART = -4;
constexpr float AFTERBURNER_EPS = 0.1f;
// src/header/ordering.h:17-19
constexpr i64 MID_CUTOFF = 64;
constexpr i64 LARGE_CUTOFF = 128;
constexpr i64 MASSIVE_CUTOFF = 15000;
// src/leidenR.cpp:38-39
constexpr i32 ORD_MAX = std::numeric_limits<i32>::max();
constexpr i32 SPLIT = 1000000000;

// xorshift32, src/local_mover.cpp:82-87 (also src/contract.cpp:91-97)
static inline u32 xorshift32(u32 x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// Kokkos::pod_hash<int32> (src/leidenR.cpp:37) = MurmurHash3_x86_32 of the
// 4 key bytes with seed 0 (public-domain algorithm by A. Appleby).
static inline u32 rotl32(u32 x, int r) { return (x << r) | (x >> (32 - r)); }
static inline u32 murmur3_i32(i32 key) {
    u32 k1 = (u32)key;
    u32 h1 = 0;
    k1 *= 0xcc9e2d51u;
    k1 = rotl32(k1, 15);
    k1 *= 0x1b873593u;
    h1 ^= k1;
    h1 = rotl32(h1, 13);
    h1 = h1 * 5u + 0xe6546b64u;
    h1 ^= 4u;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    h1 ^= h1 >> 16;
    return h1;
}

// Seeded stand-in for std::random_device (src/leidenR.cpp:289,291,399,479):
// the k-th draw of a run with seed s is the low 32 bits of splitmix64(s + k).
static inline u32 seed_draw(u64 seed, u64 k) {
    u64 z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (u32)z;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  The synthetic graphs are
// drawn from it so that the CPU and the GPU build identical graphs.
struct P4 {
    u32 v[4];
};
static inline P4 philox4x32_10(P4 c, u32 k0, u32 k1) {
    for (int r = 0; r < 10; r++) {
        const u64 p0 = (u64)0xD2511F53u * c.v[0];
        const u64 p1 = (u64)0xCD9E8D57u * c.v[2];
        P4 o;
        o.v[0] = (u32)(p1 >> 32) ^ c.v[1] ^ k0;
        o.v[1] = (u32)p1;
        o.v[2] = (u32)(p0 >> 32) ^ c.v[3] ^ k1;
        o.v[3] = (u32)p0;
        c = o;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Streams of the generators (counter word 3).
constexpr u32 STREAM_RMAT = 0, STREAM_WEIGHT = 1, STREAM_LABEL = 2;
constexpr u32 STREAM_LFR_DEG = 3, STREAM_LFR_COMM = 4, STREAM_LFR_ISTUB = 5, STREAM_LFR_ESTUB = 6,
              STREAM_LFR_PERM = 7;

// R-MAT quadrant thresholds on a u32 draw: r < t[0] -> (0,0), < t[1] ->
// (0,1), < t[2] -> (1,0), else (1,1) (Chakrabarti, Zhan, Faloutsos, SDM'04).
static void rmat_thresholds(double a, double b, double c, u64 t[3]) {
    t[0] = (u64)std::floor(a * 4294967296.0);
    t[1] = (u64)std::floor((a + b) * 4294967296.0);
    t[2] = (u64)std::floor((a + b + c) * 4294967296.0);
}
// Edge draw e: one Philox block per 4 levels, counter (e lo, e hi, block,
// stream 0), key (seed lo, seed hi); level l (MSB first) uses word l % 4.
static inline void rmat_edge(u64 e, int scale, const u64 t[3], u64 seed, u32& src, u32& dst) {
    u32 s = 0, d = 0;
    P4 blk{};
    for (int l = 0; l < scale; l++) {
        if ((l & 3) == 0) blk = philox4x32_10(P4{{(u32)e, (u32)(e >> 32), (u32)(l >> 2), STREAM_RMAT}},
                                              (u32)seed, (u32)(seed >> 32));
        const u64 r = blk.v[l & 3];
        const u32 bs = r >= t[1];
        const u32 bd = (r >= t[0] && r < t[1]) || r >= t[2];
        s = (s << 1) | bs;
        d = (d << 1) | bd;
    }
    src = s;
    dst = d;
}
// weight of the undirected pair lo < hi: uniform in [wmin, wmax]
static inline i32 pair_weight(u32 lo, u32 hi, u64 seed, i32 wmin, i32 wmax) {
    const P4 r = philox4x32_10(P4{{lo, hi, 0u, STREAM_WEIGHT}}, (u32)seed, (u32)(seed >> 32));
    return wmin + (i32)(r.v[0] % (u32)(wmax - wmin + 1));
}

struct Graph {
    i64 n = 0;
    std::vector<i64> rp;
    std::vector<i32> col;
    std::vector<i64> w;   // empty when uniform
    std::vector<i64> vw;  // vertex weights
    bool uniform = true;
    i64 nnz() const { return rp[n]; }
    i64 deg(i64 i) const { return rp[i + 1] - rp[i]; }
    i64 wt(i64 j) const { return uniform ? 1 : w[j]; }
};

// cluster_data (interface/cluster_data.hpp:18-88) with the objective's
// normalisation (modularity :109-133, constant_potts :135-156,
// normalized_lcc :158-182, base cluster_data).
struct State {
    std::vector<i64> vol;  // total_deg
    i64 uncut = 0;
    i64 top_nnz = 0;
    double obj = -1.0;
    i32 label_count = 0;
    double lambda = 1.0;
    double scale = 1.0;   // get_objective() = obj*scale + offset
    double offset = 0.0;

    void update_objective() {  // interface/cluster_data.hpp:76-87
        u128 sq = 0;
        for (i32 l = 0; l < label_count; l++) {
            i64 v = vol[l];
            u64 a = (u64)(v < 0 ? -v : v);
            sq += (u128)a * (u128)a;
        }
        double m = (double)uncut;
        m -= lambda * (double)sq;
        obj = m;
    }
    double objective() const { return obj * scale + offset; }
    void copy_from(const State& r) {  // interface/cluster_data.hpp:56-65
        if (vol.size() < (size_t)r.label_count) vol.resize(r.vol.size());
        std::copy(r.vol.begin(), r.vol.begin() + r.label_count, vol.begin());
        top_nnz = r.top_nnz;
        uncut = r.uncut;
        obj = r.obj;
        label_count = r.label_count;
        lambda = r.lambda;
    }
};

// ---------------------------------------------------------------- A1 ordering
// src/ordering.cpp:131-164 (stable degree bucketing into order1 / order2)
struct Orders {
    std::vector<i32> o1, o2;
    i64 off_large = 0, off_large2 = 0, off_mid = 0, off_mid2 = 0;
};
static Orders make_orders(const Graph& g) {
    Orders o;
    i64 n = g.n;
    o.o1.resize(n);
    o.o2.resize(n);
    i64 b[5] = {0, 0, 0, 0, 0};
    for (i64 i = 0; i < n; i++) {
        i64 d = g.deg(i);
        if (d < LARGE_CUTOFF) b[0]++;
        else if (d < MASSIVE_CUTOFF) b[1]++;
        else b[4]++;
        if (d < MID_CUTOFF) b[2]++;
        else if (d < MASSIVE_CUTOFF) b[3]++;
    }
    i64 off[5] = {0, b[0], 0, b[2], b[0] + b[1]};
    o.off_large = off[1];
    o.off_mid = off[3];
    o.off_large2 = off[4];
    o.off_mid2 = off[4];
    i64 c[5] = {0, 0, 0, 0, 0};
    for (i64 i = 0; i < n; i++) {
        i64 d = g.deg(i);
        if (d < LARGE_CUTOFF) o.o1[off[0] + c[0]++] = (i32)i;
        else if (d < MASSIVE_CUTOFF) o.o1[off[1] + c[1]++] = (i32)i;
        else {
            o.o1[off[4] + c[4]] = (i32)i;
            o.o2[off[4] + c[4]] = (i32)i;
            c[4]++;
        }
        if (d < MID_CUTOFF) o.o2[off[2] + c[2]++] = (i32)i;
        else if (d < MASSIVE_CUTOFF) o.o2[off[3] + c[3]++] = (i32)i;
    }
    return o;
}

// --------------------------------------------------------------- A12 evaluator
// app/verify_objective.cpp:64-89 + interface/cluster_data.hpp:76-87
static void evaluate(const Graph& g, const i32* labels, i64 L, std::vector<i64>& vol, i64& uncut) {
    vol.assign(L, 0);
    i64 u = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : u)
    for (i64 i = 0; i < g.n; i++) {
        i32 l = labels[i];
        for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++)
            if (labels[g.col[j]] == l) u += g.wt(j);
    }
    uncut = u;
    for (i64 i = 0; i < g.n; i++) vol[labels[i]] += g.vw[i];
}

// ------------------------------------------------------------- A10 contraction
// src/contract.cpp:234-336: coarse vertex u collects sum of w(i,j) over fine
// edges with vcmap(i)=u, vcmap(j)=v != u (self loops dropped :191,209).
// Canonical rule: coarse columns sorted ascending (the reference emits them
// in hash-slot order, :302-328).
static Graph contract(const Graph& g, const std::vector<i32>& vcmap, i64 nc) {
    std::vector<i64> cnt(nc + 1, 0);
    for (i64 i = 0; i < g.n; i++) cnt[vcmap[i] + 1]++;
    for (i64 u = 0; u < nc; u++) cnt[u + 1] += cnt[u];
    std::vector<i32> members(g.n);
    {
        std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
        for (i64 i = 0; i < g.n; i++) members[pos[vcmap[i]]++] = (i32)i;
    }
    Graph c;
    c.n = nc;
    c.uniform = false;
    c.rp.assign(nc + 1, 0);
    // pass 0 counts the distinct coarse neighbours of every coarse row, pass 1
    // writes them sorted with their summed weights
    for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel
        {
            std::vector<i64> acc(nc, 0);
            std::vector<i32> touched;
            std::vector<char> seen(nc, 0);
#pragma omp for schedule(dynamic, 256)
            for (i64 u = 0; u < nc; u++) {
                touched.clear();
                for (i64 k = cnt[u]; k < cnt[u + 1]; k++) {
                    i32 i = members[k];
                    for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++) {
                        i32 v = vcmap[g.col[j]];
                        if (v == u) continue;
                        if (!seen[v]) { seen[v] = 1; touched.push_back(v); }
                        acc[v] += g.wt(j);
                    }
                }
                if (pass == 0) {
                    c.rp[u + 1] = (i64)touched.size();
                } else {
                    std::sort(touched.begin(), touched.end());
                    i64 o = c.rp[u];
                    for (i32 v : touched) {
                        c.col[o] = v;
                        c.w[o] = acc[v];
                        o++;
                    }
                }
                for (i32 v : touched) { acc[v] = 0; seen[v] = 0; }
            }
        }
        if (pass == 0) {
            for (i64 u = 0; u < nc; u++) c.rp[u + 1] += c.rp[u];
            c.col.resize(c.rp[nc]);
            c.w.resize(c.rp[nc]);
        }
    }
    return c;
}

// ---------------------------------------------------------- local-move state
struct Mover {
    const Graph& g;
    Orders ord;
    std::vector<i32> part, dest;
    std::vector<float> gain;
    bool tables = false;  // cdata.init (src/local_mover.cpp:97)
    Mover(const Graph& gg) : g(gg) { ord = make_orders(g); }
};

// select_destinations (src/local_mover.cpp:435-563) for one vertex.
// Canonical rule: among eligible clusters keep the max j_conn, ties -> the
// smallest label; NEW_PART (-4) competes with j_conn 0 when p_conn < 0, so a
// real cluster beats it only with j_conn > 0 (the team path's min-reduce,
// :553-561, applied with one table slot per thread).
struct Agg {  // dense accumulator over labels with a touched list
    std::vector<i64> acc;
    std::vector<i32> keys;
    std::vector<char> seen;
    void ensure(size_t n) {
        if (acc.size() < n) { acc.resize(n, 0); seen.resize(n, 0); }
    }
    void add(i32 k, i64 w) {
        if (!seen[k]) { seen[k] = 1; keys.push_back(k); }
        acc[k] += w;
    }
    void clear() {
        for (i32 k : keys) { acc[k] = 0; seen[k] = 0; }
        keys.clear();
    }
};

static void score_vertex(Mover& M, State& st, i64 i, float lam, float ratio, bool constrained,
                         const i32* constraint, Agg& agg) {
    const Graph& g = M.g;
    if (M.dest[i] == NO_MOVE) return;  // :477-479
    const bool raw = !M.tables;        // first sweep scores the raw adjacency :1391-1395
    i64 w = g.vw[i];
    float multi = (float)w * lam;
    i32 p = M.part[i];
    agg.clear();
    i64 pv = 0;
    for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++) {
        i32 v = g.col[j];
        i32 key = raw ? v : M.part[v];
        if (!raw && key == p) pv += g.wt(j);
        else agg.add(key, g.wt(j));
    }
    if (raw) pv = 0;  // pvals are zero before the first table build (:1349)
    float p_conn = fmaf(-(float)(st.vol[p] - w), multi, (float)pv);
    float b = (float)((double)p_conn * (1.0 - (double)ratio));
    if (p_conn < 0) b = 0;
    bool have = false;
    float bj = 0;
    i32 bc = 0;
    for (i32 c : agg.keys) {
        i64 s = agg.acc[c];
        if (!(s > 0 && (float)s >= b)) continue;
        if (constrained && constraint[c] != constraint[p]) continue;
        float jc = fmaf(-(float)st.vol[c], multi, (float)s);
        if (jc >= b) {
            if (!have || jc > bj || (jc == bj && c < bc)) { have = true; bj = jc; bc = c; }
        }
    }
    if (p_conn < 0 && (!have || !(bj > 0.0f))) {  // NEW_PART wins ties at 0
        M.gain[i] = 0.0f - p_conn;
        M.dest[i] = NEW_PART;
        return;
    }
    if (have) {
        M.gain[i] = bj - p_conn;
        M.dest[i] = bc;
    } else {
        M.gain[i] = OBJ_MIN;
        M.dest[i] = NO_MOVE;
    }
}

// candidates_and_destinations (src/local_mover.cpp:567-634): score every
// vertex, then list the ones with dest != NO_MOVE in order1 order.
static std::vector<i32> candidates(Mover& M, State& st, float ratio, bool constrained, const i32* constraint) {
    float lam = (float)st.lambda;
    const size_t cap = std::max<size_t>(st.vol.size(), (size_t)M.g.n) + 1;
#pragma omp parallel
    {
        static thread_local Agg agg;
        agg.ensure(cap);
#pragma omp for schedule(dynamic, 1024)
        for (i64 i = 0; i < M.g.n; i++) score_vertex(M, st, i, lam, ratio, constrained, constraint, agg);
    }
    std::vector<i32> out;
    for (i64 x = 0; x < M.g.n; x++) {
        i32 i = M.ord.o1[x];
        if (M.dest[i] != NO_MOVE) out.push_back(i);
    }
    return out;
}

// afterburner_kernel (src/local_mover.cpp:280-370). Canonical FP32 order of
// the change sum: degree < 128 -> one sequential fold in adjacency order
// (the reference's serial path); 128 <= degree < 15000 -> 64 strided partial
// folds joined by a halving tree; degree >= 15000 -> 1024 strided partial
// folds joined by a halving tree (the reference's team reductions, :347-363,
// whose join order CUDA leaves unspecified).
static inline void ab_term(const Mover& M, i64 j, i32 p, i32 best, float multi,
                           float igain, i32 hi, float& acc) {
    const Graph& g = M.g;
    i32 v = g.col[j];
    float vg = M.gain[v];
#if ORC_MUTANT == 1  // planted mistake: unsigned hash comparison
    const bool lower = xorshift32((u32)v) < (u32)hi;
#else
    const bool lower = (i32)xorshift32((u32)v) < hi;
#endif
    if ((vg - igain) >= AFTERBURNER_EPS || (std::fabs(vg - igain) < AFTERBURNER_EPS && lower)) {
        i32 vpart = M.dest[v];
        float q = fmaf(-multi, (float)g.vw[v], (float)g.wt(j));
#if ORC_MUTANT == 2  // planted mistake: dest and part terms of p swapped
        acc = acc + ((vpart == p) ? q : 0.0f);
#else
        acc = acc - ((vpart == p) ? q : 0.0f);
#endif
        acc = acc + ((vpart == best && best != NEW_PART) ? q : 0.0f);
        vpart = M.part[v];
#if ORC_MUTANT == 2
        acc = acc - ((vpart == p) ? q : 0.0f);
#else
        acc = acc + ((vpart == p) ? q : 0.0f);
#endif
        acc = acc - ((vpart == best) ? q : 0.0f);
    }
}

static float tree_fold(const Mover& M, i64 i, i32 p, i32 best, float multi, float igain,
                       i32 hi, int T) {
    std::vector<float> part(T, 0.0f);
    i64 s = M.g.rp[i], e = M.g.rp[i + 1];
    for (int t = 0; t < T; t++)
        for (i64 j = s + t; j < e; j += T) ab_term(M, j, p, best, multi, igain, hi, part[t]);
    for (int off = T / 2; off >= 1; off /= 2)
        for (int t = 0; t < off; t++) part[t] = part[t] + part[t + off];
    return part[0];
}

static float ab_change(const Mover& M, i64 i, float lam) {
    const Graph& g = M.g;
    i32 best = M.dest[i];
    i32 p = M.part[i];
    float multi = (float)g.vw[i] * lam;
    float igain = M.gain[i];
    i32 hi = (i32)xorshift32((u32)i);
    i64 d = g.deg(i);
    if (d < LARGE_CUTOFF) {
        float acc = 0.0f;
        for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++) ab_term(M, j, p, best, multi, igain, hi, acc);
        return acc;
    }
    return tree_fold(M, i, p, best, multi, igain, hi, d < MASSIVE_CUTOFF ? 64 : 1024);
}

static std::vector<i32> afterburner(Mover& M, State& st, const std::vector<i32>& cands,
                                    std::vector<float>* change_out = nullptr) {
    float lam = (float)st.lambda;
    const i64 k = (i64)cands.size();
    std::vector<float> ch(k);
#pragma omp parallel for schedule(dynamic, 512)
    for (i64 x = 0; x < k; x++) ch[x] = ab_change(M, cands[x], lam);
    std::vector<i32> out;
    for (i64 x = 0; x < k; x++)
        if (M.gain[cands[x]] + ch[x] >= 0) out.push_back(cands[x]);  // :332-334
    if (change_out) *change_out = std::move(ch);
    return out;
}

// set_new_cluster_ids (src/local_mover.cpp:215-274)
static void set_new_cluster_ids(Mover& M, State& st, const std::vector<i32>& moves, bool constrained, i32* constraint) {
    std::vector<i32> nv;
    for (i32 i : moves)
        if (M.dest[i] == NEW_PART) nv.push_back(i);
    if (nv.empty()) return;
    i32 curr = st.label_count;
    std::vector<char> has(curr, 0);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < M.g.n; i++) __atomic_store_n(&has[M.part[i]], (char)1, __ATOMIC_RELAXED);
    i64 k = 0;
    for (i32 l = 0; l < curr && k < (i64)nv.size(); l++) {
        if (!has[l]) {
            i32 x = nv[k++];
            M.dest[x] = l;
            if (constrained) constraint[l] = constraint[M.part[x]];
        }
    }
    i64 avail = k;
    if (avail < (i64)nv.size()) {
        i64 needed = (i64)nv.size() - avail;
        st.label_count += (i32)needed;
        if ((i64)st.vol.size() < st.label_count) st.vol.resize(st.label_count, 0);
        for (i64 x = avail; x < (i64)nv.size(); x++) {
            i32 i = nv[x];
            i32 d = (i32)(x - avail) + curr;
            M.dest[i] = d;
            if (constrained) constraint[d] = constraint[M.part[i]];
        }
        for (i32 l = curr; l < st.label_count; l++) st.vol[l] = 0;
    }
}

// perform_moves (src/local_mover.cpp:1241-1279). The uncut update equals
// sum(pvals_after) - sum(pvals_before) (:1274-1277), computed here directly
// from the moved vertices' adjacency; the dest cache is reset to NULL_PART for
// every vertex on a full table build (:1264-1266) and otherwise for the movers
// and their neighbours (find_affected* :717,798; update_small :1072,1121).
static void perform_moves(Mover& M, State& st, const std::vector<i32>& moves) {
    const Graph& g = M.g;
    std::vector<char> mv(g.n, 0);
    for (i32 i : moves) mv[i] = 1;
    i64 delta = 0;
    const i64 k = (i64)moves.size();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : delta)
    for (i64 x = 0; x < k; x++) {
        const i32 i = moves[x];
        i32 oi = M.part[i], ni = M.dest[i];
        for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++) {
            i32 v = g.col[j];
            i32 ov = M.part[v], nvv = mv[v] ? M.dest[v] : M.part[v];
            i64 d = (i64)(ni == nvv) - (i64)(oi == ov);
            delta += d * g.wt(j) * (mv[v] ? 1 : 2);
        }
    }
    for (i32 i : moves) {
        st.vol[M.part[i]] -= g.vw[i];
        st.vol[M.dest[i]] += g.vw[i];
    }
    for (i32 i : moves) M.part[i] = M.dest[i];
    if (!M.tables) {
        M.tables = true;
        std::fill(M.dest.begin(), M.dest.end(), NULL_PART);
    } else {
#pragma omp parallel for schedule(dynamic, 256)
        for (i64 x = 0; x < k; x++) {
            const i32 i = moves[x];
            __atomic_store_n(&M.dest[i], NULL_PART, __ATOMIC_RELAXED);
            for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++)
                __atomic_store_n(&M.dest[g.col[j]], NULL_PART, __ATOMIC_RELAXED);
        }
    }
    st.uncut += delta;
    st.update_objective();
}

// relabel_contiguously (src/local_mover.cpp:100-129)
static void relabel(std::vector<i32>& labels, State& st) {
    i32 L = st.label_count;
    std::vector<i32> used(L, 0);
    for (i32 l : labels) used[l] = 1;
    i32 t = 0;
    std::vector<i64> nvol;
    for (i32 l = 0; l < L; l++) {
        if (used[l]) {
            used[l] = t++;
            nvol.push_back(st.vol[l]);
        }
    }
    for (i32 l = 0; l < t; l++) st.vol[l] = nvol[l];
    for (auto& x : labels) x = used[x];
    st.label_count = t;
}

struct Stats {
    i64 sweeps = 0, moves = 0, cands = 0;
};

// Per-sweep record of local_move / local_move_strict, kept while a test asks
// for it: ratio, candidates, accepted moves, and the objective after the
// sweep (the current state, not the best one).
struct SweepRec {
    float ratio;
    i64 cands, moves;
    double obj, seconds;
};
static std::vector<SweepRec>* g_trace = nullptr;
static std::chrono::steady_clock::time_point g_sweep_t0;
static void trace_sweep(float ratio, i64 cands, i64 moves, double obj) {
    if (g_trace) {
        const double sec =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - g_sweep_t0).count();
        g_trace->push_back(SweepRec{ratio, cands, moves, obj, sec});
    }
}
// sweeps per local_move call (0 = the reference's schedule; a positive value
// stops a call early, for bounded CPU timing samples only)
static int g_sweep_limit = 0;

// local_move (src/local_mover.cpp:1360-1417)
static void local_move(const Graph& g, std::vector<i32>& best_part, State& best, bool is_initial, bool constrained,
                       i32* constraint, bool SA, Stats* stats) {
    State curr = best;
    curr.copy_from(best);
    Mover M(g);
    M.part = best_part;
    M.dest.assign(g.n, NULL_PART);
    M.gain.assign(g.n, 0.0f);
    M.tables = !is_initial;
#if ORC_MUTANT == 3  // planted mistake: second filter ratio
    const float ratios[2] = {0.75f, 0.5f};
#else
    const float ratios[2] = {0.75f, 0.25f};
#endif
    const int limits[2] = {4, 2};
    int iter = 0;
    for (int x = 0; x < 2; x++) {
        float fr = SA ? ratios[x] : 0.0f;
        for (int c = 0; c < limits[x]; c++) {
            iter++;
            if (g_sweep_limit > 0 && iter > g_sweep_limit) break;
            g_sweep_t0 = std::chrono::steady_clock::now();
            std::vector<i32> cands = candidates(M, curr, fr, constrained, constraint);
            if (stats) { stats->sweeps++; stats->cands += (i64)cands.size(); }
            if (cands.empty()) {
                trace_sweep(fr, 0, 0, curr.obj);
                break;
            }
            std::vector<i32> moves = afterburner(M, curr, cands);
            if (moves.empty()) {
                trace_sweep(fr, (i64)cands.size(), 0, curr.obj);
                break;
            }
            if (iter > 1 || !is_initial) set_new_cluster_ids(M, curr, moves, constrained, constraint);
            perform_moves(M, curr, moves);
            if (stats) stats->moves += (i64)moves.size();
            trace_sweep(fr, (i64)cands.size(), (i64)moves.size(), curr.obj);
#if ORC_MUTANT == 4  // planted mistake: the last state is kept, not the best
            if (true) {
#else
            if (curr.obj > best.obj) {
#endif
                best.copy_from(curr);
                best_part = M.part;
            }
        }
    }
    relabel(best_part, best);
}

// afterburner_filter_strict (src/local_mover.cpp:131-213). Canonical FP32
// order: per move a sequential fold in adjacency order then over the earlier
// moves (as written, :154-177); the prefix sum is one sequential FP32 scan in
// move-list order (Kokkos parallel_scan order is unspecified in the reference).
static std::vector<i32> strict_filter(Mover& M, State& st, const std::vector<i32>& moves, double& diff) {
    const Graph& g = M.g;
    float lam = (float)st.lambda;
    i64 k = (i64)moves.size();
    std::vector<i64> prio(g.n, g.n + 1);
    for (i64 x = 0; x < k; x++) prio[moves[x]] = x;
    std::vector<float> sg(k);
    for (i64 x = 0; x < k; x++) {
        i32 i = moves[x];
        i32 best = M.dest[i], p = M.part[i];
        float multi = (float)g.vw[i] * lam;
        float oc = M.gain[i];
        for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++) {
            i32 v = g.col[j];
            i64 wgt = g.wt(j);
            if (prio[v] < prio[i]) {
                i32 vp = M.dest[v];
                oc = oc - (float)((vp == p) ? wgt : 0);
                oc = oc + (float)((vp == best && best != NEW_PART) ? wgt : 0);
                vp = M.part[v];
                oc = oc + (float)((vp == p) ? wgt : 0);
                oc = oc - (float)((vp == best) ? wgt : 0);
            }
        }
        for (i64 jx = 0; jx < x; jx++) {
            i32 v = moves[jx];
            i32 vp = M.dest[v];
            float wg = -multi * (float)g.vw[v];
            oc = oc - ((vp == p) ? wg : 0.0f);
            oc = oc + ((vp == best && best != NEW_PART) ? wg : 0.0f);
            vp = M.part[v];
            oc = oc + ((vp == p) ? wg : 0.0f);
            oc = oc - ((vp == best) ? wg : 0.0f);
        }
        sg[x] = oc;
    }
    float run = 0.0f;
    for (i64 x = 0; x < k; x++) {
        run = run + sg[x];
        sg[x] = run;
    }
    float bv = -1.0f;
    i64 bl = g.n;
    for (i64 x = 0; x < k; x++)
        if (sg[x] > bv) { bv = sg[x]; bl = x; }
    diff = 0;
    i64 trunc = 0;
    if (bl < k && bv > 0) { trunc = bl + 1; diff = (double)bv * 2; }
    for (i64 x = 0; x < k; x++) M.gain[moves[x]] = sg[x];
    return std::vector<i32>(moves.begin(), moves.begin() + trunc);
}

// local_move_strict (src/local_mover.cpp:1419-1483)
static void local_move_strict(const Graph& g, std::vector<i32>& best_part, State& best, bool is_initial,
                              bool constrained, i32* constraint, Stats* stats) {
    State curr = best;
    curr.copy_from(best);
    Mover M(g);
    M.part = best_part;
    M.dest.assign(g.n, NULL_PART);
    M.gain.assign(g.n, 0.0f);
    M.tables = !is_initial;
    for (int it = 0; it < 6; it++) {
        std::vector<i32> moves = candidates(M, curr, 0.0f, constrained, constraint);
        if (stats) { stats->sweeps++; stats->cands += (i64)moves.size(); }
        if (moves.empty()) break;
        if (it > 0 || !is_initial) set_new_cluster_ids(M, curr, moves, constrained, constraint);
        double diff = 0;
        moves = strict_filter(M, curr, moves, diff);
        if (moves.empty()) break;
        perform_moves(M, curr, moves);
        if (stats) stats->moves += (i64)moves.size();
        if (curr.obj > best.obj) {
            best.copy_from(curr);
            best_part = M.part;
        }
        std::fill(M.dest.begin(), M.dest.end(), NULL_PART);  // :1478-1480
    }
    relabel(best_part, best);
}

// own-cluster weight of every vertex under labels (the pvals_clone contract,
// src/local_mover.cpp:1354-1358, consumed by src/leidenR.cpp:279)
static std::vector<i64> own_weight(const Graph& g, const std::vector<i32>& labels) {
    std::vector<i64> pv(g.n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
    for (i64 i = 0; i < g.n; i++)
        for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++)
            if (labels[g.col[j]] == labels[i]) pv[i] += g.wt(j);
    return pv;
}

// ----------------------------------------------------------------- A11 leidenR
struct Rng {
    u64 seed;
    u64 k = 0;
    u32 draw() { return seed_draw(seed, k++); }
};

// find_trees (src/leidenR.cpp:212-260)
static void find_trees(std::vector<i32>& vcmap, i64 n, const std::vector<i32>& hn, std::vector<i32>& order) {
    for (i64 i = 0; i < n; i++) {
        i32 now = (i32)i, then = hn[i];
        if (now == then) { vcmap[i] = (i32)i; continue; }
        while (order[now] > order[then]) { now = then; then = hn[then]; }
        if (now != i) { vcmap[i] = now; vcmap[now] = now; }
    }
    for (i64 i = 0; i < n; i++)
        if (vcmap[i] == ORD_MAX) order[i] = 2 * SPLIT - order[i];
    // part 2 is race-free in effect: every chase from an unassigned vertex
    // follows the same pointer path, so sequential order gives the same labels
    for (i64 i = 0; i < n; i++) {
        if (vcmap[i] != ORD_MAX) continue;
        i32 now = (i32)i, then = hn[i];
        while (vcmap[now] == ORD_MAX && order[now] > order[then]) { now = then; then = hn[then]; }
        vcmap[i] = vcmap[now];
    }
}

// ensure_gamma_connectivity (src/leidenR.cpp:41-187). Canonical rule: the
// vertices of one tree are sorted by (order, vertex id); the reference's
// KokkosKernels sort leaves ties to its bucket insertion order.
static void ensure_gamma(const Graph& g, std::vector<i32>& vcmap, const std::vector<i32>& constraint,
                         const std::vector<i32>& order, const std::vector<i64>& total_deg,
                         const std::vector<i64>& pvals, float gamma) {
    i64 n = g.n;
    std::vector<i32> ids(n);
    std::iota(ids.begin(), ids.end(), 0);
    __gnu_parallel::sort(ids.begin(), ids.end(), [&](i32 a, i32 b) {
        if (vcmap[a] != vcmap[b]) return vcmap[a] < vcmap[b];
        if (order[a] != order[b]) return order[a] < order[b];
        return a < b;
    });
    std::vector<i64> cbeg(n + 1, 0);
    for (i64 i = 0; i < n; i++) cbeg[vcmap[i] + 1]++;
    for (i64 c = 0; c < n; c++) cbeg[c + 1] += cbeg[c];
    std::vector<i64> tsize(n);  // exclusive prefix of vertex weight, global
    {
        i64 run = 0;
        for (i64 x = 0; x < n; x++) { tsize[x] = run; run += g.vw[ids[x]]; }
    }
    std::vector<i64> inner(n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
    for (i64 i = 0; i < n; i++) {
        i64 r = 0;
        for (i64 j = g.rp[i]; j < g.rp[i + 1]; j++) {
            i32 v = g.col[j];
            if (vcmap[i] == vcmap[v] && order[v] < order[i]) r += g.wt(j);
        }
        inner[i] = r;
    }
    std::vector<i64> breakers(n, n);
    for (i64 x = 0; x < n; x++) {
        i32 i = ids[x];
        i32 c = vcmap[i];
        i64 prev = tsize[x] - tsize[cbeg[c]];
#if ORC_MUTANT == 5  // planted mistake: the vertex's own weight left out of the inner test
        if ((float)inner[i] < gamma * (float)prev) breakers[c] = std::min(breakers[c], x);
#else
        if ((float)inner[i] < gamma * (float)g.vw[i] * (float)prev) breakers[c] = std::min(breakers[c], x);
#endif
    }
    std::vector<i64> outer(n);
    {
        i64 run = 0;
        for (i64 x = 0; x < n; x++) {
            i32 i = ids[x];
            run += pvals[i] - 2 * inner[i];
            outer[i] = run;
        }
    }
    for (i64 x = 0; x < n; x++) {
        i32 i = ids[x];
        i32 c = vcmap[i];
        i64 cb = cbeg[c];
        i64 cur = tsize[x] + g.vw[i] - tsize[cb];
        i64 other = total_deg[constraint[i]] - cur;
        i64 out = outer[i];
        if (cb > 0) out -= outer[ids[cb - 1]];
#if ORC_MUTANT == 6  // planted mistake: the outer test breaks before the vertex, not after it
        if ((float)out < gamma * (float)cur * (float)other) breakers[c] = std::min(breakers[c], x);
#else
        if ((float)out < gamma * (float)cur * (float)other) breakers[c] = std::min(breakers[c], x + 1);
#endif
    }
    for (i64 x = 0; x < n; x++) {
        i32 i = ids[x];
        if (x >= breakers[vcmap[i]]) vcmap[i] = i;
    }
}

// contigitize_clusters (src/leidenR.cpp:189-210)
static i64 contigitize(std::vector<i32>& vcmap, i64 n) {
    std::vector<i32> rank(n, -1);
    i64 nc = 0;
    for (i64 u = 0; u < n; u++)
        if (vcmap[u] == u) rank[u] = (i32)nc++;
    for (i64 u = 0; u < n; u++) vcmap[u] = rank[vcmap[u]];
    return nc;
}

// coarsen_leidenR (src/leidenR.cpp:263-500); `uniform` selects the unit-weight
// neighbour rule used only on the top level of unit-weight graphs (:173).
// Trees and random order of the last leiden_refine call, kept while a test
// asks for them.
struct RefineDetail {
    std::vector<i32> trees, order;
};
static RefineDetail* g_refine = nullptr;

static std::vector<i32> leiden_refine(const Graph& g, const std::vector<i32>& constraint, const State& rfd,
                                      const std::vector<i64>& pvals, bool uniform, Rng& rng, i64& nc) {
    i64 n = g.n;
    float gamma = (float)rfd.lambda;
    std::vector<i32> vcmap(n, ORD_MAX), hn(n), well(n);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; i++) {  // :281-287
        i64 wd = g.vw[i];
        well[i] = ((float)pvals[i] >= gamma * (float)wd * (float)(rfd.vol[constraint[i]] - wd)) ? 1 : 0;
    }
    if (uniform) {
        // :296-393: first satisfactory neighbour cyclically from a hashed start
        i32 seed = (i32)rng.draw();
#pragma omp parallel for schedule(dynamic, 4096)
        for (i64 i = 0; i < n; i++) {
            i64 s = g.rp[i], e = g.rp[i + 1], width = e - s;
            float multi = gamma * (float)g.vw[i];
            u32 h = murmur3_i32((i32)((u32)seed + (u32)i));
            i32 jx32 = (i32)h;
            if (jx32 % 1000 == 0 || well[i] == 0 || width == 0) { hn[i] = (i32)i; continue; }
            i64 jx = s + (i64)(h % (u32)width);
            i32 pick = (i32)i;
            bool found = false;
            for (i64 j = jx; j < e && !found; j++) {
                i32 v = g.col[j];
                if (constraint[i] != constraint[v] || well[v] == 0) continue;
                if ((float)g.wt(j) >= multi * (float)g.vw[v]) { pick = v; found = true; }
            }
            for (i64 j = s; j < jx && !found; j++) {
                i32 v = g.col[j];
                if (constraint[i] != constraint[v] || well[v] == 0) continue;
                if ((float)g.wt(j) >= multi * (float)g.vw[v]) { pick = v; found = true; }
            }
            hn[i] = pick;
        }
    } else {
        // :395-474: weighted random argmax of rand * (w - gamma w_i w_v)^4.
        // Canonical rule: degree < 128 keeps the last maximum (serial path),
        // degree >= 128 keeps the first maximum (MaxFirstLoc team path).
        // (w - ...)^4 is evaluated as ((double)val^2)^2 in double.
        i32 seed = (i32)rng.draw();
#pragma omp parallel for schedule(dynamic, 4096)
        for (i64 i = 0; i < n; i++) {
            i64 s = g.rp[i], e = g.rp[i + 1];
            float multi = gamma * (float)g.vw[i];
            if (murmur3_i32((i32)((u32)seed + (u32)i)) % 10000u == 0 || well[i] == 0) { hn[i] = (i32)i; continue; }
            bool last = g.deg(i) < LARGE_CUTOFF;
            float m = 0.0f;
            i64 am = e;
            for (i64 j = s; j < e; j++) {
                i32 v = g.col[j];
                if (constraint[i] != constraint[v] || well[v] == 0) continue;
                float val = fmaf(-multi, (float)g.vw[v], (float)g.wt(j));
                if (val >= 0) {
                    float smp = (float)murmur3_i32((i32)((u32)seed + (u32)j)) / (float)ORD_MAX;
                    double v2 = (double)val * (double)val;
                    smp = (float)((double)smp * (v2 * v2));
                    if (last ? (smp >= m) : (smp > m || am == e)) { m = smp; am = j; }
                }
            }
            hn[i] = (am >= s && am < e) ? g.col[am] : (i32)i;
        }
    }
    std::vector<i32> order(n);
    {
        i32 seed = (i32)rng.draw();
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; i++) {  // :482-486
            i32 o = (i32)murmur3_i32((i32)((u32)i + (u32)seed));
            i32 r = o % SPLIT;
            order[i] = r < 0 ? -r : r;
        }
    }
    for (i64 i = 0; i < n; i++)  // :487-494
        if (order[i] == order[hn[i]]) hn[i] = (i32)i;
    find_trees(vcmap, n, hn, order);
    if (g_refine) {
        g_refine->trees = vcmap;
        g_refine->order = order;
    }
    ensure_gamma(g, vcmap, constraint, order, rfd.vol, pvals, gamma);
    nc = contigitize(vcmap, n);
    return vcmap;
}

// ------------------------------------------------------------- A13 drivers
static void coarsen_vtx_w(const Graph& g, const std::vector<i32>& map, Graph& c) {  // src/clustering_methods.cpp:33-38
    c.vw.assign(c.n, 0);
    for (i64 i = 0; i < g.n; i++) c.vw[map[i]] += g.vw[i];
}

// Canonical rule for downsample (src/clustering_methods.cpp:40-45) when the
// members of one cluster carry different values: the highest vertex id wins.
static void downsample(const std::vector<i32>& in, std::vector<i32>& out, const std::vector<i32>& map) {
    for (size_t i = 0; i < in.size(); i++) out[map[i]] = in[i];
}

struct Run {
    Stats stats;
    i64 levels = 0;
    i64 total_nnz = 0;
};

// leiden_part<plus, improve> (src/clustering_methods.cpp:47-137)
static std::vector<i32> leiden_part(const Graph& top, State& rfd, bool plus, bool improve,
                                    const std::vector<i32>* input, Rng& rng, Run& run) {
    std::vector<Graph> levels;
    std::vector<std::vector<i32>> parts;
    levels.push_back(top);
    std::vector<i32> part(top.n);
    if (!improve) std::iota(part.begin(), part.end(), 0);
    else part = *input;
    int stalled = 0;  // consecutive refinements that contracted nothing
    while (true) {
        const Graph& c = levels.back();
        double old_obj = rfd.obj;
        bool init = !improve && levels.size() == 1;
        local_move(c, part, rfd, init, false, nullptr, true, &run.stats);
        if (c.nnz() > 100000 && old_obj == rfd.obj) local_move(c, part, rfd, init, false, nullptr, false, &run.stats);
        if (old_obj == rfd.obj) local_move_strict(c, part, rfd, init, false, nullptr, &run.stats);
        if (rfd.label_count == c.n) { parts.push_back(part); break; }
        std::vector<i32> louv = part;
        std::vector<i64> pv = own_weight(c, louv);
        i64 nc = 0;
        bool uni = levels.size() == 1 && c.uniform;
        std::vector<i32> cmap = leiden_refine(c, louv, rfd, pv, uni, rng, nc);
        parts.push_back(cmap);
        if (nc < c.n) {
            Graph next = contract(c, cmap, nc);
            coarsen_vtx_w(c, cmap, next);
            part.assign(nc, 0);
            downsample(louv, part, cmap);
            levels.push_back(std::move(next));
            stalled = 0;
        } else {
            // the reference retries forever when no neighbour passes the
            // refinement tests (:105-110); stop after 8 tries at the current
            // clustering (the map just pushed is the identity)
            if (++stalled >= 8) {
                parts.pop_back();
                parts.push_back(part);
                break;
            }
            Graph same = c;
            levels.push_back(std::move(same));
        }
    }
    run.levels = (i64)levels.size();
    for (auto& l : levels) run.total_nnz += l.nnz();
    for (int i = (int)levels.size() - 2; i >= 0; i--) {
        const Graph& c = levels[i];
        std::vector<i32>& cp = parts[i + 1];
        std::vector<i32>& fp = parts[i];
        for (i64 x = 0; x < c.n; x++) fp[x] = cp[fp[x]];
        if (plus) local_move(c, fp, rfd, false, false, nullptr, true, &run.stats);
    }
    return parts[0];
}

// louvain_part<constrained> (src/clustering_methods.cpp:139-222)
static std::vector<i32> louvain_part(const Graph& top, State& rfd, bool constrained, std::vector<i32> constraint,
                                     Run& run) {
    if (constrained) {  // cluster_data::reset (interface/cluster_data.hpp:49-55)
        std::copy(top.vw.begin(), top.vw.end(), rfd.vol.begin());
        rfd.uncut = 0;
        rfd.label_count = (i32)top.n;
        rfd.update_objective();
    }
    std::vector<Graph> levels;
    std::vector<std::vector<i32>> parts;
    levels.push_back(top);
    bool drop = constrained;
    while (true) {
        const Graph& c = levels.back();
        std::vector<i32> part(c.n);
        std::iota(part.begin(), part.end(), 0);
        i32* cons = constrained ? constraint.data() : nullptr;
        local_move(c, part, rfd, true, constrained, cons, true, &run.stats);
        if (constrained && rfd.label_count == c.n) local_move_strict(c, part, rfd, true, constrained, cons, &run.stats);
        parts.push_back(part);
        if (rfd.label_count < c.n) {
            Graph next = contract(c, part, rfd.label_count);
            next.vw.assign(rfd.vol.begin(), rfd.vol.begin() + rfd.label_count);
            if (constrained) {
                std::vector<i32> nc(rfd.label_count, 0);
                downsample(constraint, nc, part);
                constraint = nc;
            }
            levels.push_back(std::move(next));
        } else if (drop) {
            std::fill(constraint.begin(), constraint.end(), 0);
            parts.pop_back();
            drop = false;
        } else {
            break;
        }
    }
    run.levels = (i64)levels.size();
    for (auto& l : levels) run.total_nnz += l.nnz();
    if (levels.size() > 1) { levels.pop_back(); parts.pop_back(); }
    for (int i = (int)levels.size() - 2; i >= 0; i--) {
        const Graph& c = levels[i];
        std::vector<i32>& cp = parts[i + 1];
        std::vector<i32>& fp = parts[i];
        for (i64 x = 0; x < c.n; x++) fp[x] = cp[fp[x]];
        local_move(c, fp, rfd, false, false, nullptr, true, &run.stats);
    }
    return parts[0];
}

// Symmetric CSR of sorted, deduplicated pair keys (lo << 32 | hi, lo < hi):
// both directions stored, rows sorted; weights (wmax > 0) drawn per pair.
static Graph csr_from_pairs(i64 n, const std::vector<u64>& key, u64 seed, i32 wmin, i32 wmax) {
    Graph g;
    g.n = n;
    g.rp.assign(n + 1, 0);
    for (u64 k : key) {
        g.rp[(k >> 32) + 1]++;
        g.rp[(u32)k + 1]++;
    }
    for (i64 i = 0; i < n; i++) g.rp[i + 1] += g.rp[i];
    g.col.resize(g.rp[n]);
    g.uniform = wmax <= 0;
    if (!g.uniform) g.w.resize(g.rp[n]);
    // pairs in (lo, hi) order reach every row in increasing column order
    std::vector<i64> pos(g.rp.begin(), g.rp.end() - 1);
    for (u64 k : key) {
        const u32 lo = (u32)(k >> 32), hi = (u32)k;
        const i64 a0 = pos[lo]++, a1 = pos[hi]++;
        g.col[a0] = (i32)hi;
        g.col[a1] = (i32)lo;
        if (!g.uniform) g.w[a0] = g.w[a1] = pair_weight(lo, hi, seed, wmin, wmax);
    }
    return g;
}

// sort, drop the invalid key ~0 and repeats
static void sort_unique(std::vector<u64>& key) {
    __gnu_parallel::sort(key.begin(), key.end());
    i64 u = 0;
    for (size_t e = 0; e < key.size(); e++) {
        if (key[e] == ~(u64)0) break;
        if (u == 0 || key[e] != key[u - 1]) key[u++] = key[e];
    }
    key.resize(u);
}

// Symmetric CSR of an R-MAT draw: 2^scale vertices, edge_factor * 2^scale
// draws, self loops dropped, repeated pairs merged, both directions stored,
// rows sorted; weights (wmax > 0) are drawn per undirected pair.
static Graph build_rmat(int scale, int edge_factor, double a, double b, double c, u64 seed, i32 wmin, i32 wmax) {
    const i64 n = (i64)1 << scale;
    const i64 m = (i64)edge_factor << scale;
    u64 t[3];
    rmat_thresholds(a, b, c, t);
    std::vector<u64> key(m);
#pragma omp parallel for schedule(static)
    for (i64 e = 0; e < m; e++) {
        u32 s, d;
        rmat_edge((u64)e, scale, t, seed, s, d);
        const u32 lo = std::min(s, d), hi = std::max(s, d);
        key[e] = lo == hi ? ~(u64)0 : ((u64)lo << 32) | hi;
    }
    sort_unique(key);
    return csr_from_pairs(n, key, seed, wmin, wmax);
}

// ---------------------------------------------------------------- LFR-style
// Benchmark graph in the manner of Lancichinetti, Fortunato and Radicchi
// (Phys. Rev. E 78, 046110, 2008), stated for exact CPU/GPU agreement:
//  - integer power laws: X with density ~ x^-tau on [x0, top + 1), value
//    floor(X); the degree law's x0 is solved by bisection for the mean
//    degree, the community-size law starts at cmin; a u32 Philox draw r maps
//    to the first value whose cumulative threshold floor(CDF * 2^32) exceeds r;
//  - community sizes are drawn in sequence until they cover n; the last takes
//    what is left and joins the one before if that is below cmin;
//  - vertices fill the communities in id order; internal degree
//    round((1 - mu) * degree) (mu in thousandths), at most size - 1;
//  - stubs: internal ones sorted (stably) by (community, draw), paired in
//    order inside each community; external ones sorted by draw and paired
//    in order over the graph; pairs of one vertex and repeated pairs vanish in
//    the simple-graph build;
//  - vertex ids are permuted by a stable sort on one draw per vertex.
struct PLTable {
    i32 k0 = 0;
    std::vector<u64> thr;
};

static void pl_probs(double tau, double x0, i32 top, std::vector<double>& p, i32& k0) {
    k0 = (i32)std::floor(x0);
    const double a = std::pow(x0, 1.0 - tau);
    const double b = std::pow((double)top + 1.0, 1.0 - tau);
    const double z = a - b;
    p.assign((size_t)(top - k0 + 1), 0.0);
    double prev = a;
    for (i32 k = k0; k <= top; k++) {
        const double nxt = std::pow((double)k + 1.0, 1.0 - tau);
        p[(size_t)(k - k0)] = (prev - nxt) / z;
        prev = nxt;
    }
}

static double pl_mean(double tau, double x0, i32 top) {
    std::vector<double> p;
    i32 k0;
    pl_probs(tau, x0, top, p, k0);
    double m = 0.0;
    for (size_t i = 0; i < p.size(); i++) m += (double)(k0 + (i32)i) * p[i];
    return m;
}

static PLTable power_law_table(double tau, i32 lo, i32 top, double mean) {
    double x0 = (double)lo;
    if (mean > 0.0) {
        double a = (double)lo, b = (double)top;
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (a + b);
            if (pl_mean(tau, mid, top) < mean) a = mid;
            else b = mid;
        }
        x0 = 0.5 * (a + b);
    }
    std::vector<double> p;
    PLTable t;
    pl_probs(tau, x0, top, p, t.k0);
    t.thr.resize(p.size());
    double cum = 0.0;
    for (size_t i = 0; i < p.size(); i++) {
        cum += p[i];
        double f = std::floor(cum * 4294967296.0);
        if (f < 0.0) f = 0.0;
        if (f > 4294967296.0) f = 4294967296.0;
        t.thr[i] = (u64)f;
    }
    t.thr.back() = (u64)1 << 32;
    return t;
}

static i32 pl_draw(const PLTable& t, u32 r) {
    const i32 i = (i32)(std::upper_bound(t.thr.begin(), t.thr.end(), (u64)r) - t.thr.begin());
    return t.k0 + i;
}

static inline u32 draw32(u64 x, u32 stream, u64 seed) {
    return philox4x32_10(P4{{(u32)x, (u32)(x >> 32), 0u, stream}}, (u32)seed, (u32)(seed >> 32)).v[0];
}

struct LfrSpec {
    i64 n;
    double tau1, kmean, tau2, mu;
    i32 kmax, cmin, cmax, wmin, wmax;
    u64 seed;
};

static std::vector<i64> lfr_sizes(const LfrSpec& p) {
    const PLTable t = power_law_table(p.tau2, p.cmin, p.cmax, 0.0);
    std::vector<i64> sz;
    i64 sum = 0;
    for (u64 j = 0; sum < p.n; j++) {
        const i64 s = pl_draw(t, draw32(j, STREAM_LFR_COMM, p.seed));
        sz.push_back(s);
        sum += s;
    }
    if (!sz.empty()) {
        const i64 rem = p.n - (sum - sz.back());
        sz.back() = rem;
        if (rem < p.cmin && sz.size() > 1) {
            sz[sz.size() - 2] += rem;
            sz.pop_back();
        }
    }
    return sz;
}

// stable sort of (key, vertex) records by key
static void stable_by_key(std::vector<std::pair<u64, i32>>& r) {
    __gnu_parallel::stable_sort(r.begin(), r.end(),
                                [](const std::pair<u64, i32>& a, const std::pair<u64, i32>& b) { return a.first < b.first; });
}

static Graph build_lfr(const LfrSpec& p, std::vector<i32>* gt, std::vector<i32>* degree) {
    const i64 n = p.n;
    const PLTable dt = power_law_table(p.tau1, 1, p.kmax, p.kmean);
    const std::vector<i64> sz = lfr_sizes(p);
    const i64 nc = (i64)sz.size();
    std::vector<i64> cs(nc + 1, 0);
    for (i64 k = 0; k < nc; k++) cs[k + 1] = cs[k] + sz[k];
    const i64 mu_m = std::llround(p.mu * 1000.0);
    std::vector<i32> kin(n), kext(n), comm(n);
    for (i64 c = 0; c < nc; c++) {
        for (i64 v = cs[c]; v < cs[c + 1]; v++) {
            const i64 deg = pl_draw(dt, draw32((u64)v, STREAM_LFR_DEG, p.seed));
            i64 ki = (deg * (1000 - mu_m) + 500) / 1000;
            ki = std::min(ki, sz[c] - 1);
            kin[v] = (i32)ki;
            kext[v] = (i32)(deg - ki);
            comm[v] = (i32)c;
        }
    }
    if (degree) {
        degree->resize(n);
        for (i64 v = 0; v < n; v++) (*degree)[v] = kin[v] + kext[v];
    }
    // vertex permutation: rank of each vertex's draw
    std::vector<i32> pi(n);
    {
        std::vector<std::pair<u64, i32>> r(n);
        for (i64 v = 0; v < n; v++) r[v] = {draw32((u64)v, STREAM_LFR_PERM, p.seed), (i32)v};
        stable_by_key(r);
        for (i64 k = 0; k < n; k++) pi[r[k].second] = (i32)k;
    }
    if (gt) {
        gt->resize(n);
        for (i64 v = 0; v < n; v++) (*gt)[pi[v]] = comm[v];
    }
    std::vector<u64> key;
    auto add_pair = [&](i32 a, i32 b) {
        const u32 u = (u32)pi[a], w = (u32)pi[b];
        if (u != w) key.push_back(((u64)std::min(u, w) << 32) | std::max(u, w));
    };
    {  // internal stubs, paired inside each community
        std::vector<i64> oin(n + 1, 0);
        for (i64 v = 0; v < n; v++) oin[v + 1] = oin[v] + kin[v];
        std::vector<std::pair<u64, i32>> r(oin[n]);
        for (i64 v = 0; v < n; v++)
            for (i64 q = oin[v]; q < oin[v + 1]; q++)
                r[q] = {((u64)(u32)comm[v] << 32) | draw32((u64)q, STREAM_LFR_ISTUB, p.seed), (i32)v};
        stable_by_key(r);
        for (i64 c = 0; c < nc; c++) {
            const i64 s0 = oin[cs[c]], len = oin[cs[c + 1]] - s0;
            for (i64 q = 0; q + 1 < len; q += 2) add_pair(r[s0 + q].second, r[s0 + q + 1].second);
        }
    }
    {  // external stubs, paired over the whole graph
        std::vector<i64> oext(n + 1, 0);
        for (i64 v = 0; v < n; v++) oext[v + 1] = oext[v] + kext[v];
        std::vector<std::pair<u64, i32>> r(oext[n]);
        for (i64 v = 0; v < n; v++)
            for (i64 q = oext[v]; q < oext[v + 1]; q++) r[q] = {draw32((u64)q, STREAM_LFR_ESTUB, p.seed), (i32)v};
        stable_by_key(r);
        for (i64 q = 0; q + 1 < (i64)r.size(); q += 2) add_pair(r[q].second, r[q + 1].second);
    }
    sort_unique(key);
    return csr_from_pairs(n, key, p.seed, p.wmin, p.wmax);
}

}  // namespace orc

// =====================================================================
// C ABI for the tests (host arrays only)
// =====================================================================
using namespace orc;

static Graph make_graph(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw) {
    Graph g;
    g.n = n;
    g.rp.assign(rp, rp + n + 1);
    g.col.assign(col, col + rp[n]);
    if (w32) { g.uniform = false; g.w.assign(w32, w32 + rp[n]); }
    else if (w64) { g.uniform = false; g.w.assign(w64, w64 + rp[n]); }
    if (vw) g.vw.assign(vw, vw + n);
    else g.vw.assign(n, 1);
    return g;
}

extern "C" {

void or_set_threads(int k) { omp_set_num_threads(k > 0 ? k : 1); }
void or_set_sweep_limit(int k) { g_sweep_limit = k > 0 ? k : 0; }
int or_get_threads(void) { return omp_get_max_threads(); }

void or_philox(const u32* ctr, const u32* key, u32* out) {
    P4 r = philox4x32_10(P4{{ctr[0], ctr[1], ctr[2], ctr[3]}}, key[0], key[1]);
    std::copy(r.v, r.v + 4, out);
}

// R-MAT graph held by the oracle until or_graph_free
void* or_rmat(int scale, int edge_factor, double a, double b, double c, u64 seed, i32 wmin, i32 wmax) {
    return new Graph(build_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax));
}
// LFR-style graph held by the oracle until or_graph_free; gt (n entries,
// nullable) receives the planted communities, degree (n, nullable) the drawn
// degrees in generation order, sizes (nullable, sizes_cap entries) the
// community sizes; returns the number of communities
void* or_lfr(i64 n, double tau1, double kmean, i32 kmax, double tau2, i32 cmin, i32 cmax, double mu, i32 wmin,
             i32 wmax, u64 seed, i32* gt, i32* degree, i64* sizes, i64 sizes_cap, i64* ncomm) {
    const LfrSpec p{n, tau1, kmean, tau2, mu, kmax, cmin, cmax, wmin, wmax, seed};
    std::vector<i32> g, d;
    Graph* out = new Graph(build_lfr(p, gt ? &g : nullptr, degree ? &d : nullptr));
    if (gt) std::copy(g.begin(), g.end(), gt);
    if (degree) std::copy(d.begin(), d.end(), degree);
    const std::vector<i64> sz = lfr_sizes(p);
    *ncomm = (i64)sz.size();
    if (sizes) std::copy(sz.begin(), sz.begin() + std::min<i64>(sizes_cap, (i64)sz.size()), sizes);
    return out;
}
// the integer power-law table: k0 and up to cap thresholds; returns the count
i64 or_power_law_table(double tau, i32 lo, i32 top, double mean, i32* k0, u64* thr, i64 cap) {
    const PLTable t = power_law_table(tau, lo, top, mean);
    *k0 = t.k0;
    std::copy(t.thr.begin(), t.thr.begin() + std::min<i64>(cap, (i64)t.thr.size()), thr);
    return (i64)t.thr.size();
}
void or_graph_info(void* h, i64* n, i64* nnz, int* weighted) {
    const Graph* g = (const Graph*)h;
    *n = g->n;
    *nnz = g->nnz();
    *weighted = g->uniform ? 0 : 1;
}
void or_graph_copy(void* h, i64* rp, i32* col, i32* w) {
    const Graph* g = (const Graph*)h;
    std::copy(g->rp.begin(), g->rp.end(), rp);
    std::copy(g->col.begin(), g->col.end(), col);
    if (w && !g->uniform)
        for (i64 j = 0; j < g->nnz(); j++) w[j] = (i32)g->w[j];
}
void or_graph_free(void* h) { delete (Graph*)h; }

// labels[i] = Philox(counter (i lo, i hi, 0, stream 2), key seed) word 0 mod L
void or_random_labels(i64 n, i64 L, u64 seed, i32* out) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; i++) {
        P4 r = philox4x32_10(P4{{(u32)i, (u32)((u64)i >> 32), 0u, STREAM_LABEL}}, (u32)seed, (u32)(seed >> 32));
        out[i] = (i32)(r.v[0] % (u32)L);
    }
}

u32 or_murmur3_i32(i32 key) { return murmur3_i32(key); }
u32 or_xorshift32(u32 x) { return xorshift32(x); }
u32 or_seed_draw(u64 seed, u64 k) { return seed_draw(seed, k); }

// A1: order1/order2 plus offsets {large, large2, mid, mid2}
void or_orderings(i64 n, const i64* rp, i32* o1, i32* o2, i64* offs) {
    Graph g;
    g.n = n;
    g.rp.assign(rp, rp + n + 1);
    Orders o = make_orders(g);
    std::copy(o.o1.begin(), o.o1.end(), o1);
    std::copy(o.o2.begin(), o.o2.end(), o2);
    offs[0] = o.off_large; offs[1] = o.off_large2; offs[2] = o.off_mid; offs[3] = o.off_mid2;
}

// A12: vol[L], uncut, exact sum of squares as (hi, lo) u64 words
void or_evaluate(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* vw, const i32* labels, i64 L,
                 i64* vol_out, i64* uncut_out, u64* sq_out) {
    Graph g = make_graph(n, rp, col, w32, nullptr, vw);
    std::vector<i64> vol;
    i64 uncut;
    evaluate(g, labels, L, vol, uncut);
    u128 sq = 0;
    for (i64 l = 0; l < L; l++) sq += (u128)(u64)vol[l] * (u128)(u64)vol[l];
    std::copy(vol.begin(), vol.end(), vol_out);
    *uncut_out = uncut;
    sq_out[0] = (u64)(sq >> 64);
    sq_out[1] = (u64)sq;
}

// A10: returns coarse nnz; fills out arrays if capacity suffices (else -needed)
i64 or_contract(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i32* vcmap, i64 nc,
                i64* out_rp, i32* out_col, i64* out_w, i64 cap) {
    Graph g = make_graph(n, rp, col, w32, w64, nullptr);
    std::vector<i32> m(vcmap, vcmap + n);
    Graph c = contract(g, m, nc);
    if (c.nnz() > cap) return -c.nnz();
    std::copy(c.rp.begin(), c.rp.end(), out_rp);
    std::copy(c.col.begin(), c.col.end(), out_col);
    std::copy(c.w.begin(), c.w.end(), out_w);
    return c.nnz();
}

// One sweep from a fixed state (candidates + afterburner) with the dest cache
// in/out: returns the number of accepted moves, written (order1 order) to
// moves_out; dest/gain are updated in place.  raw != 0 scores the raw
// adjacency (the first sweep of an initial call).
i64 or_sweep_detail(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                    const i32* part, const i64* vol, i64 L, double lambda, float ratio, int raw, const i32* constraint,
                    i32* dest, float* gain, i32* moves_out, i64* ncand_out, i32* cands_out, float* change_out);

i64 or_sweep(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw, const i32* part,
             const i64* vol, i64 L, double lambda, float ratio, int raw, const i32* constraint, i32* dest, float* gain,
             i32* moves_out, i64* ncand_out) {
    return or_sweep_detail(n, rp, col, w32, w64, vw, part, vol, L, lambda, ratio, raw, constraint, dest, gain,
                           moves_out, ncand_out, nullptr, nullptr);
}

// or_sweep, also returning the candidate list (order1 order) and each
// candidate's afterburner change (the sum over its priority neighbours,
// src/local_mover.cpp:314-331) in cands_out / change_out (n entries each)
i64 or_sweep_detail(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                    const i32* part, const i64* vol, i64 L, double lambda, float ratio, int raw, const i32* constraint,
                    i32* dest, float* gain, i32* moves_out, i64* ncand_out, i32* cands_out, float* change_out) {
    Graph g = make_graph(n, rp, col, w32, w64, vw);
    State st;
    st.vol.assign(vol, vol + L);
    st.label_count = (i32)L;
    st.lambda = lambda;
    Mover M(g);
    M.part.assign(part, part + n);
    M.dest.assign(dest, dest + n);
    M.gain.assign(gain, gain + n);
    M.tables = !raw;
    std::vector<i32> c = candidates(M, st, ratio, constraint != nullptr, constraint);
    std::vector<float> ch;
    std::vector<i32> mv = c.empty() ? std::vector<i32>() : afterburner(M, st, c, &ch);
    if (cands_out) std::copy(c.begin(), c.end(), cands_out);
    if (change_out) std::copy(ch.begin(), ch.end(), change_out);
    std::copy(M.dest.begin(), M.dest.end(), dest);
    std::copy(M.gain.begin(), M.gain.end(), gain);
    std::copy(mv.begin(), mv.end(), moves_out);
    *ncand_out = (i64)c.size();
    return (i64)mv.size();
}

// own-cluster weight per vertex
void or_own_weight(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i32* labels,
                   i64* out) {
    Graph g = make_graph(n, rp, col, w32, w64, nullptr);
    std::vector<i32> l(labels, labels + n);
    std::vector<i64> pv = own_weight(g, l);
    std::copy(pv.begin(), pv.end(), out);
}

// objective set-up: app/objective_helpers.hpp:9-29, interface/cluster_data.hpp
// kind: 0 Mod, 1 WMod, 2 CPM, 3 NLCC, 4 LCC (base)
static State make_state(const Graph& g, int kind, double mult) {
    State s;
    s.vol.assign(g.vw.begin(), g.vw.end());
    s.top_nnz = g.nnz();
    s.label_count = (i32)g.n;
    s.uncut = 0;
    s.lambda = 1.0;
    if (kind == 0 || kind == 1) {
        i64 gd = 0;
        if (kind == 0 || g.uniform) gd = g.nnz();
        else for (i64 v : g.vw) gd += v;
        double inv = 1.0 / (double)gd;
        s.lambda = mult * inv;
        s.scale = inv;
    } else if (kind == 2) {
        s.lambda = mult * 0.5;
        s.offset = s.lambda * (double)g.n;
    } else if (kind == 3) {
        i64 gd = 0;
        if (g.uniform) gd = g.nnz();
        else for (i64 v : g.w) gd += v;
        u64 vt = 0;
        for (i64 v : g.vw) vt += (u64)v;
        double denom = (double)(vt * vt);
        s.lambda = mult * (double)gd / denom;
        s.scale = 1.0 / (double)gd;
    } else {
        s.lambda = mult;
    }
    s.update_objective();
    return s;
}

// Full clustering as app/cluster_main.cpp:116-165 (run_clustering) for one
// trial.  algo: 0 pLouvain, 1 pLeiden, 2 pLeiden+.  Returns the objective;
// part_out[n]; info[0..5] = label_count, uncut, levels, sweeps, moves, outer
// iterations.
double or_cluster(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* vw, int kind, double mult,
                  int algo, int ex_iters, u64 seed, i32* part_out, i64* info) {
    Graph g = make_graph(n, rp, col, w32, nullptr, vw);
    State rfd = make_state(g, kind, mult);
    Rng rng{seed};
    Run run;
    std::vector<i32> part;
    if (algo == 0) part = louvain_part(g, rfd, false, {}, run);
    else part = leiden_part(g, rfd, algo == 2, false, nullptr, rng, run);
    double last = -std::numeric_limits<double>::infinity();
    int counter = 0;
    i64 outer = 1;
    while (true) {  // iteration_termination_condition (app/cluster_main.cpp:101-114)
        bool term = true;
        if (ex_iters < 0) {
            if (rfd.objective() > last) term = false;
            last = rfd.objective();
        } else if (counter++ < ex_iters) {
            term = false;
        }
        if (term) break;
        std::vector<i32> cons = part;
        if (algo == 0) part = louvain_part(g, rfd, true, cons, run);
        else part = leiden_part(g, rfd, algo == 2, true, &cons, rng, run);
        outer++;
    }
    std::copy(part.begin(), part.end(), part_out);
    info[0] = rfd.label_count;
    info[1] = rfd.uncut;
    info[2] = run.levels;
    info[3] = run.stats.sweeps;
    info[4] = run.stats.moves;
    info[5] = outer;
    return rfd.objective();
}

// local_move on one level from a given state (for parity of A8 on its own)
// vol_inout sized >= n; returns label_count; part_inout relabelled.
i64 or_local_move_trace(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                        i32* part_inout, i64* vol_inout, i64 L, i64* uncut_inout, double lambda, int is_initial,
                        int strict, int SA, double* obj_inout, double* trace_out, i64 trace_cap, i64* ntrace);

i64 or_local_move(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                  i32* part_inout, i64* vol_inout, i64 L, i64* uncut_inout, double lambda, int is_initial,
                  int strict, int SA, double* obj_inout) {
    return or_local_move_trace(n, rp, col, w32, w64, vw, part_inout, vol_inout, L, uncut_inout, lambda, is_initial,
                               strict, SA, obj_inout, nullptr, 0, nullptr);
}

// or_local_move with its per-sweep records: trace_out[5*k..5*k+4] = ratio,
// candidates, accepted moves, objective after sweep k and its wall seconds
// (at most trace_cap records)
i64 or_local_move_trace(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                        i32* part_inout, i64* vol_inout, i64 L, i64* uncut_inout, double lambda, int is_initial,
                        int strict, int SA, double* obj_inout, double* trace_out, i64 trace_cap, i64* ntrace) {
    std::vector<SweepRec> tr;
    g_trace = trace_out ? &tr : nullptr;
    Graph g = make_graph(n, rp, col, w32, w64, vw);
    State st;
    st.vol.assign(vol_inout, vol_inout + n);
    st.label_count = (i32)L;
    st.lambda = lambda;
    st.uncut = *uncut_inout;
    st.top_nnz = g.nnz();
    st.obj = *obj_inout;
    std::vector<i32> part(part_inout, part_inout + n);
    if (strict) local_move_strict(g, part, st, is_initial != 0, false, nullptr, nullptr);
    else local_move(g, part, st, is_initial != 0, false, nullptr, SA != 0, nullptr);
    g_trace = nullptr;
    if (trace_out) {
        const i64 k = std::min<i64>((i64)tr.size(), trace_cap);
        for (i64 x = 0; x < k; x++) {
            trace_out[5 * x] = tr[x].ratio;
            trace_out[5 * x + 1] = (double)tr[x].cands;
            trace_out[5 * x + 2] = (double)tr[x].moves;
            trace_out[5 * x + 3] = tr[x].obj;
            trace_out[5 * x + 4] = tr[x].seconds;
        }
        *ntrace = (i64)tr.size();
    }
    std::copy(part.begin(), part.end(), part_inout);
    std::copy(st.vol.begin(), st.vol.begin() + st.label_count, vol_inout);
    *uncut_inout = st.uncut;
    *obj_inout = st.obj;
    return st.label_count;
}

// leidenR on one level: louv = Louvain labels (constraint), vol[L] its volumes
i64 or_leiden_refine_detail(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                            const i32* louv, const i64* vol, i64 L, double lambda, int uniform, u64 seed, u64 draw0,
                            i32* vcmap_out, i32* trees_out, i32* order_out);

i64 or_leiden_refine(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                     const i32* louv, const i64* vol, i64 L, double lambda, int uniform, u64 seed, u64 draw0,
                     i32* vcmap_out) {
    return or_leiden_refine_detail(n, rp, col, w32, w64, vw, louv, vol, L, lambda, uniform, seed, draw0, vcmap_out,
                                   nullptr, nullptr);
}

// or_leiden_refine, also returning the forest before the gamma breaks
// (trees_out[i] = root of i's tree, find_trees, src/leidenR.cpp:212-260) and
// the order it is sorted by (order_out, after find_trees' flip)
i64 or_leiden_refine_detail(i64 n, const i64* rp, const i32* col, const i32* w32, const i64* w64, const i64* vw,
                            const i32* louv, const i64* vol, i64 L, double lambda, int uniform, u64 seed, u64 draw0,
                            i32* vcmap_out, i32* trees_out, i32* order_out) {
    RefineDetail det;
    g_refine = trees_out ? &det : nullptr;
    Graph g = make_graph(n, rp, col, w32, w64, vw);
    State st;
    st.vol.assign(vol, vol + L);
    st.label_count = (i32)L;
    st.lambda = lambda;
    std::vector<i32> lv(louv, louv + n);
    std::vector<i64> pv = own_weight(g, lv);
    Rng rng{seed, draw0};
    i64 nc = 0;
    std::vector<i32> m = leiden_refine(g, lv, st, pv, uniform != 0, rng, nc);
    g_refine = nullptr;
    if (trees_out) {
        std::copy(det.trees.begin(), det.trees.end(), trees_out);
        std::copy(det.order.begin(), det.order.end(), order_out);
    }
    std::copy(m.begin(), m.end(), vcmap_out);
    return nc;
}

}  // extern "C"
