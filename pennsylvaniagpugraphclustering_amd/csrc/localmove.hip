// Local mover: degree-classed neighbour-cluster aggregation with gain argmax,
// the afterburner conflict filter, move application with exact uncut deltas,
// the strict prefix filter, relabelling, and the local_move drivers.
//
// Determinism (shared with the CPU oracle):
//  * destination = max gain, ties -> smallest label; NEW_PART (-4) competes
//    with j_conn 0 when p_conn < 0 (team-path rule, src/local_mover.cpp:553-561)
//  * afterburner FP32 sums: degree < 128 sequential in adjacency order,
//    128..14999 64 strided partials + halving tree, >= 15000 1024 strided
//    partials + halving tree
//  * uncut is updated by the exact integer delta of the moved vertices
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <chrono>

#include "device.h"
#include "gnn.h"
#include "hash.h"
#include "localmove.h"

namespace pggc {

// ------------------------------------------------------------- objective
double objective_value(i64 uncut, double lambda, u64 sq_hi, u64 sq_lo) {
    unsigned __int128 sq = ((unsigned __int128)sq_hi << 64) | (unsigned __int128)sq_lo;
    double m = (double)uncut;
    m -= lambda * (double)sq;
    return m;
}

void update_objective(Ctx& c, RefineData& rd) {
    Buf<i64> sq(c, 2);
    sum_sq_u128(c, rd.vol, rd.label_count, sq);
    const i64* h = c.read(sq, 2);
    rd.obj = objective_value(rd.uncut, rd.lambda, (u64)h[0], (u64)h[1]);
}

// ------------------------------------------------------------- scratch
void MoveScratch::ensure(Ctx& c, i64 n) {
    if (cap >= n) return;
    cap = std::max<i64>(n, 1);
    part = Buf<i32>(c, cap);
    dest = Buf<i32>(c, cap);
    mflag = Buf<i32>(c, cap);
    cand = Buf<i32>(c, cap);
    moves = Buf<i32>(c, cap);
    flag = Buf<i32>(c, cap);
    tmp = Buf<i32>(c, cap);
    tmp2 = Buf<i32>(c, cap + 1);
    gain = Buf<float>(c, cap);
    abi = Buf<AbInfo>(c, cap);
    mvp = Buf<u64>(c, cap);
    prio = Buf<i64>(c, cap + 1);
    scal = Buf<i64>(c, 16);
    fill_i32(c, mflag, cap, 0);
}

// ------------------------------------------------------------- level plan
__host__ __device__ __forceinline__ int score_class(i64 d) {
    return d <= 4 ? 0 : d <= 8 ? 1 : d <= 16 ? 2 : d <= 32 ? 3 : d <= 64 ? 4 : d <= 128 ? 5 : d <= 256 ? 6
         : d <= 512 ? 7 : d <= 1024 ? 8 : d <= 2048 ? 9 : 10;
}
// LDS table of a vertex of degree d: the power of two >= spe * d slots
__host__ __device__ __forceinline__ i64 tslots(i64 d, int spe) {
    i64 s = 2;
    while (s < spe * d) s <<= 1;
    return s;
}
// vertices of the hub class scored from an LDS table (degree <= HUB_LDS_BOUND)
constexpr i64 HUB_LDS_BOUND = 4096;
// HBM hub tables: power of two >= 1.5 d, so at most 2/3 full (a hub has at
// most d distinct neighbour clusters); every probe loop ends at an empty slot
__host__ __device__ __forceinline__ i64 hub_slots(i64 d) {
    i64 s = 2;
    while (2 * s < 3 * d) s <<= 1;
    return s;
}

__global__ void k_score_classes(const i64* rp, i64 n, i32* cls) {
    const i64 i = thread_item();
    if (i >= n) return;
    cls[i] = score_class(rp[i + 1] - rp[i]);
}
__global__ void k_global_sizes(const i64* rp, const i32* list, i64 m, i64 he, i64* sz, i64* items) {
    const i64 x = thread_item();
    if (x >= m) return;
    i32 i = list[x];
    i64 d = rp[i + 1] - rp[i];
    sz[x] = hub_slots(d);
    items[x] = (d + he - 1) / he;
}
__global__ void k_slot_items(const i64* goff, i64 m, i64 hs, i64* sitems) {
    const i64 x = thread_item();
    if (x >= m) return;
    sitems[x] = (goff[x + 1] - goff[x] + hs - 1) / hs;
}
// Slots per hub-table slice: one slice is built in LDS by one block, 128 KB
// of LDS for packed u64 slots, 96 KB for separate i32 keys and i64 weights.
constexpr i64 HUB_SLICE_PACKED = 16384, HUB_SLICE_WIDE = 8192;
constexpr i64 HUB_MAX_SLICES = 16;  // larger tables are built by the fallback
__host__ __device__ __forceinline__ i64 hub_slice(bool packed) { return packed ? HUB_SLICE_PACKED : HUB_SLICE_WIDE; }
__global__ void k_slice_counts(const i64* goff, i64 m, i64 sl, i64* soff) {
    const i64 x = thread_item();
    if (x >= m) return;
    const i64 S = goff[x + 1] - goff[x];
    soff[x] = S > sl ? S / sl : 1;
}
__global__ void k_hub_deg_keys(const i64* rp, const i32* list, i64 m, u64* keys, i32* vals) {
    const i64 x = thread_item();
    if (x >= m) return;
    const i32 i = list[x];
    keys[x] = (u64)(rp[i + 1] - rp[i]);
    vals[x] = i;
}
// first x with degree(list[x]) > bound (list sorted by degree)
__global__ void k_first_above(const i64* rp, const i32* list, i64 m, i64 bound, i64* out) {
    if (threadIdx.x != 0) return;
    i64 lo = 0, hi = m;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        const i32 i = list[mid];
        if (rp[i + 1] - rp[i] > bound) hi = mid;
        else lo = mid + 1;
    }
    *out = lo;
}

// row weight of each hub (one block per edge item)
template <class WT>
__global__ void __launch_bounds__(256) k_hub_wsum(const i64* __restrict__ rp, WT W, const i32* __restrict__ hubs,
                                                  i64 m, const i64* __restrict__ items, i64 he, i64* hsum) {
    const i64 b = block_id();
    if (b >= items[m]) return;
    const i64 hi = upper_row(items, 0, m, b);
    const i32 i = hubs[hi];
    const i64 e = rp[i + 1];
    const i64 j_lo = rp[i] + (b - items[hi]) * he;
    const i64 j_hi = j_lo + he < e ? j_lo + he : e;
    i64 acc = 0;
    for (i64 jb = j_lo; jb < j_hi; jb += 256) {
        const i64 j = jb + threadIdx.x;
        if (j < j_hi) acc += W(j);
    }
    acc = block_sum_i64<256>(acc);
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)&hsum[hi], (unsigned long long)acc);
}

// Block -> work-item maps that keep each hub on one XCD: hubs are assigned
// longest-first to the least loaded of the 8 XCDs (load = insert items + scan
// items); block 8k + x runs the k-th item of XCD x.
constexpr int NXCD = 8;
static void build_map(const std::vector<i64>& pre, const std::vector<int>& xcd_of, i64 m, std::vector<i32>& map,
                      i64& grid, i64 lo = 0, bool largest_first = false) {
    std::vector<std::vector<i32>> per(NXCD);
    for (i64 k = 0; k < m; k++) {
        const i64 h = largest_first ? m - 1 - k : k;  // hubs are sorted by degree
        for (i64 it = std::max(pre[h], lo); it < pre[h + 1]; it++) per[xcd_of[h]].push_back((i32)it);
    }
    size_t mx = 0;
    for (auto& v : per) mx = std::max(mx, v.size());
    grid = (i64)mx * NXCD;
    map.assign(std::max<i64>(grid, 1), -1);
    for (int x = 0; x < NXCD; x++)
        for (size_t k = 0; k < per[x].size(); k++) map[k * NXCD + x] = per[x][k];
}

static void xcd_maps(Ctx& c, LevelPlan& p, i64 m) {
    std::vector<i64> it(m + 1), st(m + 1), so(m + 1), go(m + 1);
    PGGC_HIP(hipMemcpyAsync(go.data(), p.goff, (m + 1) * sizeof(i64), hipMemcpyDeviceToHost, c.stream));
    PGGC_HIP(hipMemcpyAsync(it.data(), p.items, (m + 1) * sizeof(i64), hipMemcpyDeviceToHost, c.stream));
    PGGC_HIP(hipMemcpyAsync(st.data(), p.sitems, (m + 1) * sizeof(i64), hipMemcpyDeviceToHost, c.stream));
    PGGC_HIP(hipMemcpyAsync(so.data(), p.soff, (m + 1) * sizeof(i64), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::vector<i64> order(m);
    for (i64 h = 0; h < m; h++) order[h] = h;
    auto load = [&](i64 h) { return (it[h + 1] - it[h]) + (st[h + 1] - st[h]); };
    std::stable_sort(order.begin(), order.end(), [&](i64 a, i64 b) { return load(a) > load(b); });
    std::vector<i64> tot(NXCD, 0);
    std::vector<int> xcd_of(m, 0);
    for (i64 h : order) {
        int best = 0;
        for (int x = 1; x < NXCD; x++)
            if (tot[x] < tot[best]) best = x;
        xcd_of[h] = best;
        tot[best] += load(h);
    }
    std::vector<i32> imap, smap;
    build_map(it, xcd_of, m, imap, p.igrid);
    p.imap = Buf<i32>(c, imap.size());
    copy_h2d(c, p.imap, imap.data(), imap.size() * sizeof(i32));
    build_map(st, xcd_of, m, smap, p.sgrid);
    p.smap = Buf<i32>(c, smap.size());
    copy_h2d(c, p.smap, smap.data(), smap.size() * sizeof(i32));
    // slice classes (hubs are sorted by degree, so table sizes ascend):
    // [0, hs1) tables of at most 4096 slots, [hs1, hs2) packed tables of
    // 8192, then the tables of more slots than one block's slice class
    // (several slices each), which go through the XCD map
    auto first_above = [&](i64 slots) {
        i64 h = 0;
        while (h < m && go[h + 1] - go[h] <= slots) h++;
        return so[h];
    };
    p.hs1 = first_above(4096);
    p.hs2 = p.packed ? first_above(8192) : p.hs1;
    // the largest hubs (more than HUB_MAX_SLICES slices) are built by the
    // fallback: first such hub, its first table slot and first build item
    p.hg = 0;
    while (p.hg < m && go[p.hg + 1] - go[p.hg] <= HUB_MAX_SLICES * hub_slice(p.packed)) p.hg++;
    p.ghg = go[p.hg];
    p.ihg = it[p.hg];
    p.shg = so[p.hg];
    std::vector<i32> bmap;
    build_map(so, xcd_of, m, bmap, p.bgrid, p.hs2, true);
    if (p.hs2 == so[m]) p.bgrid = 0;
    p.bmap = Buf<i32>(c, bmap.size());
    copy_h2d(c, p.bmap, bmap.data(), bmap.size() * sizeof(i32));
}

__global__ void k_hub_index(const i32* hubs, i64 m, i32* hidx) {
    const i64 x = thread_item();
    if (x >= m) return;
    hidx[hubs[x]] = (i32)x;
}

void plan_level(Ctx& c, const DGraph& g, LevelPlan& p) {
    if (p.key == (const void*)g.rp && p.n == g.n && p.epoch == c.knob_epoch) return;
    p.key = g.rp;
    p.n = g.n;
    p.epoch = c.knob_epoch;
    p.he = c.hub_edges_item;
    p.hs = c.hub_slots_item;
    make_orders(c, g, p.ord);
    Buf<i32> cls(c, std::max<i64>(g.n, 1));
    p.sc_list = Buf<i32>(c, std::max<i64>(g.n, 1));
    k_score_classes<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.rp, g.n, cls);
    PGGC_LAUNCH_CHECK();
    partition_classes(c, cls, g.n, NCLS, p.sc_list, p.sc_start);
    // hub class by degree (stable: ties by id), so that the hubs scored in
    // LDS form a prefix; every hub is scored on its own, so the order within
    // the class does not change results
    const i64 nh = p.sc_start[CL_HUB + 1] - p.sc_start[CL_HUB];
    p.nlds = 0;
    if (nh > 1) {
        i32* hl = p.sc_list.p + p.sc_start[CL_HUB];
        Buf<u64> hk(c, nh);
        k_hub_deg_keys<<<grid_items(nh, 256), 256, 0, c.stream>>>(g.rp, hl, nh, hk, hl);
        PGGC_LAUNCH_CHECK();
        radix_sort_pairs(c, hk, hl, nh, 40);
    }
    if (nh > 0 && c.hub_lds_max > 2048) {
        Buf<i64> f(c, 1);
        k_first_above<<<1, 64, 0, c.stream>>>(g.rp, p.sc_list.p + p.sc_start[CL_HUB], nh,
                                               std::min<i64>(c.hub_lds_max, HUB_LDS_BOUND), f);
        PGGC_LAUNCH_CHECK();
        p.nlds = c.read1(f);
    }
    p.hub0 = p.sc_start[CL_HUB] + p.nlds;
    p.nhub = nh - p.nlds;
    const i64 m7 = p.nhub;
    p.gslots = 0;
    if (m7 > 0) {
        p.goff = Buf<i64>(c, m7 + 1);
        p.items = Buf<i64>(c, m7 + 1);
        p.sitems = Buf<i64>(c, m7 + 1);
        p.hub_pv = Buf<i64>(c, m7);
        p.hub_best = Buf<u64>(c, m7);
        k_global_sizes<<<grid_items(m7, 256), 256, 0, c.stream>>>(g.rp, p.sc_list.p + p.hub0, m7, p.he, p.goff,
                                                                        p.items);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, p.goff, p.goff, m7);
        scan_i64(c, p.items, p.items, m7);
        k_slot_items<<<grid_items(m7, 256), 256, 0, c.stream>>>(p.goff, m7, p.hs, p.sitems);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, p.sitems, p.sitems, m7);
        p.gslots = c.read1(p.goff.p + m7);
        p.nitems = c.read1(p.items.p + m7);
        p.nsitems = c.read1(p.sitems.p + m7);
        const i32* hubs = p.sc_list.p + p.hub0;
        // packed slots need every hub row weight below 2^32
        p.packed = true;
        if (g.wk != W_UNIT) {
            Buf<i64> hsum(c, m7);
            fill_i64(c, hsum, m7, 0);
            PGGC_WDISPATCH(g, (k_hub_wsum<<<grid_blocks(p.nitems), 256, 0, c.stream>>>(g.rp, W, hubs, m7, p.items,
                                                                                    p.he, hsum)));
            PGGC_LAUNCH_CHECK();
            std::vector<i64> hs(m7);
            PGGC_HIP(hipMemcpyAsync(hs.data(), hsum, m7 * sizeof(i64), hipMemcpyDeviceToHost, c.stream));
            c.sync();
            for (i64 x = 0; x < m7; x++)
                if (hs[x] >= ((i64)1 << 32) - 1) p.packed = false;
        }
        // the tables themselves (and their slot lists) are allocated by
        // hub_tables_acquire when a local move starts
        p.gtab.reset();
        p.gkeys.reset();
        p.gvals.reset();
        p.hlist.reset();
        p.soff = Buf<i64>(c, m7 + 1);
        k_slice_counts<<<grid_items(m7, 256), 256, 0, c.stream>>>(p.goff, m7, hub_slice(p.packed), p.soff);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, p.soff, p.soff, m7);
        p.nslices = c.read1(p.soff.p + m7);
        p.hflag = Buf<i32>(c, m7);
        p.hovf = Buf<i32>(c, 1);
        p.hcnt = Buf<i32>(c, m7);
        p.hsm = Buf<u32>(c, m7);
        xcd_maps(c, p, m7);
        p.hidx = Buf<i32>(c, std::max<i64>(g.n, 1));
        fill_i32(c, p.hidx, g.n, -1);
        k_hub_index<<<grid_items(m7, 256), 256, 0, c.stream>>>(hubs, m7, p.hidx);
        PGGC_LAUNCH_CHECK();
    }
    p.hub_live = false;
}

// Hub tables and slot lists, the bulk of a plan (about 1.5-3 slots per hub
// edge), exist only while a local move of the level runs; they are rebuilt
// at the first table sweep of every call anyway.  Released at the end of
// the call (to the context's cache, which usually hands the same block back
// at the level's next call), they leave the room to the contraction and the
// other levels of the hierarchy on graphs of billions of edges.
static void hub_tables_acquire(Ctx& c, LevelPlan& p) {
    if (p.nhub == 0 || p.gslots == 0) return;
    if (p.packed) {
        if (!p.gtab.p) p.gtab = Buf<u64>(c, p.gslots);
    } else if (!p.gkeys.p) {
        p.gkeys = Buf<i32>(c, p.gslots);
        p.gvals = Buf<i64>(c, p.gslots);
    }
    if (!p.hlist.p) p.hlist = Buf<u32>(c, p.gslots);
}
static void hub_tables_settle(Ctx& c, LevelPlan& p) {
    (void)c;
    p.gtab.reset();
    p.gkeys.reset();
    p.gvals.reset();
    p.hlist.reset();
    p.hub_live = false;
}

// ------------------------------------------------------------- scoring
__device__ __forceinline__ u64 pack_cand(float jc, i32 key) {
    return ((u64)ord_f32(jc) << 32) | (u64)(0xFFFFFFFFu - (u32)(key + 4));
}

__device__ __forceinline__ void decide(i32 i, float p_conn, u64 bk, i32* dest, float* gain) {
    bool have = bk != 0;
    float bj = have ? unord_f32((u32)(bk >> 32)) : 0.0f;
    i32 bc = have ? (i32)(0xFFFFFFFFu - (u32)(bk & 0xffffffffull)) - 4 : 0;
    if (p_conn < 0.0f && (!have || !(bj > 0.0f))) {
        gain[i] = 0.0f - p_conn;
        dest[i] = NEW_PART;
    } else if (have) {
        gain[i] = bj - p_conn;
        dest[i] = bc;
    } else {
        gain[i] = -FLT_MAX;
        dest[i] = NO_MOVE;
    }
}

// lambda of a vertex's objective and the sweep's filter ratio: one value
// each for a single graph, per-vertex tables for a batch of graphs held as
// one disjoint union (each vertex carries its instance's values), whose
// instances can be at different sweeps of their ratio schedules.  The batch
// tables are filled per sweep by k_vertex_lam rather than looked up as
// lam[inst[v]] inside the scoring kernels: that two-step lookup there gave
// wrong values in about 1% of the lanes on gfx950, run to run different,
// while the same lookup in its own kernel and the per-vertex tables were
// exact.
struct LamOne {
    float v;
    __device__ __forceinline__ float operator()(i64) const { return v; }
    __device__ __forceinline__ float rt(i64, float r) const { return r; }
};
struct LamVert {
    const float* lam;    // lambda of vertex i's instance
    const float* ratio;  // filter ratio of vertex i's instance this sweep
    __device__ __forceinline__ float operator()(i64 i) const { return lam[i]; }
    __device__ __forceinline__ float rt(i64 i, float) const { return ratio[i]; }
};
// afterburner priority hash of a vertex (src/local_mover.cpp:319): xorshift32
// of its id, or of its id within its instance for a batch held as one union
// (loc[v] = v - first vertex of v's instance)
struct PrioOne {
    __device__ __forceinline__ i32 operator()(i32 v) const { return (i32)xorshift32((u32)v); }
};
struct PrioInst {
    const i32* loc;
    __device__ __forceinline__ i32 operator()(i32 v) const { return (i32)xorshift32((u32)loc[v]); }
};

// p_conn and the filter threshold (src/local_mover.cpp:481-487)
__device__ __forceinline__ void own_terms(i64 pv, i64 volp, i64 wv, float lam, float ratio, float& multi,
                                          float& p_conn, float& b) {
    multi = (float)wv * lam;
    p_conn = fmaf(-(float)(volp - wv), multi, (float)pv);
    b = (float)((double)p_conn * (1.0 - (double)ratio));
    if (p_conn < 0.0f) b = 0.0f;
}

// G lanes per vertex, one edge per lane (degree <= G): duplicate clusters
// are merged with in-group shuffles.
template <int G, class WT, class LT>
__global__ void __launch_bounds__(256) k_score_group(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                     const i64* __restrict__ vw, const i32* __restrict__ list, i64 m,
                                                     const i32* __restrict__ part, const i64* __restrict__ vol,
                                                     const i32* __restrict__ cons, int raw, LT lam, float ratio,
                                                     i32* __restrict__ dest, float* __restrict__ gain) {
    constexpr int GPB = 256 / G;
    const int gl = threadIdx.x & (G - 1);
    const i64 gi = block_id() * GPB + threadIdx.x / G;
    if (gi >= m) return;
    {
        const i32 i = list[gi];
        if (dest[i] == NO_MOVE) return;
        const i64 s = rp[i];
        const int d = (int)(rp[i + 1] - s);
        const i32 p = part[i];
        const i64 wv = vw[i];
        i32 key = -5;
        i64 wt = 0;
        const bool has = gl < d;
        if (has) {
            i32 v = col[s + gl];
            key = raw == 1 ? v : part[v];
            wt = W(s + gl);
        }
        i64 agg = wt;
        bool own = false;
        i64 pv = 0;
        if (!raw) {  // raw rows have distinct columns and no own-cluster edge
            agg = 0;
            xlane_fence(key);
            xlane_fence(wt);
#pragma unroll
            for (int k = 0; k < G; k++) {
                i32 kk = __shfl(key, k, G);
                i64 ww = __shfl(wt, k, G);
                agg += (kk == key) ? ww : 0;
            }
            own = has && key == p;
            pv = group_max_i64<G>(own ? agg : (i64)0);
        }
        float multi, p_conn, b;
        own_terms(pv, vol[p], wv, lam(i), lam.rt(i, ratio), multi, p_conn, b);
        u64 bk = 0;
        if (has && !own && agg > 0 && (float)agg >= b && (cons == nullptr || cons[key] == cons[p])) {
            float jc = fmaf(-(float)vol[key], multi, (float)agg);
            if (jc >= b) bk = pack_cand(jc, key);
        }
        bk = group_max_u64<G>(bk);
        if (gl == 0) decide(i, p_conn, bk, dest, gain);
    }
}

// One block per vertex with a hash table in LDS (or HBM for the largest
// degrees): SLOTS slots of LDS hold the table of any degree up to
// SLOTS / SPE, each table the power of two >= SPE * degree.
template <int BS, int SLOTS, int SPE, bool GLOBAL, class WT, class LT>
__global__ void __launch_bounds__(BS) k_score_table(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                    const i64* __restrict__ vw, const i32* __restrict__ list, i64 m,
                                                    const i32* __restrict__ part, const i64* __restrict__ vol,
                                                    const i32* __restrict__ cons, int raw, LT lam, float ratio,
                                                    i32* __restrict__ dest, float* __restrict__ gain, i32* gkeys,
                                                    i64* gvals, const i64* __restrict__ goff) {
    __shared__ i32 skeys[GLOBAL ? 1 : SLOTS];
    __shared__ i64 svals[GLOBAL ? 1 : SLOTS];
    const int t = threadIdx.x;
    const i64 li = block_id();
    if (li >= m) return;
    {
        const i32 i = list[li];
        if (dest[i] == NO_MOVE) return;
        const i64 s = rp[i];
        const i64 d = rp[i + 1] - s;
        const i64 S = tslots(d, SPE);
        const u32 mask = (u32)(S - 1);
        i32* keys = GLOBAL ? gkeys + goff[li] : skeys;
        i64* vals = GLOBAL ? gvals + goff[li] : svals;
        for (i64 q = t; q < S; q += BS) {
            keys[q] = EMPTY_KEY;
            vals[q] = 0;
        }
        if (GLOBAL) __threadfence();
        __syncthreads();
        const i32 p = part[i];
        i64 pv = 0;
        // LDS classes hold d <= SLOTS / SPE: a thread first gathers the
        // clusters of all its (at most R) edges, so the gathers are in flight
        // together, then inserts them
        constexpr int R = GLOBAL ? 0 : (SLOTS / SPE + BS - 1) / BS;
        i64 e_done = 0;
        if (!GLOBAL) {
            i32 kk[R > 0 ? R : 1];
            i64 wv[R > 0 ? R : 1];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const i64 e = (i64)r * BS + t;
                kk[r] = EMPTY_KEY;
                wv[r] = 0;
                if (e < d) {
                    const i32 v = col[s + e];
                    kk[r] = raw == 1 ? v : part[v];
                    wv[r] = W(s + e);
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                if (kk[r] == EMPTY_KEY) continue;
                if (!raw && kk[r] == p) pv += wv[r];
                else table_add_lds(keys, vals, mask, kk[r], wv[r]);
            }
            e_done = (i64)R * BS;
        }
        for (i64 eb = e_done; eb < d; eb += BS) {  // block-uniform loop control
            const i64 e = eb + t;
            if (e >= d) continue;
            i32 v = col[s + e];
            i32 key = raw == 1 ? v : part[v];
            i64 ww = W(s + e);
            if (!raw && key == p) pv += ww;
            else if (GLOBAL) table_add_global(keys, vals, mask, key, ww);
            else table_add_lds(keys, vals, mask, key, ww);
        }
        if (GLOBAL) __threadfence();
        pv = block_sum_i64<BS>(pv);
        float multi, p_conn, b;
        own_terms(pv, vol[p], vw[i], lam(i), lam.rt(i, ratio), multi, p_conn, b);
        u64 bk = 0;
        for (i64 qb = 0; qb < S; qb += BS) {  // block-uniform loop control
            const i64 q = qb + t;
            if (q >= S) continue;
            i32 key;
            i64 agg;
            if (GLOBAL) {
                key = __hip_atomic_load(&keys[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (key == EMPTY_KEY) continue;
                agg = __hip_atomic_load(&vals[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                key = keys[q];
                if (key == EMPTY_KEY) continue;
                agg = vals[q];
            }
            if (agg > 0 && (float)agg >= b && (cons == nullptr || cons[key] == cons[p])) {
                float jc = fmaf(-(float)vol[key], multi, (float)agg);
                if (jc >= b) {
                    u64 k2 = pack_cand(jc, key);
                    bk = k2 > bk ? k2 : bk;
                }
            }
        }
        bk = block_max_u64<BS>(bk);
        if (t == 0) decide(i, p_conn, bk, dest, gain);
        __syncthreads();
    }
}

// Raw sweep (first sweep of an initial local_move, src/local_mover.cpp:1391-1395):
// every neighbour is its own cluster with weight w_ij (rows carry no duplicate
// columns and no self loop), so the own-cluster weight is 0 and the best
// candidate is a streaming arg-max over the adjacency, with the same packed
// (gain, smallest label) order as the table path.
template <class WT>
__device__ __forceinline__ u64 raw_cand(const i32* __restrict__ col, WT W, const i64* __restrict__ vol,
                                        const i32* __restrict__ cons, const i32* __restrict__ lab, i32 cp, i64 j,
                                        float multi, float b) {
    const i32 v = lab ? lab[col[j]] : col[j];  // the neighbour's (singleton) cluster
    const i64 agg = W(j);
    if (agg > 0 && (float)agg >= b && (cons == nullptr || cons[v] == cp)) {
        const float jc = fmaf(-(float)vol[v], multi, (float)agg);
        if (jc >= b) return pack_cand(jc, v);
    }
    return 0;
}

// one block per vertex (64 < degree <= 2048)
template <int BS, class WT, class LT>
__global__ void __launch_bounds__(BS) k_score_raw(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                  const i64* __restrict__ vw, const i32* __restrict__ list, i64 m,
                                                  const i32* __restrict__ part, const i64* __restrict__ vol,
                                                  const i32* __restrict__ cons, const i32* __restrict__ lab, LT lam,
                                                  float ratio, i32* __restrict__ dest, float* __restrict__ gain) {
    const i64 li = block_id();
    if (li >= m) return;
    const i32 i = list[li];
    if (dest[i] == NO_MOVE) return;
    const i64 s = rp[i], e = rp[i + 1];
    const i32 p = part[i];
    const i32 cp = cons ? cons[p] : 0;
    float multi, p_conn, b;
    own_terms(0, vol[p], vw[i], lam(i), lam.rt(i, ratio), multi, p_conn, b);
    u64 bk = 0;
    for (i64 jb = s; jb < e; jb += BS) {  // block-uniform loop control
        const i64 j = jb + threadIdx.x;
        if (j < e) {
            const u64 k2 = raw_cand(col, W, vol, cons, lab, cp, j, multi, b);
            bk = k2 > bk ? k2 : bk;
        }
    }
    bk = block_max_u64<BS>(bk);
    if (threadIdx.x == 0) decide(i, p_conn, bk, dest, gain);
}

// hubs: one block per edge item, folded into hub_best with atomicMax
template <class WT, class LT>
__global__ void __launch_bounds__(256) k_hub_raw(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                 const i64* __restrict__ vw, const i32* __restrict__ hubs, i64 m,
                                                 const i64* __restrict__ items, const i32* __restrict__ part,
                                                 const i64* __restrict__ vol, const i32* __restrict__ cons,
                                                 const i32* __restrict__ lab, LT lam, float ratio,
                                                 const i32* __restrict__ dest, i64 he, u64* hub_best) {
    const i64 b = block_id();
    if (b >= items[m]) return;
    const i64 hi = upper_row(items, 0, m, b);
    const i32 i = hubs[hi];
    if (dest[i] == NO_MOVE) return;
    const i64 s = rp[i], e = rp[i + 1];
    const i64 j_lo = s + (b - items[hi]) * he;
    const i64 j_hi = j_lo + he < e ? j_lo + he : e;
    const i32 p = part[i];
    const i32 cp = cons ? cons[p] : 0;
    float multi, p_conn, bth;
    own_terms(0, vol[p], vw[i], lam(i), lam.rt(i, ratio), multi, p_conn, bth);
    u64 bk = 0;
    for (i64 jb = j_lo; jb < j_hi; jb += 256) {  // block-uniform loop control
        const i64 j = jb + threadIdx.x;
        if (j < j_hi) {
            const u64 k2 = raw_cand(col, W, vol, cons, lab, cp, j, multi, bth);
            bk = k2 > bk ? k2 : bk;
        }
    }
    bk = block_max_u64<256>(bk);
    if (threadIdx.x == 0 && bk) atomicMax((unsigned long long*)&hub_best[hi], (unsigned long long)bk);
}

// Hubs: an HBM table per hub of its neighbour-cluster weights (own cluster
// included), built at the first table sweep of a call and kept current from
// the movers by later sweeps (k_hub_update_*).  Per sweep, k_hub_pv reads
// the own-cluster weight, slot work items scan the table and fold the packed
// best candidate with atomicMax, and a last pass decides.
//
// Build: one block per table slice (k_hub_build) reads all of its hub's
// edges, keeps the (cluster, weight) pairs whose hash falls in its slice in
// an LDS table with the slice's probe sequence, and writes the slice out
// whole (empty slots included), so the build needs no fill, no HBM atomics
// and only coalesced stores.  The slices of one hub run on one XCD, where
// the hub's edges and neighbour clusters are read from L2 after the first
// slice.  A slice whose keys do not fit flags its hub (hflag), which is then
// cleared (k_hub_clear) and built by edge items with HBM atomics and
// whole-table probing (k_hub_insert).

// Slice masks and fallback flags before a build.  hflag 1: built by the
// fallback from the start (hub_build mode 1: every hub, mode 2: the hubs of
// odd index, a mixed set; mode 0: the hubs from hg on, whose tables have
// more than HUB_MAX_SLICES slices: every slice block reads all of a hub's
// edges, and for the largest hubs the fallback's edge items spread better
// over the CUs).  hflag 2: a slice of the hub was full (k_hub_build), and
// *hovf is set.
__global__ void k_hub_sm_init(const i64* __restrict__ goff, i64 m, i64 sl, int mode, i64 hg, u32* hsm, i32* hflag,
                              i32* hovf, i32* hcnt) {
    const i64 h = thread_item();
    if (h == 0) *hovf = 0;
    if (h >= m) return;
    hcnt[h] = 0;
    const i64 S = goff[h + 1] - goff[h];
    const bool fb = mode == 1 || (mode == 2 && (h & 1)) || (mode == 0 && h >= hg);
    hflag[h] = fb ? 1 : 0;
    hsm[h] = (u32)((fb || S < sl ? S : sl) - 1);
}

// LDS insert with the slice's probe sequence; false when the slice is full
__device__ __forceinline__ bool slice_add_packed(u64* lt, u32 smask, u32 q, i32 key, u64 w) {
    for (u32 probe = 0; probe <= smask; probe++) {
        const u64 cur = __hip_atomic_load(&lt[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((u32)(cur >> 32) == (u32)key) {
            atomicAdd((unsigned long long*)&lt[q], (unsigned long long)w);
            return true;
        }
        if (cur == PEMPTY) {
            const u64 old = atomicCAS((unsigned long long*)&lt[q], (unsigned long long)PEMPTY,
                                      (unsigned long long)(((u64)(u32)key << 32) | w));
            if (old == PEMPTY) return true;
            if ((u32)(old >> 32) == (u32)key) {
                atomicAdd((unsigned long long*)&lt[q], (unsigned long long)w);
                return true;
            }
        }
        q = (q + 1) & smask;
    }
    return false;
}
__device__ __forceinline__ bool slice_add_wide(i32* lk, i64* lv, u32 smask, u32 q, i32 key, i64 w) {
    for (u32 probe = 0; probe <= smask; probe++) {
        const i32 cur = __hip_atomic_load(&lk[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        bool hit = cur == key;
        if (!hit && cur == EMPTY_KEY) {
            const i32 old = atomicCAS(&lk[q], EMPTY_KEY, key);
            hit = old == EMPTY_KEY || old == key;
        }
        if (hit) {
            atomicAdd((unsigned long long*)&lv[q], (unsigned long long)w);
            return true;
        }
        q = (q + 1) & smask;
    }
    return false;
}

// SL: LDS slots of the block (>= the slice size of every slice it builds),
// BS threads.  Slice b0 + block id, or bmap[block id] when bmap is given.
template <bool PACKED, int SL, int BS, class WT>
__global__ void __launch_bounds__(BS) k_hub_build(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                  const i32* __restrict__ hubs, i64 m,
                                                  const i64* __restrict__ soff, i64 b0, const i32* __restrict__ bmap,
                                                  i64 nblk, const i64* __restrict__ goff,
                                                  const i32* __restrict__ part, u32* hsm, i32* hflag,
                                                  i32* hovf, i32* gkeys, i64* gvals, u64* gtab, u32* hlist,
                                                  i32* hcnt) {
    constexpr i64 SLMAX = PACKED ? HUB_SLICE_PACKED : HUB_SLICE_WIDE;
    __shared__ u64 lt[PACKED ? SL : 1];
    __shared__ i32 lk[PACKED ? 1 : SL];
    __shared__ i64 lv[PACKED ? 1 : SL];
    __shared__ int full, nocc, lbase;
    const i64 bb = block_id();
    if (bb >= nblk) return;
    const i64 b = bmap ? (i64)bmap[bb] : b0 + bb;
    if (b < 0) return;
    const i64 hi = upper_row(soff, 0, m, b);
    if (hflag[hi]) return;  // built by the fallback
    const int t = threadIdx.x;
    const i32 i = hubs[hi];
    const i64 S = goff[hi + 1] - goff[hi];
    const u32 mask = (u32)(S - 1);
    const u32 smask = (u32)((S < SLMAX ? S : SLMAX) - 1);
    if (smask >= (u32)SL) return;  // not this instance's slices (never launched so)
    const u32 base = (u32)(b - soff[hi]) * (smask + 1);  // first slot of the slice
    for (u32 q = t; q <= smask; q += BS) {
        if (PACKED) {
            lt[q] = PEMPTY;
        } else {
            lk[q] = EMPTY_KEY;
            lv[q] = 0;
        }
    }
    if (t == 0) {
        full = 0;
        nocc = 0;
    }
    __syncthreads();
    const i64 s = rp[i], e = rp[i + 1];
    bool ok = true;
    constexpr int U = 4;  // edges per thread in flight
    for (i64 jb = s; jb < e; jb += (i64)BS * U) {  // block-uniform loop control
        i32 kk[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const i64 j = jb + (i64)u * BS + t;
            kk[u] = j < e ? col[j] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++) kk[u] = kk[u] >= 0 ? part[kk[u]] : EMPTY_KEY;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (kk[u] == EMPTY_KEY) continue;
            const u32 h = slot_hash(kk[u], mask);
            if ((h & ~smask) != base) continue;
            const i64 w = W(jb + (i64)u * BS + t);
            ok = (PACKED ? slice_add_packed(lt, smask, h & smask, kk[u], (u64)w)
                         : slice_add_wide(lk, lv, smask, h & smask, kk[u], w)) && ok;
        }
    }
    if (!ok) full = 1;
    __syncthreads();
    if (full) {  // the hub goes to the fallback build
        if (t == 0) {
            hflag[hi] = 2;
            hsm[hi] = mask;
            *hovf = 1;
        }
        return;
    }
    const i64 g0 = goff[hi] + base;
    int mine = 0;
    for (u32 q = t; q <= smask; q += BS) {
        if (PACKED) {
            gtab[g0 + q] = lt[q];
            mine += lt[q] != PEMPTY;
        } else {
            gkeys[g0 + q] = lk[q];
            gvals[g0 + q] = lv[q];
            mine += lk[q] != EMPTY_KEY;
        }
    }
    // the slice's occupied slots join the hub's list (in any order)
    if (mine) atomicAdd(&nocc, mine);
    __syncthreads();
    if (t == 0) {
        lbase = nocc ? atomicAdd(&hcnt[hi], nocc) : 0;
        nocc = 0;
    }
    __syncthreads();
    u32* lst = hlist + goff[hi] + lbase;
    for (u32 q = t; q <= smask; q += BS)
        if (PACKED ? lt[q] != PEMPTY : lk[q] != EMPTY_KEY) lst[atomicAdd(&nocc, 1)] = base + q;
}

// empty every slice of a hub flagged for the fallback build
// after a full slice (grid-stride over the slices; nothing to do unless
// *hovf)
template <bool PACKED>
__global__ void __launch_bounds__(256) k_hub_clear(i64 m, const i64* __restrict__ soff, i64 nslices,
                                                   const i64* __restrict__ goff, const i32* __restrict__ hflag,
                                                   const i32* __restrict__ hovf, i32* gkeys, i64* gvals, u64* gtab,
                                                   i32* hcnt) {
    if (*hovf == 0) return;
    const i64 nblk = (i64)gridDim.x * gridDim.y;
    for (i64 b = block_id(); b < nslices; b += nblk) {
        const i64 hi = upper_row(soff, 0, m, b);
        if (hflag[hi] != 2) continue;
        if (b == soff[hi] && threadIdx.x == 0) hcnt[hi] = 0;  // relisted after the fallback build
        const i64 S = goff[hi + 1] - goff[hi];
        const i64 ns = soff[hi + 1] - soff[hi];
        const i64 len = S / ns;
        const i64 g0 = goff[hi] + (b - soff[hi]) * len;
        for (i64 q = threadIdx.x; q < len; q += 256) {
            if (PACKED) {
                gtab[g0 + q] = PEMPTY;
            } else {
                gkeys[g0 + q] = EMPTY_KEY;
                gvals[g0 + q] = 0;
            }
        }
    }
}

// Occupied-slot lists of the tables the fallback built: the slices
// b0 + k (grid-stride, n of them) of hubs with hflag == want; nothing to do
// when gate is given and *gate is 0.
template <bool PACKED>
__global__ void __launch_bounds__(256) k_hub_relist(i64 m, const i64* __restrict__ soff, i64 b0, i64 n,
                                                    const i64* __restrict__ goff, const i32* __restrict__ hflag,
                                                    i32 want, const i32* __restrict__ gate,
                                                    const i32* __restrict__ gkeys, const u64* __restrict__ gtab,
                                                    u32* hlist, i32* hcnt) {
    if (gate && *gate == 0) return;
    __shared__ int nocc, lbase;
    const i64 nblk = (i64)gridDim.x * gridDim.y;
    for (i64 k = block_id(); k < n; k += nblk) {  // block-uniform loop control
        const i64 b = b0 + k;
        const i64 hi = upper_row(soff, 0, m, b);
        if (hflag[hi] != want) continue;
        const i64 S = goff[hi + 1] - goff[hi];
        const i64 len = S / (soff[hi + 1] - soff[hi]);
        const i64 s0 = (b - soff[hi]) * len;
        const i64 g0 = goff[hi] + s0;
        if (threadIdx.x == 0) nocc = 0;
        __syncthreads();
        int mine = 0;
        for (i64 q = threadIdx.x; q < len; q += 256) mine += PACKED ? gtab[g0 + q] != PEMPTY : gkeys[g0 + q] != EMPTY_KEY;
        if (mine) atomicAdd(&nocc, mine);
        __syncthreads();
        if (threadIdx.x == 0) {
            lbase = nocc ? atomicAdd(&hcnt[hi], nocc) : 0;
            nocc = 0;
        }
        __syncthreads();
        u32* lst = hlist + goff[hi] + lbase;
        for (i64 q = threadIdx.x; q < len; q += 256)
            if (PACKED ? gtab[g0 + q] != PEMPTY : gkeys[g0 + q] != EMPTY_KEY) lst[atomicAdd(&nocc, 1)] = (u32)(s0 + q);
        __syncthreads();
    }
}

// Fallback build of the hubs with hflag == want (into emptied tables): edge
// work items (he = HL / 2 edges each) pre-aggregate in LDS, then add into
// the HBM table with atomics, probing over the whole table.  Grid-stride
// over the items b0 + k, or imap[k] when imap is given; nothing to do when
// gate is given and *gate is 0.
template <bool PACKED, int HL, class WT>
__global__ void __launch_bounds__(256) k_hub_insert(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                    const i32* __restrict__ hubs, i64 m, const i64* __restrict__ items,
                                                    const i32* __restrict__ imap, i64 b0, i64 nmap,
                                                    const i64* __restrict__ goff, const i32* __restrict__ part,
                                                    const i32* __restrict__ hflag, i32 want,
                                                    const i32* __restrict__ gate, i32* gkeys, i64* gvals, u64* gtab) {
    if (gate && *gate == 0) return;
    constexpr i64 he = HL / 2;
    __shared__ i32 lk[HL];
    __shared__ i64 lv[HL];
    const i64 nblk = (i64)gridDim.x * gridDim.y;
    for (i64 bb = block_id(); bb < nmap; bb += nblk) {  // block-uniform loop control
        const i64 b = imap ? (i64)imap[bb] : b0 + bb;
        if (b < 0) continue;
        const i64 hi = upper_row(items, 0, m, b);
        if (hflag[hi] != want) continue;
        const i32 i = hubs[hi];
        const i64 s = rp[i], e = rp[i + 1];
        const i64 j_lo = s + (b - items[hi]) * he;
        const i64 j_hi = j_lo + he < e ? j_lo + he : e;
        lds_agg_clear<256, HL>(lk, lv);
        for (i64 jb = j_lo; jb < j_hi; jb += 256) {  // block-uniform loop control
            const i64 j = jb + threadIdx.x;
            if (j >= j_hi) continue;
            table_add_lds(lk, lv, HL - 1, part[col[j]], W(j));
        }
        const u32 gmask = (u32)(goff[hi + 1] - goff[hi] - 1);
        if (PACKED) lds_agg_flush_packed<256, HL>(lk, lv, gtab + goff[hi], gmask);
        else lds_agg_flush<256, HL>(lk, lv, gkeys + goff[hi], gvals + goff[hi], gmask);
        __syncthreads();  // the flush reads LDS the next item clears
    }
}

// own-cluster weight of every active hub, read from its table
template <bool PACKED>
__global__ void k_hub_pv(const i32* __restrict__ hubs, i64 m, const i64* __restrict__ goff,
                         const u32* __restrict__ hsm, const i32* __restrict__ part, const i32* __restrict__ dest,
                         const i32* gkeys, const i64* gvals, const u64* gtab, i64* hub_pv) {
    const i64 hi = thread_item();
    if (hi >= m) return;
    const i32 i = hubs[hi];
    if (dest[i] == NO_MOVE) return;
    const u32 gmask = (u32)(goff[hi + 1] - goff[hi] - 1);
    hub_pv[hi] = PACKED ? ptable_get(gtab + goff[hi], gmask, hsm[hi], part[i])
                        : table_get(gkeys + goff[hi], gvals + goff[hi], gmask, hsm[hi], part[i]);
}

// Movers' edges into hubs: weight w of edge (i, u) leaves key part[i] and
// joins key dest[i] in hub u's table (run before the moves are applied).
// Same degree split as the uncut delta.  A failed update sets *stale.
template <bool PACKED, class WT>
__device__ __forceinline__ void hub_update_edge(const i32* col, WT W, const i32* hidx, const i64* goff,
                                                const u32* hsm, i32* gkeys, i64* gvals, u64* gtab, u32* hlist,
                                                i32* hcnt, i64 j, i32 oi, i32 ni, i64* stale) {
    const i32 h = hidx[col[j]];
    if (h < 0) return;
    const i64 w = W(j);
    const i64 g = goff[h];
    const u32 gmask = (u32)(goff[h + 1] - g - 1);
    const u32 sm = hsm[h];
    bool ok;
    if (PACKED) {
        ok = ptable_update(gtab + g, gmask, sm, oi, -w, hlist + g, hcnt + h);
        ok = ptable_update(gtab + g, gmask, sm, ni, w, hlist + g, hcnt + h) && ok;
    } else {
        ok = table_update(gkeys + g, gvals + g, gmask, sm, oi, -w, hlist + g, hcnt + h);
        ok = table_update(gkeys + g, gvals + g, gmask, sm, ni, w, hlist + g, hcnt + h) && ok;
    }
    if (!ok) *stale = 1;
}
template <bool PACKED, class WT>
__global__ void __launch_bounds__(256) k_hub_update_thread(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                           WT W, const i32* __restrict__ moves, i64 k,
                                                           const i32* __restrict__ part, const i32* __restrict__ dest,
                                                           const i32* __restrict__ hidx, const i64* __restrict__ goff,
                                                           const u32* __restrict__ hsm, i32* gkeys, i64* gvals,
                                                           u64* gtab, u32* hlist, i32* hcnt, i64* stale) {
    const i64 x = thread_item();
    if (x >= k) return;
    const i32 i = moves[x];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s >= LARGE_CUTOFF) return;
    const i32 oi = part[i], ni = dest[i];
    if (ni < 0) return;  // not applied (k_apply_moves)
#pragma unroll 4
    for (i64 j = s; j < e; j++) hub_update_edge<PACKED>(col, W, hidx, goff, hsm, gkeys, gvals, gtab, hlist, hcnt, j, oi, ni, stale);
}
template <bool PACKED, class WT>
__global__ void __launch_bounds__(256) k_hub_update_wave(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                         WT W, const i32* __restrict__ moves, i64 k, i64 nw,
                                                         const i32* __restrict__ part, const i32* __restrict__ dest,
                                                         const i32* __restrict__ hidx, const i64* __restrict__ goff,
                                                         const u32* __restrict__ hsm, i32* gkeys, i64* gvals,
                                                         u64* gtab, u32* hlist, i32* hcnt, i64* stale) {
    const int lane = threadIdx.x & 63;
    const i64 w = block_id() * 4 + (threadIdx.x >> 6);
    if (w >= nw) return;
    const i32 i = moves[k - 1 - w];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s < LARGE_CUTOFF || e - s >= MASSIVE_CUTOFF) return;
    const i32 oi = part[i], ni = dest[i];
    if (ni < 0) return;  // not applied (k_apply_moves)
    for (i64 base = s; base < e; base += 64) {  // wave-uniform loop control
        const i64 j = base + lane;
        if (j < e) hub_update_edge<PACKED>(col, W, hidx, goff, hsm, gkeys, gvals, gtab, hlist, hcnt, j, oi, ni, stale);
    }
}
// movers of degree >= MASSIVE_CUTOFF (the last nb entries of moves at most):
// a 1024-thread block each, so one huge row is not a single wave's serial tail
template <bool PACKED, class WT>
__global__ void __launch_bounds__(1024) k_hub_update_block(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                           WT W, const i32* __restrict__ moves, i64 k, i64 nb,
                                                           const i32* __restrict__ part, const i32* __restrict__ dest,
                                                           const i32* __restrict__ hidx, const i64* __restrict__ goff,
                                                           const u32* __restrict__ hsm, i32* gkeys, i64* gvals,
                                                           u64* gtab, u32* hlist, i32* hcnt, i64* stale) {
    const i64 w = block_id();
    if (w >= nb) return;
    const i32 i = moves[k - 1 - w];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s < MASSIVE_CUTOFF) return;
    const i32 oi = part[i], ni = dest[i];
    if (ni < 0) return;  // not applied (k_apply_moves)
    for (i64 j = s + threadIdx.x; j < e; j += 1024)
        hub_update_edge<PACKED>(col, W, hidx, goff, hsm, gkeys, gvals, gtab, hlist, hcnt, j, oi, ni, stale);
}

// Scan items cover hs entries of a hub's occupied-slot list (the grid is
// sized for full tables; items past the list's end exit at once).
template <bool PACKED, class LT>
__global__ void __launch_bounds__(256) k_hub_scan(const i32* __restrict__ hubs, i64 m, const i64* __restrict__ sitems,
                                                  const i32* __restrict__ smap, i64 nmap,
                                                  const i64* __restrict__ goff, const i32* __restrict__ part,
                                                  const i32* __restrict__ dest, const i64* __restrict__ vw,
                                                  const i64* __restrict__ vol, const i32* __restrict__ cons, LT lam,
                                                  float ratio, const i32* gkeys, const i64* gvals, const u64* gtab,
                                                  const u32* __restrict__ hlist, const i32* __restrict__ hcnt,
                                                  const i64* __restrict__ hub_pv, i64 hs, u64* hub_best) {
    const i64 bb = block_id();
    if (bb >= nmap) return;
    const i64 b = smap[bb];
    if (b < 0) return;
    const i64 hi = upper_row(sitems, 0, m, b);
    const i64 k_lo = (b - sitems[hi]) * hs;
    const i64 cnt = hcnt[hi];
    if (k_lo >= cnt) return;
    const i32 i = hubs[hi];
    if (dest[i] == NO_MOVE) return;
    const i64 k_hi = k_lo + hs < cnt ? k_lo + hs : cnt;
    const i64 g0 = goff[hi];
    const u32* lst = hlist + g0;
    const i32 p = part[i];
    float multi, p_conn, bth;
    own_terms(hub_pv[hi], vol[p], vw[i], lam(i), lam.rt(i, ratio), multi, p_conn, bth);
    u64 bk = 0;
    for (i64 kb = k_lo; kb < k_hi; kb += 256) {  // block-uniform loop control
        const i64 k = kb + threadIdx.x;
        if (k >= k_hi) continue;
        const i64 q = g0 + lst[k];
        i32 key;
        i64 agg;
        if (PACKED) {
            const u64 e = gtab[q];
            key = (i32)(e >> 32);
            agg = (i64)(e & 0xffffffffull);
        } else {
            key = gkeys[q];
            agg = gvals[q];
        }
        if (key != p && agg > 0 && (float)agg >= bth && (cons == nullptr || cons[key] == cons[p])) {
            const float jc = fmaf(-(float)vol[key], multi, (float)agg);
            if (jc >= bth) {
                const u64 k2 = pack_cand(jc, key);
                bk = k2 > bk ? k2 : bk;
            }
        }
    }
    bk = block_max_u64<256>(bk);
    if (threadIdx.x == 0 && bk) atomicMax((unsigned long long*)&hub_best[hi], (unsigned long long)bk);
}

template <class LT>
__global__ void k_hub_decide(const i32* hubs, i64 m, const i32* part, const i64* vw, const i64* vol, LT lam,
                             float ratio, const i64* hub_pv, const u64* hub_best, i32* dest, float* gain) {
    const i64 hi = thread_item();
    if (hi >= m) return;
    const i32 i = hubs[hi];
    if (dest[i] == NO_MOVE) return;
    float multi, p_conn, bth;
    own_terms(hub_pv[hi], vol[part[i]], vw[i], lam(i), lam.rt(i, ratio), multi, p_conn, bth);
    decide(i, p_conn, hub_best[hi], dest, gain);
}


// Algorithmic bytes of one scoring sweep, counted (on timing runs only)
// before it runs, since dest decides which vertices it scores.  A scored
// vertex reads its row bounds, destination, cluster, vertex weight and its
// cluster's volume, every edge's column (and weight) and, after the first
// sweep, the neighbour's cluster, and writes its destination and gain:
// 48 + d (4 + weight bytes + 4) bytes.  A hub whose table is kept current reads
// the table instead of its edges: 48 + 8 per slot, plus the edge bytes on the
// sweep that builds the table.  Every listed vertex also costs its list entry
// and destination read.  acc3[0..2] = group, table and hub classes.
struct ScoreBounds {
    i64 s[NCLS + 1];
    i64 hub0;  // first HBM-table hub (the hubs before it are scored in LDS)
};
template <int WB>
__global__ void __launch_bounds__(256) k_score_bytes(const i64* __restrict__ rp, const i32* __restrict__ list,
                                                     ScoreBounds sb, const i32* __restrict__ dest, int raw, int build,
                                                     const i64* __restrict__ goff, i64* acc3) {
    const i64 x = thread_item();
    i64 b0 = 0, b1 = 0, b2 = 0;
    if (x < sb.s[NCLS]) {
        const i32 i = list[x];
        const int grp = x < sb.s[CL_TABLE] ? 0 : (x < sb.hub0 ? 1 : 2);
        i64 by = 8;
        if (dest[i] != NO_MOVE) {
            const i64 eb = (rp[i + 1] - rp[i]) * (4 + WB + (raw == 1 ? 0 : 4));
            if (grp < 2 || raw) {
                by += 48 + eb;
            } else {
                const i64 h = x - sb.hub0;
                by += 48 + 8 * (goff[h + 1] - goff[h]) + (build ? eb : 0);
            }
        }
        if (grp == 0) b0 = by;
        else if (grp == 1) b1 = by;
        else b2 = by;
    }
    b0 = block_sum_i64<256>(b0);
    b1 = block_sum_i64<256>(b1);
    b2 = block_sum_i64<256>(b2);
    if (threadIdx.x == 0) {
        if (b0) atomicAdd((unsigned long long*)&acc3[0], (unsigned long long)b0);
        if (b1) atomicAdd((unsigned long long*)&acc3[1], (unsigned long long)b1);
        if (b2) atomicAdd((unsigned long long*)&acc3[2], (unsigned long long)b2);
    }
}

// alpha + beta * degree bytes per listed vertex (per candidate of the
// afterburner, per mover of perform_moves), plus `extra` once
__global__ void __launch_bounds__(256) k_list_bytes(const i64* __restrict__ rp, const i32* __restrict__ list, i64 k,
                                                    i64 alpha, i64 beta, i64 extra, i64* acc) {
    const i64 x = thread_item();
    i64 b = 0;
    if (x < k) {
        const i32 i = list[x];
        b = alpha + beta * (rp[i + 1] - rp[i]);
    }
    if (x == 0) b += extra;
    b = block_sum_i64<256>(b);
    if (threadIdx.x == 0 && b) atomicAdd((unsigned long long*)acc, (unsigned long long)b);
}

// LDS-table scoring of the list entries [x0, x1) (one block per vertex):
// BS threads, SLOTS LDS slots, SPE slots per edge of the degree bound.
template <int BS, int SLOTS, int SPE, class WT, class LT>
static void launch_table(Ctx& c, const DGraph& g, WT W, const i32* L, i64 x0, i64 x1, const i32* part,
                         const i64* vol, const i32* cons, LT lam, float ratio, i32* dest, float* gain) {
    const i64 m = x1 - x0;
    if (m <= 0) return;
    k_score_table<BS, SLOTS, SPE, false><<<grid_blocks(m), BS, 0, c.stream>>>(
        g.rp, g.col, W, g.vw, L + x0, m, part, vol, cons, 0, lam, ratio, dest, gain, nullptr, nullptr, nullptr);
}
// class with degree bound D (128, 256, 512): threads per vertex and slots per
// edge from the knobs
template <int D, class WT, class LT>
static void launch_table_class(Ctx& c, const DGraph& g, WT W, const i32* L, i64 x0, i64 x1, const i32* part,
                               const i64* vol, const i32* cons, LT lam, float ratio, i32* dest, float* gain) {
    const int bs = c.table_threads, spe = c.table_spe;
    if (bs == 64 && spe == 2) launch_table<64, 2 * D, 2>(c, g, W, L, x0, x1, part, vol, cons, lam, ratio, dest, gain);
    else if (bs == 64) launch_table<64, 4 * D, 4>(c, g, W, L, x0, x1, part, vol, cons, lam, ratio, dest, gain);
    else if (bs == 128 && spe == 2)
        launch_table<128, 2 * D, 2>(c, g, W, L, x0, x1, part, vol, cons, lam, ratio, dest, gain);
    else if (bs == 128) launch_table<128, 4 * D, 4>(c, g, W, L, x0, x1, part, vol, cons, lam, ratio, dest, gain);
    else if (spe == 2) launch_table<256, 2 * D, 2>(c, g, W, L, x0, x1, part, vol, cons, lam, ratio, dest, gain);
    else launch_table<256, 4 * D, 4>(c, g, W, L, x0, x1, part, vol, cons, lam, ratio, dest, gain);
}

// Hub tables of the current clustering: slices built in LDS, then the
// fallback build of the hubs flagged (by a full slice or the hub_build knob).
// occupied-slot lists of the fallback-built tables (slices b0 .. b0+n-1)
static void hub_relist(Ctx& c, LevelPlan& lp, i64 m, i64 b0, i64 n, i64 cap, i32 want, const i32* gate) {
    if (n <= 0) return;
    const dim3 grid = grid_blocks(cap > 0 ? std::min(n, cap) : n);
    if (lp.packed)
        k_hub_relist<true><<<grid, 256, 0, c.stream>>>(m, lp.soff, b0, n, lp.goff, lp.hflag, want, gate, nullptr,
                                                       lp.gtab, lp.hlist, lp.hcnt);
    else
        k_hub_relist<false><<<grid, 256, 0, c.stream>>>(m, lp.soff, b0, n, lp.goff, lp.hflag, want, gate, lp.gkeys,
                                                        nullptr, lp.hlist, lp.hcnt);
    PGGC_LAUNCH_CHECK();
}

// fallback insert launch over n item slots (b0 + k, or imap[k]); grid
// capped at cap blocks (grid-stride) when cap > 0
template <class WT>
static void hub_insert_items(Ctx& c, const DGraph& g, WT W, LevelPlan& lp, const i32* hubs, i64 m, const i32* part,
                             const i32* imap, i64 b0, i64 n, i64 cap, i32 want, const i32* gate) {
    if (n <= 0) return;
    i32* gk = lp.packed ? nullptr : lp.gkeys.p;
    i64* gv = lp.packed ? nullptr : lp.gvals.p;
    u64* gt = lp.packed ? lp.gtab.p : nullptr;
    const dim3 grid = grid_blocks(cap > 0 ? std::min(n, cap) : n);
#define PGGC_HUB_INSERT(PK, HL)                                                                                   \
    k_hub_insert<PK, HL><<<grid, 256, 0, c.stream>>>(g.rp, g.col, W, hubs, m, lp.items, imap, b0, n, lp.goff, part, \
                                                     lp.hflag, want, gate, gk, gv, gt)
    if (lp.packed) {
        if (lp.he == 512) PGGC_HUB_INSERT(true, 1024);
        else if (lp.he == 2048) PGGC_HUB_INSERT(true, 4096);
        else PGGC_HUB_INSERT(true, 2048);
    } else {
        if (lp.he == 512) PGGC_HUB_INSERT(false, 1024);
        else if (lp.he == 2048) PGGC_HUB_INSERT(false, 4096);
        else PGGC_HUB_INSERT(false, 2048);
    }
#undef PGGC_HUB_INSERT
    PGGC_LAUNCH_CHECK();
}

template <class WT>
static void hub_build_all(Ctx& c, const DGraph& g, WT W, LevelPlan& lp, const i32* hubs, i64 m, const i32* part) {
    i32* gk = lp.packed ? nullptr : lp.gkeys.p;
    i64* gv = lp.packed ? nullptr : lp.gvals.p;
    u64* gt = lp.packed ? lp.gtab.p : nullptr;
    const int mode = c.hub_build;
    k_hub_sm_init<<<grid_items(m, 256), 256, 0, c.stream>>>(lp.goff, m, hub_slice(lp.packed), mode, lp.hg, lp.hsm,
                                                            lp.hflag, lp.hovf, lp.hcnt);
    // empty the tables the fallback builds from the start (before the slice
    // builds, which write the other tables whole)
    const i64 f0 = mode == 0 ? lp.ghg : 0;
    if (lp.gslots > f0) {
        if (lp.packed) {
            fill_i64(c, (i64*)lp.gtab.p + f0, lp.gslots - f0, (i64)PEMPTY);
        } else {
            fill_i32(c, lp.gkeys.p + f0, lp.gslots - f0, EMPTY_KEY);
            fill_i64(c, lp.gvals.p + f0, lp.gslots - f0, 0);
        }
    }
    if (mode != 1) {
        // single-slice tables of 4096 and (packed) 8192 slots in smaller
        // blocks, so several run per CU; the multi-slice tables through the
        // XCD map
        const i64 n1 = lp.hs1, n2 = lp.hs2 - lp.hs1;
#define PGGC_HUB_BUILD(PK, SL, BS, B0, MAP, N)                                                                    \
    k_hub_build<PK, SL, BS><<<grid_blocks(N), BS, 0, c.stream>>>(g.rp, g.col, W, hubs, m, lp.soff, B0, MAP, N, lp.goff, \
                                                                 part, lp.hsm, lp.hflag, lp.hovf, gk, gv, gt, lp.hlist, \
                                                                 lp.hcnt)
        if (lp.packed) {
            if (n1 > 0) PGGC_HUB_BUILD(true, 4096, 256, 0, nullptr, n1);
            if (n2 > 0) PGGC_HUB_BUILD(true, 8192, 512, lp.hs1, nullptr, n2);
            if (lp.bgrid > 0) PGGC_HUB_BUILD(true, 16384, 1024, 0, lp.bmap.p, lp.bgrid);
        } else {
            if (n1 > 0) PGGC_HUB_BUILD(false, 4096, 256, 0, nullptr, n1);
            if (lp.bgrid > 0) PGGC_HUB_BUILD(false, 8192, 512, 0, lp.bmap.p, lp.bgrid);
        }
#undef PGGC_HUB_BUILD
        // hubs with a full slice (normally none: the gated kernels exit)
        constexpr i64 FB_GRID = 4096;
        const dim3 cgrid = grid_blocks(std::min(lp.nslices, FB_GRID));
        if (lp.packed)
            k_hub_clear<true><<<cgrid, 256, 0, c.stream>>>(m, lp.soff, lp.nslices, lp.goff, lp.hflag, lp.hovf, gk, gv,
                                                           gt, lp.hcnt);
        else
            k_hub_clear<false><<<cgrid, 256, 0, c.stream>>>(m, lp.soff, lp.nslices, lp.goff, lp.hflag, lp.hovf, gk,
                                                            gv, gt, lp.hcnt);
        hub_insert_items(c, g, W, lp, hubs, m, part, lp.imap.p, 0, lp.igrid, FB_GRID, 2, lp.hovf.p);
        hub_relist(c, lp, m, 0, lp.nslices, FB_GRID, 2, lp.hovf.p);
    }
    // the hubs built by the fallback from the start: in mode 0 the items
    // from the first of the largest hubs on, else all items through the map
    if (mode == 0) {
        if (lp.nitems > lp.ihg) {
            hub_insert_items(c, g, W, lp, hubs, m, part, nullptr, lp.ihg, lp.nitems - lp.ihg, 0, 1, nullptr);
            hub_relist(c, lp, m, lp.shg, lp.nslices - lp.shg, 0, 1, nullptr);
        }
    } else {
        hub_insert_items(c, g, W, lp, hubs, m, part, lp.imap.p, 0, lp.igrid, 0, 1, nullptr);
        hub_relist(c, lp, m, 0, lp.nslices, 0, 1, nullptr);
    }
    PGGC_LAUNCH_CHECK();
}

template <class LT>
static void score_all(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, const i32* part, const i64* vol,
                      const i32* cons, int raw, LT lam, float ratio, i32* dest, float* gain) {
    // raw 1: singletons labelled by vertex id; raw 2: singletons whose labels
    // are read from part (a renumbered graph); 0: any clustering
    const i32* lab = raw == 2 ? part : nullptr;
    const i64* st = lp.sc_start;
    const i32* L = lp.sc_list.p;
    Buf<i64> acct;
    if (c.kt.on && g.n > 0) {  // timing runs only
        acct = Buf<i64>(c, 3);
        fill_i64(c, acct, 3, 0);
        ScoreBounds sb;
        for (int k = 0; k <= NCLS; k++) sb.s[k] = st[k];
        sb.hub0 = lp.hub0;
        const int build = (!raw && !lp.hub_live) ? 1 : 0;
        const i64* goff = lp.nhub > 0 ? lp.goff.p : nullptr;
        PGGC_WDISPATCH(g, (k_score_bytes<decltype(W)::bytes><<<grid_items(g.n, 256), 256, 0, c.stream>>>(
                              g.rp, L, sb, dest, raw, build, goff, acct)));
        PGGC_LAUNCH_CHECK();
    }
    PGGC_WDISPATCH(g, {
        i64 m;
        int tk = c.kt_begin(KS_SCORE_GROUP);
        if ((m = st[1] - st[0]) > 0)
            k_score_group<4><<<grid_items(m, 64), 256, 0, c.stream>>>(
                g.rp, g.col, W, g.vw, L + st[0], m, part, vol, cons, raw, lam, ratio, dest, gain);
        if ((m = st[2] - st[1]) > 0)
            k_score_group<8><<<grid_items(m, 32), 256, 0, c.stream>>>(
                g.rp, g.col, W, g.vw, L + st[1], m, part, vol, cons, raw, lam, ratio, dest, gain);
        if ((m = st[3] - st[2]) > 0)
            k_score_group<16><<<grid_items(m, 16), 256, 0, c.stream>>>(
                g.rp, g.col, W, g.vw, L + st[2], m, part, vol, cons, raw, lam, ratio, dest, gain);
        if ((m = st[4] - st[3]) > 0)
            k_score_group<32><<<grid_items(m, 8), 256, 0, c.stream>>>(
                g.rp, g.col, W, g.vw, L + st[3], m, part, vol, cons, raw, lam, ratio, dest, gain);
        if ((m = st[5] - st[4]) > 0)
            k_score_group<64><<<grid_items(m, 4), 256, 0, c.stream>>>(
                g.rp, g.col, W, g.vw, L + st[4], m, part, vol, cons, raw, lam, ratio, dest, gain);
        c.kt_end(tk, 0.0);
        if (tk >= 0) acct_add(c, c.kt_counter(tk), acct.p);
        tk = c.kt_begin(KS_SCORE_TABLE);
        if (raw) {
            if ((m = st[CL_BIG] - st[CL_TABLE]) > 0)
                k_score_raw<64><<<grid_blocks(m), 64, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[CL_TABLE], m,
                                                                     part, vol, cons, lab, lam, ratio, dest, gain);
            if ((m = lp.hub0 - st[CL_BIG]) > 0)  // big rows and the LDS-scored hubs
                k_score_raw<256><<<grid_blocks(m), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[CL_BIG], m,
                                                                       part, vol, cons, lab, lam, ratio, dest, gain);
        } else {
            // LDS tables sized for the class, so the many low-degree rows keep
            // more waves resident per CU; threads per vertex and slots per
            // edge are knobs
            launch_table_class<128>(c, g, W, L, st[5], st[6], part, vol, cons, lam, ratio, dest, gain);
            launch_table_class<256>(c, g, W, L, st[6], st[7], part, vol, cons, lam, ratio, dest, gain);
            launch_table_class<512>(c, g, W, L, st[7], st[8], part, vol, cons, lam, ratio, dest, gain);
            if (c.table_spe == 2) {
                launch_table<256, 2048, 2>(c, g, W, L, st[8], st[9], part, vol, cons, lam, ratio, dest, gain);
                launch_table<512, 4096, 2>(c, g, W, L, st[9], st[10], part, vol, cons, lam, ratio, dest, gain);
            } else {
                launch_table<256, 4096, 4>(c, g, W, L, st[8], st[9], part, vol, cons, lam, ratio, dest, gain);
                launch_table<512, 8192, 4>(c, g, W, L, st[9], st[10], part, vol, cons, lam, ratio, dest, gain);
            }
            // hubs of degree <= hub_lds_max: 8192 LDS slots (96 KB of LDS)
            launch_table<512, 8192, 2>(c, g, W, L, st[CL_HUB], lp.hub0, part, vol, cons, lam, ratio, dest, gain);
        }
        c.kt_end(tk, 0.0);
        if (tk >= 0) acct_add(c, c.kt_counter(tk), acct.p + 1);
        if ((m = lp.nhub) > 0) {
            const i32* hubs = L + lp.hub0;
            tk = c.kt_begin(KS_SCORE_HUB);
            fill_i64(c, lp.hub_pv, m, 0);
            fill_i64(c, (i64*)lp.hub_best.p, m, 0);
            if (raw) {
                k_hub_raw<<<grid_blocks(lp.nitems), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, hubs, m, lp.items, part,
                                                                       vol, cons, lab, lam, ratio, dest, lp.he,
                                                                       lp.hub_best);
            } else {
                if (!lp.hub_live) {  // first table sweep of the call (or stale tables): build
                    hub_build_all(c, g, W, lp, hubs, m, part);
                    lp.hub_live = true;
                }
                if (lp.packed) {
                    k_hub_pv<true><<<grid_items(m, 256), 256, 0, c.stream>>>(hubs, m, lp.goff, lp.hsm, part, dest, nullptr,
                                                                            nullptr, lp.gtab, lp.hub_pv);
                    k_hub_scan<true><<<grid_blocks(lp.sgrid), 256, 0, c.stream>>>(
                        hubs, m, lp.sitems, lp.smap, lp.sgrid, lp.goff, part, dest, g.vw, vol, cons, lam, ratio,
                        nullptr, nullptr, lp.gtab, lp.hlist, lp.hcnt, lp.hub_pv, lp.hs, lp.hub_best);
                } else {
                    k_hub_pv<false><<<grid_items(m, 256), 256, 0, c.stream>>>(hubs, m, lp.goff, lp.hsm, part, dest, lp.gkeys,
                                                                             lp.gvals, nullptr, lp.hub_pv);
                    k_hub_scan<false><<<grid_blocks(lp.sgrid), 256, 0, c.stream>>>(
                        hubs, m, lp.sitems, lp.smap, lp.sgrid, lp.goff, part, dest, g.vw, vol, cons, lam, ratio,
                        lp.gkeys, lp.gvals, nullptr, lp.hlist, lp.hcnt, lp.hub_pv, lp.hs, lp.hub_best);
                }
            }
            k_hub_decide<<<grid_items(m, 256), 256, 0, c.stream>>>(hubs, m, part, g.vw, vol, lam, ratio, lp.hub_pv,
                                                                       lp.hub_best, dest, gain);
            c.kt_end(tk, 0.0);
            if (tk >= 0) acct_add(c, c.kt_counter(tk), acct.p + 2);
        }
    });
    PGGC_LAUNCH_CHECK();
}

// ------------------------------------------------------------- candidates
__global__ void k_cand_flags(const i32* o1, i64 n, const i32* dest, i32* flag) {
    const i64 x = thread_item();
    if (x >= n) return;
    flag[x] = dest[o1[x]] != NO_MOVE ? 1 : 0;
}
// first list positions whose degree reaches LARGE_CUTOFF / MASSIVE_CUTOFF
__global__ void k_list_bounds(const i32* list, i64 k, const i64* rp, i64* out) {
    if (threadIdx.x >= 2) return;
    i64 cut = threadIdx.x == 0 ? LARGE_CUTOFF : MASSIVE_CUTOFF;
    i64 lo = 0, hi = k;
    while (lo < hi) {
        i64 m = (lo + hi) >> 1;
        i32 i = list[m];
        if (rp[i + 1] - rp[i] >= cut) hi = m;
        else lo = m + 1;
    }
    out[threadIdx.x] = lo;
}

// candidates in order1 order; returns count, b[0]/b[1] = class boundaries
static i64 gather_candidates(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, const i32* dest, i32* out,
                             i64* b) {
    k_cand_flags<<<grid_items(g.n, 256), 256, 0, c.stream>>>(lp.ord.o1, g.n, dest, s.flag);
    PGGC_LAUNCH_CHECK();
    i64 k = compact_flags(c, s.flag, lp.ord.o1, g.n, out, s.scal);
    if (k > 0) {
        k_list_bounds<<<1, 64, 0, c.stream>>>(out, k, g.rp, s.scal.p + 1);
        PGGC_LAUNCH_CHECK();
        const i64* h = c.read(s.scal.p + 1, 2);
        b[0] = h[0];
        b[1] = h[1];
    } else {
        b[0] = b[1] = 0;
    }
    return k;
}

// ------------------------------------------------------------- afterburner
// afterburner_filter (src/local_mover.cpp:374-433): the change in the
// candidate's gain if every neighbour with a larger gain (ties broken by the
// xorshift hash) also makes its move
template <class WT, class PT>
__device__ __forceinline__ void ab_term(const i32* col, WT W, const AbInfo* info, i64 j, i32 p, i32 best,
                                        float multi, float igain, i32 hi, PT prio, float& acc) {
    const i32 v = col[j];
    const AbInfo q4 = info[v];
    const float vg = q4.gain;
    if ((vg - igain) >= AFTERBURNER_EPS || (fabsf(vg - igain) < AFTERBURNER_EPS && prio(v) < hi)) {
        i32 vpart = q4.dest;
        const float q = fmaf(-multi, q4.vwf, (float)W(j));
        acc = acc - ((vpart == p) ? q : 0.0f);
        acc = acc + ((vpart == best && best != NEW_PART) ? q : 0.0f);
        vpart = q4.part;
        acc = acc + ((vpart == p) ? q : 0.0f);
        acc = acc - ((vpart == best) ? q : 0.0f);
    }
}

template <class WT, class LT, class PT>
__global__ void __launch_bounds__(256) k_ab_serial(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                   const AbInfo* __restrict__ info, const i32* __restrict__ cand,
                                                   i64 k, LT lam, PT prio, i32* keep, i32* mflag) {
    const i64 x = thread_item();
    if (x >= k) return;
    const i32 i = cand[x];
    const AbInfo me = info[i];
    const i32 best = me.dest, p = me.part;
    const float multi = me.vwf * lam(i);
    const float igain = me.gain;
    const i32 hi = prio(i);
    float acc = 0.0f;
    const i64 e = rp[i + 1];
#pragma unroll 4  // independent neighbour gathers in flight; the adds keep their order
    for (i64 j = rp[i]; j < e; j++) ab_term(col, W, info, j, p, best, multi, igain, hi, prio, acc);
    const int kp = (igain + acc >= 0.0f) ? 1 : 0;
    keep[x] = kp;
    mflag[i] = kp;  // candidates start with mflag 0: an unconditional store
}

template <class WT, class LT, class PT>
__global__ void __launch_bounds__(256) k_ab_wave(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                 const AbInfo* __restrict__ info, const i32* __restrict__ cand, i64 k0,
                                                 i64 k1, LT lam, PT prio, i32* keep, i32* mflag) {
    const int lane = threadIdx.x & 63;
    const i64 x = k0 + block_id() * 4 + (threadIdx.x >> 6);
    if (x >= k1) return;
    {
        const i32 i = cand[x];
        const AbInfo me = info[i];
        const i32 best = me.dest, p = me.part;
        const float multi = me.vwf * lam(i);
        const float igain = me.gain;
        const i32 hi = prio(i);
        float acc = 0.0f;
        const i64 e = rp[i + 1];
#pragma unroll 4
        for (i64 base = rp[i]; base < e; base += 64) {  // wave-uniform loop control
            const i64 j = base + lane;
            if (j < e) ab_term(col, W, info, j, p, best, multi, igain, hi, prio, acc);
        }
        xlane_fence(acc);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off);
        if (lane == 0) {
            const int kp = (igain + acc >= 0.0f) ? 1 : 0;
            keep[x] = kp;
            if (kp) mflag[i] = 1;
        }
    }
}

template <class WT, class LT, class PT>
__global__ void __launch_bounds__(1024) k_ab_block(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                   const AbInfo* __restrict__ info, const i32* __restrict__ cand,
                                                   i64 k0, i64 k1, LT lam, PT prio, i32* keep, i32* mflag) {
    __shared__ float sh[1024];
    const int t = threadIdx.x;
    const i64 x = k0 + block_id();
    if (x >= k1) return;
    {
        const i32 i = cand[x];
        const AbInfo me = info[i];
        const i32 best = me.dest, p = me.part;
        const float multi = me.vwf * lam(i);
        const float igain = me.gain;
        const i32 hi = prio(i);
        float acc = 0.0f;
        const i64 e = rp[i + 1];
        for (i64 base = rp[i]; base < e; base += 1024) {  // block-uniform loop control
            const i64 j = base + t;
            if (j < e) ab_term(col, W, info, j, p, best, multi, igain, hi, prio, acc);
        }
        sh[t] = acc;
        __syncthreads();
        for (int off = 512; off >= 64; off >>= 1) {
            if (t < off) sh[t] = sh[t] + sh[t + off];
            __syncthreads();
        }
        if (t < 64) {
            acc = sh[t];
            xlane_fence(acc);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off);
            if (t == 0) {
                const int kp = (igain + acc >= 0.0f) ? 1 : 0;
                keep[x] = kp;
                if (kp) mflag[i] = 1;
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_ab_pack(i64 n, const float* __restrict__ gain, const i32* __restrict__ dest,
                                                 const i32* __restrict__ part, const i64* __restrict__ vw,
                                                 AbInfo* __restrict__ info) {
    const i64 i = thread_item();
    if (i >= n) return;
    AbInfo a;
    a.gain = gain[i];
    a.dest = dest[i];
    a.part = part[i];
    a.vwf = (float)vw[i];
    info[i] = a;
}

template <class LT, class PT>
static i64 afterburner(Ctx& c, MoveScratch& s, const DGraph& g, const i32* cand, i64 k, const i64* b, LT lam, PT prio,
                       i32* moves) {
    const int tk = c.kt_begin(KS_AFTERBURNER);
    k_ab_pack<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.n, s.gain, s.dest, s.part, g.vw, s.abi);
    PGGC_LAUNCH_CHECK();
    PGGC_WDISPATCH(g, {
        if (b[0] > 0)
            k_ab_serial<<<grid_items(b[0], 256), 256, 0, c.stream>>>(g.rp, g.col, W, s.abi, cand, b[0], lam, prio,
                                                                     s.flag, s.mflag);
        if (b[1] > b[0])
            k_ab_wave<<<grid_items(b[1] - b[0], 4), 256, 0, c.stream>>>(g.rp, g.col, W, s.abi, cand, b[0], b[1], lam,
                                                                        prio, s.flag, s.mflag);
        if (k > b[1])
            k_ab_block<<<grid_blocks(k - b[1]), 1024, 0, c.stream>>>(g.rp, g.col, W, s.abi, cand, b[1], k, lam,
                                                                     prio, s.flag, s.mflag);
    });
    c.kt_end(tk, 0.0);
    PGGC_LAUNCH_CHECK();
    if (tk >= 0) {  // timing runs only: count the bytes outside the timed region
        // the pack reads gain, dest, part and vertex weight and writes a
        // 16-byte record per vertex; the filter reads per candidate edge the
        // column, weight and the neighbour's record, per candidate its id, row
        // bounds and keep flag
        Buf<i64> acc(c, 1);
        fill_i64(c, acc, 1, 0);
        k_list_bytes<<<grid_items(std::max<i64>(k, 1), 256), 256, 0, c.stream>>>(
            g.rp, cand, k, 24, 20 + (g.wk == W_UNIT ? 0 : g.wk == W_I32 ? 4 : 8), 36 * g.n, acc);
        PGGC_LAUNCH_CHECK();
        acct_add(c, c.kt_counter(tk), acc);
    }
    return compact_flags(c, s.flag, cand, k, moves, s.scal);
}

// ------------------------------------------------------------- new labels
__global__ void k_new_flags(const i32* moves, i64 k, const i32* dest, i32* flag) {
    const i64 x = thread_item();
    if (x >= k) return;
    flag[x] = dest[moves[x]] == NEW_PART ? 1 : 0;
}
__global__ void k_mark_used(const i32* part, i64 n, i32* used) {
    const i64 i = thread_item();
    if (i >= n) return;
    used[part[i]] = 1;
}
__global__ void k_unused_flags(const i32* used, i64 L, i32* flag) {
    const i64 l = thread_item();
    if (l >= L) return;
    flag[l] = used[l] == 0 ? 1 : 0;
}
__global__ void k_assign_new(const i32* newv, i64 K, const i32* avail, i64 A, i32 curr, i32* dest, i32* cons,
                             const i32* part) {
    const i64 x = thread_item();
    if (x >= K) return;
    i32 i = newv[x];
    i32 l = x < A ? avail[x] : (i32)(x - A) + curr;
    dest[i] = l;
    if (cons) cons[l] = cons[part[i]];
}

// key of a NEW_PART mover in the caller's move order: order class
// (src/header/ordering.h:17-19), then the caller's id
__global__ void k_newid_keys(const i32* newv, i64 K, const i64* rp, const i32* ids, u64* key) {
    const i64 x = thread_item();
    if (x >= K) return;
    const i32 i = newv[x];
    const i64 d = rp[i + 1] - rp[i];
    const u64 cls = d < LARGE_CUTOFF ? 0 : (d < MASSIVE_CUTOFF ? 1 : 2);
    key[x] = cls << 32 | (u32)ids[i];
}

// set_new_cluster_ids (src/local_mover.cpp:215-274).  ids non-null: the
// graph is renumbered and ids[i] is vertex i's id in the caller's order; the
// movers are then taken in the caller's move order.
static void set_new_cluster_ids(Ctx& c, MoveScratch& s, const DGraph& g, RefineData& rd, const i32* moves, i64 k,
                                i32* cons, const i32* ids) {
    k_new_flags<<<grid_items(k, 256), 256, 0, c.stream>>>(moves, k, s.dest, s.flag);
    PGGC_LAUNCH_CHECK();
    i64 K = compact_flags(c, s.flag, moves, k, s.tmp, s.scal);
    if (K == 0) return;
    if (ids && K > 1) {
        Buf<u64> key(c, K);
        k_newid_keys<<<grid_items(K, 256), 256, 0, c.stream>>>(s.tmp, K, g.rp, ids, key);
        PGGC_LAUNCH_CHECK();
        radix_sort_pairs(c, key, s.tmp.p, K, 34);
    }
    const i32 curr = rd.label_count;
    fill_i32(c, s.flag, curr, 0);
    k_mark_used<<<grid_items(g.n, 256), 256, 0, c.stream>>>(s.part, g.n, s.flag);
    PGGC_LAUNCH_CHECK();
    Buf<i32> uf(c, std::max<i64>(curr, 1));
    k_unused_flags<<<grid_items(curr, 256), 256, 0, c.stream>>>(s.flag, curr, uf);
    PGGC_LAUNCH_CHECK();
    i64 A = compact_flags(c, uf, nullptr, curr, s.tmp2, s.scal);
    if (K > A) {
        i64 need = K - A;
        if (curr + need > rd.cap) throw Error(-4, "label capacity exceeded while assigning new clusters");
        fill_i64(c, rd.vol + curr, need, 0);
        rd.label_count = (i32)(curr + need);
    }
    k_assign_new<<<grid_items(K, 256), 256, 0, c.stream>>>(s.tmp, K, s.tmp2, A, curr, s.dest, cons, s.part);
    PGGC_LAUNCH_CHECK();
}

// ------------------------------------------------------------- perform moves
// Exact uncut delta of the moves (src/local_mover.cpp:1274-1277 computes the
// same quantity from pvals): per mover, the sum over its edges of
// w * ([new_i == new_v] - [old_i == old_v]), counted once for mover pairs and
// twice for edges to vertices that stay.  Move lists are in order1 order, so
// movers of degree >= LARGE_CUTOFF form a suffix: they take one wave each
// (counted from the end); the others one thread each.  out[x] = delta of mover x.
// mvp[v]: low word part[v] | mflag[v] << 31 (cluster ids are >= 0), high
// word the cluster after the moves (dest[v] for a mover, else part[v]), so
// the delta reads one 8-byte word per edge
__global__ void k_mv_pack(i64 n, const i32* __restrict__ part, const i32* __restrict__ dest,
                          const i32* __restrict__ mflag, u64* __restrict__ mvp) {
    const i64 i = thread_item();
    if (i >= n) return;
    const i32 ov = part[i];
    const u32 mv = mflag[i] ? 1u : 0u;
    const i32 nv = mv ? dest[i] : ov;
    mvp[i] = (u64)((u32)ov | (mv << 31)) | ((u64)(u32)nv << 32);
}

template <class WT>
__device__ __forceinline__ i64 delta_term(const i32* col, WT W, const u64* mvp, i64 j, i32 oi, i32 ni) {
    const u64 q = mvp[col[j]];
    const u32 lo = (u32)q;
    const i32 ov = (i32)(lo & 0x7fffffffu);
    const bool mv = (lo >> 31) != 0;
    const i32 nv = (i32)(u32)(q >> 32);
    const i64 dd = (i64)(ni == nv) - (i64)(oi == ov);
    return dd * W(j) * (mv ? 1 : 2);
}
template <class WT>
__global__ void __launch_bounds__(256) k_delta_uncut_thread(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                            WT W, const i32* __restrict__ moves, i64 k,
                                                            const i32* __restrict__ part,
                                                            const i32* __restrict__ dest,
                                                            const u64* __restrict__ mvp, i64* out) {
    const i64 x = thread_item();
    if (x >= k) return;
    const i32 i = moves[x];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s >= LARGE_CUTOFF) return;
    const i32 oi = part[i], ni = dest[i];
    i64 acc = 0;
#pragma unroll 4
    for (i64 j = s; j < e; j++) acc += delta_term(col, W, mvp, j, oi, ni);
    out[x] = acc;
}
template <class WT>
__global__ void __launch_bounds__(256) k_delta_uncut_wave(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                          WT W, const i32* __restrict__ moves, i64 k, i64 nw,
                                                          const i32* __restrict__ part, const i32* __restrict__ dest,
                                                          const u64* __restrict__ mvp, i64* out) {
    const int lane = threadIdx.x & 63;
    const i64 w = block_id() * 4 + (threadIdx.x >> 6);
    if (w >= nw) return;
    const i64 x = k - 1 - w;
    const i32 i = moves[x];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s < LARGE_CUTOFF || e - s >= MASSIVE_CUTOFF) return;
    const i32 oi = part[i], ni = dest[i];
    i64 acc = 0;
    for (i64 base = s; base < e; base += 64) {  // wave-uniform loop control
        const i64 j = base + lane;
        if (j < e) acc += delta_term(col, W, mvp, j, oi, ni);
    }
    acc = wave_sum_i64(acc);
    if (lane == 0) out[x] = acc;
}
template <class WT>
__global__ void __launch_bounds__(1024) k_delta_uncut_block(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                            WT W, const i32* __restrict__ moves, i64 k, i64 nb,
                                                            const i32* __restrict__ part,
                                                            const i32* __restrict__ dest,
                                                            const u64* __restrict__ mvp, i64* out) {
    const i64 w = block_id();
    if (w >= nb) return;
    const i64 x = k - 1 - w;
    const i32 i = moves[x];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s < MASSIVE_CUTOFF) return;  // block-uniform
    const i32 oi = part[i], ni = dest[i];
    i64 acc = 0;
    for (i64 j = s + threadIdx.x; j < e; j += 1024) acc += delta_term(col, W, mvp, j, oi, ni);
    acc = block_sum_i64<1024>(acc);
    if (threadIdx.x == 0) out[x] = acc;
}

// The volume changes of a block's 256 movers are merged per cluster in an
// LDS table first (many movers of a sweep leave or join the same clusters),
// then each distinct cluster gets one HBM atomic; the sums are exact
// integers, so the order does not matter.
constexpr int AM_SLOTS = 1024;  // >= 2 keys per mover of a 256-thread block, at most half full
__global__ void __launch_bounds__(256) k_apply_moves(const i32* moves, i64 k, const i64* vw, i32* part,
                                                     const i32* dest, i64* vol) {
    __shared__ i32 lk[AM_SLOTS];
    __shared__ i64 lv[AM_SLOTS];
    lds_agg_clear<256, AM_SLOTS>(lk, lv);
    const i64 x = thread_item();
    if (x < k) {
        const i32 i = moves[x];
        const i32 p = part[i], b = dest[i];
        const i64 w = vw[i];
        if (b >= 0) {  // unresolved NEW_PART cannot reach here (p_conn is 0 for singletons)
            table_add_lds(lk, lv, AM_SLOTS - 1, p, -w);
            table_add_lds(lk, lv, AM_SLOTS - 1, b, w);
            part[i] = b;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < AM_SLOTS; q += 256) {
        const i32 key = lk[q];
        const i64 v = lv[q];
        if (key != EMPTY_KEY && v != 0) atomicAdd((unsigned long long*)&vol[key], (unsigned long long)v);
    }
}

// dest := NULL_PART for the movers and their neighbours (src/local_mover.cpp:717,798),
// same degree split as the uncut delta
__global__ void __launch_bounds__(256) k_reset_affected_thread(const i64* __restrict__ rp,
                                                               const i32* __restrict__ col,
                                                               const i32* __restrict__ moves, i64 k, i32* dest,
                                                               i32* mflag) {
    const i64 x = thread_item();
    if (x >= k) return;
    const i32 i = moves[x];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s >= LARGE_CUTOFF) return;
#pragma unroll 4
    for (i64 j = s; j < e; j++) dest[col[j]] = NULL_PART;
    dest[i] = NULL_PART;
    mflag[i] = 0;
}
__global__ void __launch_bounds__(256) k_reset_affected_wave(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                             const i32* __restrict__ moves, i64 k, i64 nw, i32* dest,
                                                             i32* mflag) {
    const int lane = threadIdx.x & 63;
    const i64 w = block_id() * 4 + (threadIdx.x >> 6);
    if (w >= nw) return;
    const i32 i = moves[k - 1 - w];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s < LARGE_CUTOFF || e - s >= MASSIVE_CUTOFF) return;
    for (i64 j = s + lane; j < e; j += 64) dest[col[j]] = NULL_PART;
    if (lane == 0) {
        dest[i] = NULL_PART;
        mflag[i] = 0;
    }
}
__global__ void __launch_bounds__(1024) k_reset_affected_block(const i64* __restrict__ rp,
                                                               const i32* __restrict__ col,
                                                               const i32* __restrict__ moves, i64 k, i64 nb,
                                                               i32* dest, i32* mflag) {
    const i64 w = block_id();
    if (w >= nb) return;
    const i32 i = moves[k - 1 - w];
    const i64 s = rp[i], e = rp[i + 1];
    if (e - s < MASSIVE_CUTOFF) return;
    for (i64 j = s + threadIdx.x; j < e; j += 1024) dest[col[j]] = NULL_PART;
    if (threadIdx.x == 0) {
        dest[i] = NULL_PART;
        mflag[i] = 0;
    }
}
__global__ void k_clear_mflag(const i32* moves, i64 k, i32* mflag) {
    const i64 x = thread_item();
    if (x >= k) return;
    mflag[moves[x]] = 0;
}
__global__ void k_set_mflag(const i32* moves, i64 k, i32* mflag) {
    const i64 x = thread_item();
    if (x >= k) return;
    mflag[moves[x]] = 1;
}

// perform_moves (src/local_mover.cpp:1241-1279).  moves is in order1 order,
// so movers by degree class: below LARGE_CUTOFF (a thread each), below
// MASSIVE_CUTOFF (a wave each), then the rest (a 1024-thread block each).
// n_large / n_massive = vertices of the level with degree >= LARGE_CUTOFF /
// MASSIVE_CUTOFF, which bound the movers of the last two classes.
// The moves themselves: hub-table updates (before part changes), the exact
// uncut delta of every mover into du[x], volume moves and part updates.  A
// hub table that runs out of slots sets s.scal[3].
static void apply_moves_core(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i64* vol, const i32* moves,
                             i64 k, i64 n_large, i64 n_massive, i64* du) {
    const i64 nw = std::min(k, n_large);
    const i64 nb = std::min(k, n_massive);
    const int tk = c.kt_begin(KS_MOVES);
    fill_i64(c, s.scal.p + 3, 1, 0);
    const i64 m7 = lp.nhub;
    // many movers: updating costs more than rebuilding the tables next sweep
    if (lp.hub_live && m7 > 0 && (c.hub_update_div <= 0 || k * c.hub_update_div > g.n)) lp.hub_live = false;
    const bool hub_upd = lp.hub_live && m7 > 0;
    if (hub_upd) {  // keep the hub tables current (before part changes)
        i64* stale = s.scal.p + 3;
        PGGC_WDISPATCH(g, {
            if (lp.packed) {
                k_hub_update_thread<true><<<grid_items(k, 256), 256, 0, c.stream>>>(
                    g.rp, g.col, W, moves, k, s.part, s.dest, lp.hidx, lp.goff, lp.hsm, nullptr, nullptr, lp.gtab, lp.hlist, lp.hcnt, stale);
                if (nw > 0)
                    k_hub_update_wave<true><<<grid_items(nw, 4), 256, 0, c.stream>>>(
                        g.rp, g.col, W, moves, k, nw, s.part, s.dest, lp.hidx, lp.goff, lp.hsm, nullptr, nullptr, lp.gtab, lp.hlist, lp.hcnt,
                        stale);
                if (nb > 0)
                    k_hub_update_block<true><<<grid_blocks(nb), 1024, 0, c.stream>>>(
                        g.rp, g.col, W, moves, k, nb, s.part, s.dest, lp.hidx, lp.goff, lp.hsm, nullptr, nullptr, lp.gtab, lp.hlist, lp.hcnt,
                        stale);
            } else {
                k_hub_update_thread<false><<<grid_items(k, 256), 256, 0, c.stream>>>(
                    g.rp, g.col, W, moves, k, s.part, s.dest, lp.hidx, lp.goff, lp.hsm, lp.gkeys, lp.gvals, nullptr, lp.hlist, lp.hcnt, stale);
                if (nw > 0)
                    k_hub_update_wave<false><<<grid_items(nw, 4), 256, 0, c.stream>>>(
                        g.rp, g.col, W, moves, k, nw, s.part, s.dest, lp.hidx, lp.goff, lp.hsm, lp.gkeys, lp.gvals, nullptr, lp.hlist, lp.hcnt,
                        stale);
                if (nb > 0)
                    k_hub_update_block<false><<<grid_blocks(nb), 1024, 0, c.stream>>>(
                        g.rp, g.col, W, moves, k, nb, s.part, s.dest, lp.hidx, lp.goff, lp.hsm, lp.gkeys, lp.gvals, nullptr, lp.hlist, lp.hcnt,
                        stale);
            }
        });
        PGGC_LAUNCH_CHECK();
    }
    {
        k_mv_pack<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.n, s.part, s.dest, s.mflag, s.mvp);
        PGGC_LAUNCH_CHECK();
        PGGC_WDISPATCH(g, {
            k_delta_uncut_thread<<<grid_items(k, 256), 256, 0, c.stream>>>(g.rp, g.col, W, moves, k, s.part, s.dest,
                                                                          s.mvp, du);
            if (nw > 0)
                k_delta_uncut_wave<<<grid_items(nw, 4), 256, 0, c.stream>>>(g.rp, g.col, W, moves, k, nw, s.part,
                                                                           s.dest, s.mvp, du);
            if (nb > 0)
                k_delta_uncut_block<<<grid_blocks(nb), 1024, 0, c.stream>>>(g.rp, g.col, W, moves, k, nb, s.part,
                                                                           s.dest, s.mvp, du);
        });
        PGGC_LAUNCH_CHECK();
    }
    k_apply_moves<<<grid_items(k, 256), 256, 0, c.stream>>>(moves, k, g.vw, s.part, s.dest, vol);
    c.kt_end(tk, 0.0);
    PGGC_LAUNCH_CHECK();
    if (tk >= 0) {  // timing runs only
        // per vertex the packed record (part, dest, flag read, 8 bytes
        // written); per mover row bounds, part, dest, the partial delta and
        // its sum, vertex weight, two volume updates and the part write; per
        // mover edge column, weight and the neighbour's record, plus the
        // column, weight and hub index again when hub tables are updated
        const int wb = g.wk == W_UNIT ? 0 : g.wk == W_I32 ? 4 : 8;
        Buf<i64> acc(c, 1);
        fill_i64(c, acc, 1, 0);
        k_list_bytes<<<grid_items(std::max<i64>(k, 1), 256), 256, 0, c.stream>>>(
            g.rp, moves, k, 72, 12 + wb + (hub_upd ? 8 + wb : 0), 20 * g.n, acc);
        PGGC_LAUNCH_CHECK();
        acct_add(c, c.kt_counter(tk), acc);
    }
}

static void perform_moves(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, RefineData& curr,
                          const i32* moves, i64 k, i64 n_large, i64 n_massive, bool& tables) {
    const i64 nw = std::min(k, n_large);
    const i64 nb = std::min(k, n_massive);
    {
        Buf<i64> du(c, std::max<i64>(k, 1));
        apply_moves_core(c, s, lp, g, curr.vol, moves, k, n_large, n_massive, du);
        sum_i64(c, du, k, s.scal);
    }
    if (!tables) {
        tables = true;
        fill_i32(c, s.dest, g.n, NULL_PART);
        k_clear_mflag<<<grid_items(k, 256), 256, 0, c.stream>>>(moves, k, s.mflag);
    } else {
        k_reset_affected_thread<<<grid_items(k, 256), 256, 0, c.stream>>>(g.rp, g.col, moves, k, s.dest, s.mflag);
        if (nw > 0)
            k_reset_affected_wave<<<grid_items(nw, 4), 256, 0, c.stream>>>(g.rp, g.col, moves, k, nw, s.dest,
                                                                           s.mflag);
        if (nb > 0)
            k_reset_affected_block<<<grid_blocks(nb), 1024, 0, c.stream>>>(g.rp, g.col, moves, k, nb, s.dest,
                                                                           s.mflag);
    }
    PGGC_LAUNCH_CHECK();
    sum_sq_u128(c, curr.vol, curr.label_count, s.scal.p + 1);
    const i64* h = c.read(s.scal, 4);
    if (h[3]) lp.hub_live = false;  // a hub table ran out of slots: rebuild next sweep
    curr.uncut += h[0];
    curr.obj = objective_value(curr.uncut, curr.lambda, (u64)h[1], (u64)h[2]);
}

// ------------------------------------------------------------- relabel
__global__ void k_relabel_vol(const i32* used, const i64* pos, i64 L, const i64* vol, i64* nvol) {
    const i64 l = thread_item();
    if (l >= L) return;
    if (used[l]) nvol[pos[l]] = vol[l];
}
__global__ void k_relabel_part(const i64* pos, i64 n, i32* part) {
    const i64 i = thread_item();
    if (i >= n) return;
    part[i] = (i32)pos[part[i]];
}

// relabel_contiguously (src/local_mover.cpp:100-129)
static void relabel(Ctx& c, MoveScratch& s, i64 n, i32* labels, RefineData& rd) {
    const i64 L = rd.label_count;
    fill_i32(c, s.flag, L, 0);
    k_mark_used<<<grid_items(n, 256), 256, 0, c.stream>>>(labels, n, s.flag);
    PGGC_LAUNCH_CHECK();
    Buf<i64> pos(c, L + 1), nvol(c, std::max<i64>(L, 1));
    scan_i32_to_i64(c, s.flag, pos, L);
    k_relabel_vol<<<grid_items(L, 256), 256, 0, c.stream>>>(s.flag, pos, L, rd.vol, nvol);
    PGGC_LAUNCH_CHECK();
    i64 t = c.read1(pos.p + L);
    copy_dev(c, rd.vol, nvol, t * sizeof(i64));
    k_relabel_part<<<grid_items(n, 256), 256, 0, c.stream>>>(pos, n, labels);
    PGGC_LAUNCH_CHECK();
    rd.label_count = (i32)t;
}

static void copy_state(Ctx& c, RefineData& dst, const RefineData& src) {
    if (src.label_count > dst.cap) throw Error(-4, "label capacity exceeded");
    copy_dev(c, dst.vol, src.vol, (size_t)src.label_count * sizeof(i64));
    dst.top_nnz = src.top_nnz;
    dst.uncut = src.uncut;
    dst.obj = src.obj;
    dst.label_count = src.label_count;
    dst.lambda = src.lambda;
}

// ------------------------------------------------------------- drivers
struct CurrVol {
    Buf<i64> v;
};

// ------------------------------------------------------------ sweep trace
__global__ void __launch_bounds__(256) k_count_active(const i32* dest, i64 n, i64* out) {
    const i64 i = thread_item();
    const i64 a = (i < n && dest[i] != NO_MOVE) ? 1 : 0;
    const i64 s = block_sum_i64<256>(a);
    if (threadIdx.x == 0 && s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
}

// Host side of the per-sweep trace: begin() counts the vertices the sweep
// will score and starts the clock, end() reports the record.  Both are
// no-ops unless a trace callback is set.
struct SweepTrace {
    Ctx& c;
    const DGraph& g;
    i64 phase, sweep = 0, active = 0;
    std::chrono::steady_clock::time_point t0;
    SweepTrace(Ctx& cc, const DGraph& gg, i64 ph) : c(cc), g(gg), phase(ph) {}
    void begin(const i32* dest, bool all) {
        if (!c.sweep_trace) return;
        if (all) {
            active = g.n;
        } else {
            Buf<i64> cnt(c, 1);
            fill_i64(c, cnt, 1, 0);
            if (g.n > 0) {
                k_count_active<<<grid_items(g.n, 256), 256, 0, c.stream>>>(dest, g.n, cnt);
                PGGC_LAUNCH_CHECK();
            }
            active = c.read1(cnt);
        }
        c.sync();
        t0 = std::chrono::steady_clock::now();
    }
    void end(i64 cands, i64 moves) {
        if (!c.sweep_trace) return;
        c.sync();
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        const int64_t rec[8] = {g.n, g.nnz, phase, sweep++, active, cands, moves, (int64_t)us};
        c.sweep_trace(c.sweep_user, rec, 8);
    }
};

static RefineData begin_curr(Ctx& c, const RefineData& best, Buf<i64>& cv, i64 n) {
    i64 cap = std::max<i64>(best.cap, n);
    cv = Buf<i64>(c, cap + 1);
    RefineData curr = best;
    curr.vol = cv;
    curr.cap = cap;
    copy_state(c, curr, best);
    return curr;
}

template <class PT>
static void local_move_t(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i32* best_part, RefineData& best,
                         bool is_initial, bool constrained, i32* cons, bool SA, MoveStats* stats, PT prio,
                         const i32* ids);

void local_move(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i32* best_part, RefineData& best,
                bool is_initial, bool constrained, i32* cons, bool SA, MoveStats* stats, const i32* ids) {
    if (ids) {
        local_move_t(c, s, lp, g, best_part, best, is_initial, constrained, cons, SA, stats, PrioInst{ids}, ids);
    } else {
        local_move_t(c, s, lp, g, best_part, best, is_initial, constrained, cons, SA, stats, PrioOne{}, nullptr);
    }
}

template <class PT>
static void local_move_t(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i32* best_part, RefineData& best,
                         bool is_initial, bool constrained, i32* cons, bool SA, MoveStats* stats, PT prio,
                         const i32* ids) {
    const i64 n = g.n;
    s.ensure(c, n);
    plan_level(c, g, lp);
    hub_tables_acquire(c, lp);
    lp.hub_live = false;  // the clustering changed since the last call
    c.mem_note("local move: plan ready");
    Buf<i64> cv;
    RefineData curr = begin_curr(c, best, cv, n);
    copy_dev(c, s.part, best_part, n * sizeof(i32));
    fill_i32(c, s.dest, n, NULL_PART);
    bool tables = !is_initial;
    const float ratios[2] = {0.75f, 0.25f};
    const int limits[2] = {4, 2};
    int iter = 0;
    const i32* ccons = constrained ? cons : nullptr;
    // the raw first sweep (every neighbour its own cluster) reads a
    // neighbour's cluster from part when the graph is renumbered
    const int raw_mode = c.raw_sweeps ? (ids ? 2 : 1) : 0;
    SweepTrace tr(c, g, is_initial ? 0 : 1);
    // active-set predictor: on levels of at least min_nnz nonzeros, from the
    // second sweep of a call that starts at singletons (whose first sweep
    // moves most vertices), from the first sweep otherwise, and only while
    // the previous sweep left at least n/16 candidates (below that the
    // active set is mostly the movers' neighbourhood and little is to gain)
    Gnn* G = (s.gnn && s.gnn->cfg.mode != GNN_OFF && g.nnz >= s.gnn->cfg.min_nnz) ? s.gnn : nullptr;
    const bool train = G && G->cfg.mode == GNN_TRAIN;
    if (G) gnn_begin(c, *G, n, train);
    i64 kprev = n;
    i64 npred = 0;  // predicted sweeps in this call
    for (int x = 0; x < 2; x++) {
        const float fr = SA ? ratios[x] : 0.0f;
        for (int cnt = 0; cnt < limits[x]; cnt++) {
            iter++;
            const LamOne lam{(float)curr.lambda};
            const bool use = G && !(is_initial && iter == 1) && kprev * 16 >= n &&
                             (c.gnn_max_sweeps <= 0 || npred < c.gnn_max_sweeps);
            const bool full = use && G->cfg.full_every > 0 && iter % G->cfg.full_every == 0;
            if (full) gnn_restore(c, *G, n, s.dest);
            if (use && !full) {
                gnn_predict(c, *G, lp, g, s.part, s.dest, s.gain, curr.vol, iter, iter == 1,
                            G->cfg.mode == GNN_PRUNE, train);
                npred++;
            }
            tr.begin(s.dest, iter == 1 && !use);
            score_all(c, s, lp, g, s.part, curr.vol, ccons, tables ? 0 : raw_mode, lam, fr, s.dest, s.gain);
            i64 b[2];
            i64 k = gather_candidates(c, s, lp, g, s.dest, s.cand, b);
            i64 km = k > 0 ? afterburner(c, s, g, s.cand, k, b, lam, prio, s.moves) : 0;
            if (km == 0 && G && gnn_restore(c, *G, n, s.dest)) {
                // nothing to move among the vertices let through: score the
                // pruned ones too, so the call ends as the plain method does
                score_all(c, s, lp, g, s.part, curr.vol, ccons, tables ? 0 : raw_mode, lam, fr, s.dest, s.gain);
                k = gather_candidates(c, s, lp, g, s.dest, s.cand, b);
                km = k > 0 ? afterburner(c, s, g, s.cand, k, b, lam, prio, s.moves) : 0;
            }
            if (stats) {
                stats->sweeps++;
                stats->cands += k;
            }
            kprev = k;
            if (G) gnn_stamp(c, *G, s.moves, km, iter);
            if (train && use && !full) gnn_backward(c, *G, lp, g, nullptr, iter, true);
            if (k == 0) {
                tr.end(0, 0);
                break;
            }
            if (km == 0) {
                tr.end(k, 0);
                break;
            }
            if (iter > 1 || !is_initial)
                set_new_cluster_ids(c, s, g, curr, s.moves, km, constrained ? cons : nullptr, ids);
            perform_moves(c, s, lp, g, curr, s.moves, km, g.n - lp.ord.off_large, g.n - lp.ord.off_large2, tables);
            if (stats) stats->moves += km;
            if (c.verbose >= 3)
                fprintf(stderr, "sweep n=%lld it=%d cands=%lld moves=%lld labels=%d uncut=%lld obj=%.17g\n",
                        (long long)n, iter, (long long)k, (long long)km, curr.label_count, (long long)curr.uncut,
                        curr.obj);
            if (curr.obj > best.obj) {
                copy_state(c, best, curr);
                copy_dev(c, best_part, s.part, n * sizeof(i32));
            }
            tr.end(k, km);
        }
    }
    relabel(c, s, n, best_part, best);
    hub_tables_settle(c, lp);
}

// afterburner_filter_strict (src/local_mover.cpp:131-213): exact sequential
// FP32 evaluation of each move given all earlier moves in list order.
__global__ void k_prio(const i32* moves, i64 k, i64* prio) {
    const i64 x = thread_item();
    if (x >= k) return;
    prio[moves[x]] = x;
}

template <class WT>
__global__ void __launch_bounds__(256) k_strict_gain(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                     const i64* __restrict__ vw, const i32* __restrict__ moves,
                                                     i64 k, const i64* __restrict__ prio,
                                                     const i32* __restrict__ part, const i32* __restrict__ dest,
                                                     const float* __restrict__ gain, float lam, float* sg) {
    const i64 x = thread_item();
    if (x >= k) return;
    const i32 i = moves[x];
    const i32 best = dest[i], p = part[i];
    const float multi = (float)vw[i] * lam;
    float oc = gain[i];
    const i64 pi = prio[i];
    for (i64 j = rp[i]; j < rp[i + 1]; j++) {
        const i32 v = col[j];
        const i64 wgt = W(j);
        if (prio[v] < pi) {
            i32 vp = dest[v];
            oc = oc - (float)((vp == p) ? wgt : 0);
            oc = oc + (float)((vp == best && best != NEW_PART) ? wgt : 0);
            vp = part[v];
            oc = oc + (float)((vp == p) ? wgt : 0);
            oc = oc - (float)((vp == best) ? wgt : 0);
        }
    }
    for (i64 jx = 0; jx < x; jx++) {
        const i32 v = moves[jx];
        i32 vp = dest[v];
        const float wg = -multi * (float)vw[v];
        oc = oc - ((vp == p) ? wg : 0.0f);
        oc = oc + ((vp == best && best != NEW_PART) ? wg : 0.0f);
        vp = part[v];
        oc = oc + ((vp == p) ? wg : 0.0f);
        oc = oc - ((vp == best) ? wg : 0.0f);
    }
    sg[x] = oc;
}

// sequential inclusive FP32 prefix sum and first arg-max (one thread)
__global__ void k_strict_scan(const i32* moves, i64 k, float* sg, float* gain, i64* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float run = 0.0f, bv = -1.0f;
    i64 bl = -1;
    for (i64 x = 0; x < k; x++) {
        run = run + sg[x];
        sg[x] = run;
        gain[moves[x]] = run;
        if (run > bv) {
            bv = run;
            bl = x;
        }
    }
    out[0] = (bl >= 0 && bv > 0.0f) ? bl + 1 : 0;
}

void local_move_strict(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i32* best_part, RefineData& best,
                       bool is_initial, bool constrained, i32* cons, MoveStats* stats) {
    const i64 n = g.n;
    s.ensure(c, n);
    plan_level(c, g, lp);
    hub_tables_acquire(c, lp);
    lp.hub_live = false;  // the clustering changed since the last call
    Buf<i64> cv;
    RefineData curr = begin_curr(c, best, cv, n);
    copy_dev(c, s.part, best_part, n * sizeof(i32));
    fill_i32(c, s.dest, n, NULL_PART);
    bool tables = !is_initial;
    const i32* ccons = constrained ? cons : nullptr;
    Buf<float> sg(c, std::max<i64>(n, 1));
    SweepTrace trc(c, g, 2);
    for (int it = 0; it < 6; it++) {
        const float lam = (float)curr.lambda;
        trc.begin(s.dest, true);
        score_all(c, s, lp, g, s.part, curr.vol, ccons, !tables && c.raw_sweeps, LamOne{lam}, 0.0f, s.dest, s.gain);
        i64 b[2];
        i64 k = gather_candidates(c, s, lp, g, s.dest, s.moves, b);
        if (stats) {
            stats->sweeps++;
            stats->cands += k;
        }
        if (k == 0) {
            trc.end(0, 0);
            break;
        }
        if (it > 0 || !is_initial) set_new_cluster_ids(c, s, g, curr, s.moves, k, constrained ? cons : nullptr, nullptr);
        fill_i64(c, s.prio, n, n + 1);
        k_prio<<<grid_items(k, 256), 256, 0, c.stream>>>(s.moves, k, s.prio);
        PGGC_LAUNCH_CHECK();
        PGGC_WDISPATCH(g, (k_strict_gain<<<grid_items(k, 256), 256, 0, c.stream>>>(
                              g.rp, g.col, W, g.vw, s.moves, k, s.prio, s.part, s.dest, s.gain, lam, sg)));
        PGGC_LAUNCH_CHECK();
        k_strict_scan<<<1, 64, 0, c.stream>>>(s.moves, k, sg, s.gain, s.scal);
        PGGC_LAUNCH_CHECK();
        i64 tr = c.read1(s.scal);
        if (tr == 0) {
            trc.end(k, 0);
            break;
        }
        k_set_mflag<<<grid_items(tr, 256), 256, 0, c.stream>>>(s.moves, tr, s.mflag);
        PGGC_LAUNCH_CHECK();
        perform_moves(c, s, lp, g, curr, s.moves, tr, g.n - lp.ord.off_large, g.n - lp.ord.off_large2, tables);
        if (stats) stats->moves += tr;
        if (curr.obj > best.obj) {
            copy_state(c, best, curr);
            copy_dev(c, best_part, s.part, n * sizeof(i32));
        }
        fill_i32(c, s.dest, n, NULL_PART);
        trc.end(k, tr);
    }
    relabel(c, s, n, best_part, best);
    hub_tables_settle(c, lp);
}

i64 sweep_once(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, const i32* part, const i64* vol, i64 L,
               double lambda, float ratio, bool raw, const i32* cons, i32* dest, float* gain, i32* moves_out,
               i64* ncand) {
    const i64 n = g.n;
    (void)L;
    s.ensure(c, n);
    plan_level(c, g, lp);
    hub_tables_acquire(c, lp);
    lp.hub_live = false;
    copy_dev(c, s.part, part, n * sizeof(i32));
    copy_dev(c, s.dest, dest, n * sizeof(i32));
    copy_dev(c, s.gain, gain, n * sizeof(float));
    const float lam = (float)lambda;
    score_all(c, s, lp, g, s.part, vol, cons, raw, LamOne{lam}, ratio, s.dest, s.gain);
    i64 b[2];
    i64 k = gather_candidates(c, s, lp, g, s.dest, s.cand, b);
    *ncand = k;
    i64 km = 0;
    if (k > 0) {
        km = afterburner(c, s, g, s.cand, k, b, LamOne{lam}, PrioOne{}, moves_out);
        k_clear_mflag<<<grid_items(km, 256), 256, 0, c.stream>>>(moves_out, km, s.mflag);
        PGGC_LAUNCH_CHECK();
    }
    copy_dev(c, dest, s.dest, n * sizeof(i32));
    copy_dev(c, gain, s.gain, n * sizeof(float));
    c.sync();
    hub_tables_settle(c, lp);
    return km;
}


// ================================================================ batches
// A batch of graphs as one disjoint union (see BatchLevel): the sweep
// kernels run once over the union with per-instance lambda, ratio and
// priority ids; the control flow of local_move runs per instance on the host
// from per-instance counts read back once per sweep.
__global__ void k_batch_ids(i64 n, const i64* __restrict__ voff, i64 K, i32* __restrict__ inst,
                            i32* __restrict__ loc) {
    const i64 v = thread_item();
    if (v >= n) return;
    const i64 k = upper_row(voff, 0, K, v);  // last instance starting at or before v (skips empty ones)
    inst[v] = (i32)k;
    loc[v] = (i32)(v - voff[k]);
}

void BatchLevel::build(Ctx& c, const std::vector<i64>& offsets) {
    voff = offsets;
    K = (i64)voff.size() - 1;
    const i64 n = voff[K];
    dvoff = Buf<i64>(c, K + 1);
    copy_h2d(c, dvoff.p, voff.data(), (K + 1) * sizeof(i64));
    inst = Buf<i32>(c, std::max<i64>(n, 1));
    loc = Buf<i32>(c, std::max<i64>(n, 1));
    if (n > 0) {
        k_batch_ids<<<grid_items(n, 256), 256, 0, c.stream>>>(n, dvoff, K, inst, loc);
        PGGC_LAUNCH_CHECK();
    }
    c.sync();
}

template <class T>
static void upload(Ctx& c, T* dst, const std::vector<T>& src) {
    copy_h2d(c, dst, src.data(), src.size() * sizeof(T));
}
template <class T>
static void download(Ctx& c, std::vector<T>& dst, const T* src, size_t count) {
    dst.resize(count);
    if (count == 0) return;
    PGGC_HIP(hipMemcpyAsync(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

__global__ void k_inst_count(const i32* __restrict__ list, i64 k, const i32* __restrict__ owner, i64* cnt) {
    const i64 x = thread_item();
    if (x >= k) return;
    atomicAdd((unsigned long long*)&cnt[owner[list[x]]], 1ull);
}
// vertices of instances outside this sweep are never scored
__global__ void k_inst_gate(i64 n, const i32* __restrict__ inst, const i32* __restrict__ on, i32* dest) {
    const i64 v = thread_item();
    if (v >= n) return;
    if (!on[inst[v]]) dest[v] = NO_MOVE;
}
__global__ void k_inst_fill(i64 n, const i32* __restrict__ inst, const i32* __restrict__ mask, i32* a, i32 val) {
    const i64 v = thread_item();
    if (v >= n) return;
    if (mask[inst[v]]) a[v] = val;
}
__global__ void k_inst_add(const i32* __restrict__ moves, i64 k, const i32* __restrict__ inst,
                           const i64* __restrict__ du, i64* acc) {
    const i64 x = thread_item();
    if (x >= k) return;
    atomicAdd((unsigned long long*)&acc[inst[moves[x]]], (unsigned long long)du[x]);
}
// exact 128-bit sum of squared volumes over each masked instance's labels
// [loff[k], loff[k] + L[k]); one block per instance, out[2k] = high word
__global__ void __launch_bounds__(256) k_inst_sumsq(const i64* __restrict__ vol, const i64* __restrict__ loff,
                                                    const i64* __restrict__ L, const i32* __restrict__ mask, i64 K,
                                                    u64* out) {
    const i64 k = block_id();
    if (k >= K) return;
    u64 lo = 0, hi = 0;
    if (mask[k]) {
        const i64 e = loff[k] + L[k];
        for (i64 l = loff[k] + threadIdx.x; l < e; l += 256) {
            const i64 v = vol[l];
            const u64 a = (u64)(v < 0 ? -v : v);
            const u64 plo = a * a, phi = __umul64hi(a, a);
            lo += plo;
            hi += phi + (lo < plo ? 1 : 0);
        }
    }
    __shared__ u64 slo[256], shi[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const u64 a = slo[threadIdx.x], b = slo[threadIdx.x + o];
            slo[threadIdx.x] = a + b;
            shi[threadIdx.x] = shi[threadIdx.x] + shi[threadIdx.x + o] + (a + b < a ? 1 : 0);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * k] = shi[0];
        out[2 * k + 1] = slo[0];
    }
}
__global__ void k_inst_copy_v(i64 n, const i32* __restrict__ inst, const i32* __restrict__ mask,
                              const i32* __restrict__ src, i32* dst) {
    const i64 v = thread_item();
    if (v >= n) return;
    if (mask[inst[v]]) dst[v] = src[v];
}
// labels l < loff[k] + L[k] of masked instances
__global__ void k_inst_copy_l(i64 nl, const i32* __restrict__ linst, const i64* __restrict__ loff,
                              const i64* __restrict__ L, const i32* __restrict__ mask, const i64* __restrict__ src,
                              i64* dst) {
    const i64 l = thread_item();
    if (l >= nl) return;
    const i32 k = linst[l];
    if (mask[k] && l - loff[k] < L[k]) dst[l] = src[l];
}
__global__ void k_unused_inst(i64 nl, const i32* __restrict__ used, const i32* __restrict__ linst,
                              const i64* __restrict__ loff, const i64* __restrict__ L, i32* flag) {
    const i64 l = thread_item();
    if (l >= nl) return;
    const i32 k = linst[l];
    flag[l] = (used[l] == 0 && l - loff[k] < L[k]) ? 1 : 0;
}
__global__ void k_inst_keys(const i32* __restrict__ list, i64 k, const i32* __restrict__ inst, u64* keys) {
    const i64 x = thread_item();
    if (x >= k) return;
    keys[x] = (u64)(u32)inst[list[x]];
}
// x-th NEW_PART mover of instance q (movers grouped by instance, in move
// order) takes the instance's x-th unused label, then labels appended after
// its label count (src/local_mover.cpp:215-274)
__global__ void k_assign_new_batch(const i32* __restrict__ np, i64 K_, const i32* __restrict__ inst,
                                   const i64* __restrict__ npstart, const i64* __restrict__ ustart,
                                   const i64* __restrict__ acnt, const i32* __restrict__ unused,
                                   const i64* __restrict__ loff, const i64* __restrict__ L, i32* dest) {
    const i64 x = thread_item();
    if (x >= K_) return;
    const i32 v = np[x];
    const i32 q = inst[v];
    const i64 r = x - npstart[q];
    dest[v] = r < acnt[q] ? unused[ustart[q] + r] : (i32)(loff[q] + L[q] + (r - acnt[q]));
}
__global__ void k_mark_used_inst(const i32* __restrict__ part, i64 n, const i32* __restrict__ inst,
                                 const i32* __restrict__ mask, i32* used) {
    const i64 v = thread_item();
    if (v >= n) return;
    if (mask[inst[v]]) used[part[v]] = 1;
}
__global__ void k_relabel_vol_inst(i64 nl, const i32* __restrict__ used, const i64* __restrict__ pos,
                                   const i32* __restrict__ linst, const i64* __restrict__ loff,
                                   const i64* __restrict__ vol, i64* nvol) {
    const i64 l = thread_item();
    if (l >= nl) return;
    if (!used[l]) return;
    const i32 k = linst[l];
    nvol[loff[k] + pos[l] - pos[loff[k]]] = vol[l];
}
__global__ void k_relabel_part_inst(i64 n, const i32* __restrict__ inst, const i32* __restrict__ mask,
                                    const i64* __restrict__ pos, const i64* __restrict__ loff, i32* part) {
    const i64 v = thread_item();
    if (v >= n) return;
    const i32 k = inst[v];
    if (mask[k]) part[v] = (i32)(loff[k] + pos[part[v]] - pos[loff[k]]);
}
__global__ void k_inst_newcount(i64 K, const i64* __restrict__ pos, const i64* __restrict__ loff,
                                const i64* __restrict__ L, i64* out) {
    const i64 k = thread_item();
    if (k >= K) return;
    out[k] = pos[loff[k] + L[k]] - pos[loff[k]];
}

static std::vector<i64> label_counts(const std::vector<RefineData>& rd) {
    std::vector<i64> L(rd.size());
    for (size_t k = 0; k < rd.size(); k++) L[k] = rd[k].label_count;
    return L;
}

// set_new_cluster_ids per instance over the batch's move list
static void set_new_cluster_ids_batch(Ctx& c, MoveScratch& s, const DGraph& g, const BatchLevel& bl,
                                      const BatchState& st, std::vector<RefineData>& cur, i64* vol,
                                      const i32* moves, i64 k) {
    const i64 K = bl.K, NL = st.loff[K];
    k_new_flags<<<grid_items(k, 256), 256, 0, c.stream>>>(moves, k, s.dest, s.flag);
    PGGC_LAUNCH_CHECK();
    const i64 KN = compact_flags(c, s.flag, moves, k, s.tmp, s.scal);
    if (KN == 0) return;
    Buf<i64> Ld(c, K);
    upload(c, Ld.p, label_counts(cur));
    fill_i32(c, s.flag, NL, 0);
    k_mark_used<<<grid_items(g.n, 256), 256, 0, c.stream>>>(s.part, g.n, s.flag);
    Buf<i32> uf(c, std::max<i64>(NL, 1));
    k_unused_inst<<<grid_items(NL, 256), 256, 0, c.stream>>>(NL, s.flag, st.linst, st.dloff, Ld, uf);
    PGGC_LAUNCH_CHECK();
    const i64 A = compact_flags(c, uf, nullptr, NL, s.tmp2, s.scal);
    Buf<i64> acnt(c, K), kcnt(c, K);
    fill_i64(c, acnt, K, 0);
    fill_i64(c, kcnt, K, 0);
    if (A > 0) k_inst_count<<<grid_items(A, 256), 256, 0, c.stream>>>(s.tmp2, A, st.linst, acnt);
    k_inst_count<<<grid_items(KN, 256), 256, 0, c.stream>>>(s.tmp, KN, bl.inst, kcnt);
    PGGC_LAUNCH_CHECK();
    {  // group the NEW_PART movers by instance, keeping move order inside each
        Buf<u64> keys(c, KN);
        k_inst_keys<<<grid_items(KN, 256), 256, 0, c.stream>>>(s.tmp, KN, bl.inst, keys);
        PGGC_LAUNCH_CHECK();
        int bits = 1;
        while (((i64)1 << bits) < K) bits++;
        radix_sort_pairs(c, keys, s.tmp.p, KN, bits);
    }
    std::vector<i64> ha, hk;
    download(c, ha, acnt.p, K);
    download(c, hk, kcnt.p, K);
    std::vector<i64> us(K), ns(K);
    i64 u0 = 0, n0 = 0;
    for (i64 q = 0; q < K; q++) {
        us[q] = u0;
        ns[q] = n0;
        u0 += ha[q];
        n0 += hk[q];
    }
    Buf<i64> dus(c, K), dns(c, K);
    upload(c, dus.p, us);
    upload(c, dns.p, ns);
    k_assign_new_batch<<<grid_items(KN, 256), 256, 0, c.stream>>>(s.tmp, KN, bl.inst, dns, dus, acnt, s.tmp2,
                                                                  st.dloff, Ld, s.dest);
    PGGC_LAUNCH_CHECK();
    for (i64 q = 0; q < K; q++) {
        const i64 need = hk[q] - ha[q];
        if (need <= 0) continue;
        const i64 L = cur[q].label_count;
        if (L + need > st.loff[q + 1] - st.loff[q]) throw Error(-4, "label capacity exceeded while assigning new clusters");
        fill_i64(c, vol + st.loff[q] + L, need, 0);
        cur[q].label_count = (i32)(L + need);
    }
}

// relabel_contiguously (src/local_mover.cpp:100-129) of every active instance
static void relabel_batch(Ctx& c, MoveScratch& s, const BatchLevel& bl, i32* labels, BatchState& st,
                          const std::vector<char>& active) {
    const i64 K = bl.K, NL = st.loff[K], n = bl.voff[K];
    std::vector<i32> ha(K);
    for (i64 q = 0; q < K; q++) ha[q] = active[q] ? 1 : 0;
    Buf<i32> act(c, K);
    upload(c, act.p, ha);
    Buf<i64> Ld(c, K);
    upload(c, Ld.p, label_counts(st.rd));
    fill_i32(c, s.flag, NL, 0);
    if (n > 0) k_mark_used_inst<<<grid_items(n, 256), 256, 0, c.stream>>>(labels, n, bl.inst, act, s.flag);
    PGGC_LAUNCH_CHECK();
    Buf<i64> pos(c, NL + 1), nvol(c, std::max<i64>(NL, 1)), nl(c, K);
    scan_i32_to_i64(c, s.flag, pos, NL);
    k_relabel_vol_inst<<<grid_items(NL, 256), 256, 0, c.stream>>>(NL, s.flag, pos, st.linst, st.dloff, st.vol, nvol);
    k_inst_newcount<<<grid_items(K, 256), 256, 0, c.stream>>>(K, pos, st.dloff, Ld, nl);
    k_inst_copy_l<<<grid_items(NL, 256), 256, 0, c.stream>>>(NL, st.linst, st.dloff, nl, act, nvol, st.vol);
    if (n > 0) k_relabel_part_inst<<<grid_items(n, 256), 256, 0, c.stream>>>(n, bl.inst, act, pos, st.dloff, labels);
    PGGC_LAUNCH_CHECK();
    std::vector<i64> hn;
    download(c, hn, nl.p, K);
    for (i64 q = 0; q < K; q++)
        if (active[q]) st.rd[q].label_count = (i32)hn[q];
}

// per-vertex lambda and filter ratio of a batch sweep (LamVert's tables)
__global__ void k_vertex_lam(i64 n, const i32* __restrict__ inst, const float* __restrict__ lam,
                             const float* __restrict__ ratio, float* __restrict__ vlam, float* __restrict__ vrat) {
    const i64 v = thread_item();
    if (v >= n) return;
    const i32 k = inst[v];
    vlam[v] = lam[k];
    vrat[v] = ratio[k];
}

void local_move_batch(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, const BatchLevel& bl,
                      i32* best_part, BatchState& best, bool is_initial, const std::vector<char>& active,
                      std::vector<MoveStats>& stats) {
    const i64 n = g.n, K = bl.K, NL = best.loff[K];
    s.ensure(c, std::max(n, NL));
    plan_level(c, g, lp);
    hub_tables_acquire(c, lp);
    lp.hub_live = false;
    Buf<i64> cvol(c, std::max<i64>(NL, 1));
    copy_dev(c, cvol, best.vol, NL * sizeof(i64));
    std::vector<RefineData> cur = best.rd;
    copy_dev(c, s.part, best_part, n * sizeof(i32));
    fill_i32(c, s.dest, n, NULL_PART);
    Buf<i32> on(c, K), full(c, K), imp(c, K);
    Buf<float> rat(c, K);
    Buf<i64> cnt2(c, 2 * K), acc(c, K), Ld(c, K);
    Buf<u64> sq(c, 2 * K);
    const float ratios[2] = {0.75f, 0.25f};
    const int limits[2] = {4, 2};
    std::vector<int> phase(K, 0), cnt(K, 0), iter(K, 0);
    std::vector<char> tables(K, is_initial ? 0 : 1), done(K);
    for (i64 q = 0; q < K; q++) done[q] = !active[q];
    Buf<float> vlam(c, std::max<i64>(n, 1)), vrat(c, std::max<i64>(n, 1));
    const LamVert lam{vlam.p, vrat.p};
    const PrioInst prio{bl.loc.p};
    const i64 n_large = g.n - lp.ord.off_large, n_massive = g.n - lp.ord.off_large2;
    std::vector<i32> hon(K), hfull(K), himp(K);
    std::vector<float> hrat(K);
    std::vector<i64> hc, hacc;
    std::vector<u64> hsq;
    while (true) {
        bool any = false;
        for (i64 q = 0; q < K; q++) {
            hon[q] = done[q] ? 0 : 1;
            any = any || !done[q];
            hrat[q] = ratios[std::min(phase[q], 1)];
        }
        if (!any) break;
        upload(c, on.p, hon);
        upload(c, rat.p, hrat);
        if (n > 0) {
            k_inst_gate<<<grid_items(n, 256), 256, 0, c.stream>>>(n, bl.inst, on, s.dest);
            k_vertex_lam<<<grid_items(n, 256), 256, 0, c.stream>>>(n, bl.inst, best.dlam, rat, vlam, vrat);
            PGGC_LAUNCH_CHECK();
        }
        // table scoring throughout: on a singleton clustering it gives what
        // the raw first sweep gives (every neighbour its own cluster)
        score_all(c, s, lp, g, s.part, cvol, nullptr, false, lam, 0.0f, s.dest, s.gain);
        i64 b[2];
        const i64 kc = gather_candidates(c, s, lp, g, s.dest, s.cand, b);
        i64 km = 0;
        if (kc > 0) {
            km = afterburner(c, s, g, s.cand, kc, b, lam, prio, s.moves);
        }
        fill_i64(c, cnt2, 2 * K, 0);
        if (kc > 0) k_inst_count<<<grid_items(kc, 256), 256, 0, c.stream>>>(s.cand, kc, bl.inst, cnt2);
        if (km > 0) k_inst_count<<<grid_items(km, 256), 256, 0, c.stream>>>(s.moves, km, bl.inst, cnt2.p + K);
        PGGC_LAUNCH_CHECK();
        download(c, hc, cnt2.p, 2 * K);
        for (i64 q = 0; q < K; q++) {
            if (!hon[q]) continue;
            iter[q]++;
            stats[q].sweeps++;
            stats[q].cands += hc[q];
        }
        if (km > 0) {
            set_new_cluster_ids_batch(c, s, g, bl, best, cur, cvol, s.moves, km);
            {
                Buf<i64> du(c, km);
                apply_moves_core(c, s, lp, g, cvol, s.moves, km, n_large, n_massive, du);
                fill_i64(c, acc, K, 0);
                k_inst_add<<<grid_items(km, 256), 256, 0, c.stream>>>(s.moves, km, bl.inst, du, acc);
                PGGC_LAUNCH_CHECK();
            }
            // dest := NULL_PART for movers and their neighbours, and for
            // every vertex of an instance whose first moves these are
            const i64 nw = std::min(km, n_large), nb = std::min(km, n_massive);
            k_reset_affected_thread<<<grid_items(km, 256), 256, 0, c.stream>>>(g.rp, g.col, s.moves, km, s.dest,
                                                                              s.mflag);
            if (nw > 0)
                k_reset_affected_wave<<<grid_items(nw, 4), 256, 0, c.stream>>>(g.rp, g.col, s.moves, km, nw, s.dest,
                                                                               s.mflag);
            if (nb > 0)
                k_reset_affected_block<<<grid_blocks(nb), 1024, 0, c.stream>>>(g.rp, g.col, s.moves, km, nb, s.dest,
                                                                               s.mflag);
            PGGC_LAUNCH_CHECK();
            bool anyfull = false;
            for (i64 q = 0; q < K; q++) {
                const bool moved = hc[K + q] > 0;
                hfull[q] = (moved && !tables[q]) ? 1 : 0;
                anyfull = anyfull || hfull[q];
                if (moved) tables[q] = 1;
                himp[q] = moved ? 1 : 0;  // instances whose objective changes
            }
            if (anyfull) {
                upload(c, full.p, hfull);
                k_inst_fill<<<grid_items(n, 256), 256, 0, c.stream>>>(n, bl.inst, full, s.dest, NULL_PART);
                PGGC_LAUNCH_CHECK();
            }
            upload(c, Ld.p, label_counts(cur));
            upload(c, imp.p, himp);
            k_inst_sumsq<<<grid_blocks(K), 256, 0, c.stream>>>(cvol, best.dloff, Ld, imp, K, sq);
            PGGC_LAUNCH_CHECK();
            const i64 stale = c.read1(s.scal.p + 3);
            if (stale) lp.hub_live = false;  // a hub table ran out of slots: rebuild next sweep
            download(c, hacc, acc.p, K);
            download(c, hsq, sq.p, 2 * K);
            bool anyimp = false;
            for (i64 q = 0; q < K; q++) {
                const bool moved = himp[q] != 0;
                himp[q] = 0;
                if (!moved) continue;
                stats[q].moves += hc[K + q];
                cur[q].uncut += hacc[q];
                cur[q].obj = objective_value(cur[q].uncut, cur[q].lambda, hsq[2 * q], hsq[2 * q + 1]);
                if (c.verbose >= 3)
                    fprintf(stderr, "sweep n=%lld it=%d cands=%lld moves=%lld labels=%d uncut=%lld obj=%.17g inst=%lld\n",
                            (long long)(bl.voff[q + 1] - bl.voff[q]), iter[q], (long long)hc[q],
                            (long long)hc[K + q], cur[q].label_count, (long long)cur[q].uncut, cur[q].obj,
                            (long long)q);
                if (cur[q].obj > best.rd[q].obj) {
                    best.rd[q] = cur[q];
                    himp[q] = 1;
                    anyimp = true;
                }
            }
            if (anyimp) {  // best state of the improved instances (copy_state)
                upload(c, imp.p, himp);
                k_inst_copy_v<<<grid_items(n, 256), 256, 0, c.stream>>>(n, bl.inst, imp, s.part, best_part);
                k_inst_copy_l<<<grid_items(NL, 256), 256, 0, c.stream>>>(NL, best.linst, best.dloff, Ld, imp, cvol,
                                                                        best.vol);
                PGGC_LAUNCH_CHECK();
            }
        }
        for (i64 q = 0; q < K; q++) {  // ratio schedule: 0.75 x4, then 0.25 x2; a sweep without moves ends a phase
            if (!hon[q]) continue;
            if (hc[q] == 0 || hc[K + q] == 0) {
                phase[q]++;
                cnt[q] = 0;
            } else if (++cnt[q] == limits[phase[q]]) {
                phase[q]++;
                cnt[q] = 0;
            }
            if (phase[q] >= 2) done[q] = 1;
        }
    }
    relabel_batch(c, s, bl, best_part, best, active);
    hub_tables_settle(c, lp);
}

}  // namespace pggc
