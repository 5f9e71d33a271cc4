// Multilevel drivers (src/clustering_methods.cpp), the run_clustering outer
// loop (app/cluster_main.cpp:101-165) and the C ABI declared in pggc.h.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"
#include "device.h"
#include "gen.h"
#include "gnn.h"
#include "hash.h"
#include "leiden.h"
#include "localmove.h"
#include "meme.h"
#include "noalloc.h"
#include "pggc.h"
#include "renum.h"

using namespace pggc;

struct pggc_devgraph {
    DevCSR g;
};

struct pggc_ctx {
    Ctx c;
    MoveScratch ms;
    Gnn gnn;
    // worker contexts of pggc_cluster_batch, each on a stream of its own
    std::vector<pggc_ctx*> workers;
};

namespace pggc {

struct RunStats {
    MoveStats mv;
    i64 levels = 0;
    i64 total_nnz = 0;
    double contract_s = 0.0;  // wall time in contraction (clustering_methods.cpp:117,220)
    double refine_s = 0.0;    // wall time in Leiden refinement (:89)
};

static double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- kernels
__global__ void k_project(i64 n, const i32* coarse, i32* fine) {
    const i64 x = thread_item();
    if (x >= n) return;
    fine[x] = coarse[fine[x]];
}
// highest fine id per coarse id; a block's 256 vertices are merged per coarse
// id in LDS first, then each distinct id gets one HBM atomicMax
constexpr int DW_SLOTS = 512;
__global__ void __launch_bounds__(256) k_down_winner(i64 n, const i32* map, i32* win) {
    __shared__ i32 lk[DW_SLOTS], lm[DW_SLOTS];
    for (int q = threadIdx.x; q < DW_SLOTS; q += 256) {
        lk[q] = EMPTY_KEY;
        lm[q] = -1;
    }
    __syncthreads();
    const i64 i = thread_item();
    if (i < n) {
        const i32 u = map[i];
        u32 h = slot_hash(u, DW_SLOTS - 1);
        while (true) {  // at most 256 keys in 512 slots
            const i32 old = atomicCAS(&lk[h], EMPTY_KEY, u);
            if (old == EMPTY_KEY || old == u) break;
            h = (h + 1) & (DW_SLOTS - 1);
        }
        atomicMax(&lm[h], (i32)i);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < DW_SLOTS; q += 256)
        if (lk[q] != EMPTY_KEY) atomicMax(&win[lk[q]], lm[q]);
}
__global__ void k_down_gather(i64 nc, const i32* win, const i32* in, i32* out) {
    const i64 c = thread_item();
    if (c >= nc) return;
    const i32 w = win[c];
    if (w >= 0) out[c] = in[w];
}
__global__ void k_downsample(i64 n, const i32* map, const i32* in, i32* out) {
    const i64 i = thread_item();
    if (i >= n) return;
    out[map[i]] = in[i];
}
template <class WT>
__global__ void __launch_bounds__(256) k_wdeg(const i64* rp, WT W, i64 n, i64* out) {
    const int lane = threadIdx.x & 63;
    const i64 i = block_id() * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    i64 acc = 0;
    const i64 e = rp[i + 1];
    for (i64 base = rp[i]; base < e; base += 64) {
        const i64 j = base + lane;
        if (j < e) acc += W(j);
    }
    acc = wave_sum_i64(acc);
    if (lane == 0) out[i] = acc;
}
__global__ void k_deg(const i64* rp, i64 n, i64* out) {
    const i64 i = thread_item();
    if (i >= n) return;
    out[i] = rp[i + 1] - rp[i];
}

static void project(Ctx& c, i64 n, const i32* coarse, i32* fine) {
    k_project<<<grid_items(n, 256), 256, 0, c.stream>>>(n, coarse, fine);
    PGGC_LAUNCH_CHECK();
}

// downsample (src/clustering_methods.cpp:40-45); when members of a cluster
// carry different values the highest vertex id wins.
static void downsample_max(Ctx& c, i64 n, const i32* map, i64 nc, const i32* in, i32* out) {
    Buf<i32> win(c, std::max<i64>(nc, 1));
    fill_i32(c, win, nc, -1);
    k_down_winner<<<grid_items(n, 256), 256, 0, c.stream>>>(n, map, win);
    k_down_gather<<<grid_items(nc, 256), 256, 0, c.stream>>>(nc, win, in, out);
    PGGC_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- drivers
struct Level {
    DGraph g;
    std::unique_ptr<OwnedGraph> own;
    std::unique_ptr<LevelPlan> plan;
};

static Level top_level(const DGraph& g) {
    Level l;
    l.g = g;
    l.plan.reset(new LevelPlan());
    return l;
}
// On the way back up the hierarchy a level is done once its clustering is
// projected onto the finer one: its graph, plan and clustering are freed,
// so the finer levels' local moves have that memory (the whole hierarchy of
// a billion-edge graph does not fit next to the top level's hub tables).
static void release_level(Level& l, Buf<i32>& part) {
    l.own.reset();
    l.plan.reset();
    part.reset();
}

static Level owned_level(std::unique_ptr<OwnedGraph> og) {
    Level l;
    l.g = og->view();
    l.own = std::move(og);
    l.plan.reset(new LevelPlan());
    return l;
}

// consecutive non-contracting refinements tolerated at one level
constexpr int LEIDEN_MAX_STALLS = 8;

// The top level renumbered by degree for the local mover (results are the
// caller's order's, bit for bit): always with the renumber knob at 1, with
// it at 2 (the default) when the copy serves several top-level passes
// (reuse), never at 0 or with the active-set predictor on, whose features
// follow the vertex order.
static Renum* top_renum(Ctx& c, MoveScratch& ms, const DGraph& g, Renum& r, bool reuse) {
    if (c.renumber == 0 || (c.renumber == 2 && !reuse) || (ms.gnn && ms.gnn->cfg.mode != GNN_OFF)) return nullptr;
    renum_build(c, g, r);
    return &r;
}

// local_move at one level; rn (when set) is the top level renumbered
static void move_level(Ctx& c, MoveScratch& ms, LevelPlan& lp, const DGraph& g, i32* part, RefineData& rfd,
                       bool is_initial, bool constrained, i32* cons, bool SA, RunStats& rs, Renum* rn) {
    if (rn) local_move_renum(c, ms, *rn, part, rfd, is_initial, constrained, cons, SA, &rs.mv);
    else local_move(c, ms, lp, g, part, rfd, is_initial, constrained, cons, SA, &rs.mv);
}

// Contraction of a level; at the top level with rn set, through the
// renumbered copy (the same coarse graph: coarse ids are the labels, coarse
// rows are sorted and their weights are integer sums), whose rows gather
// the map with more locality.  rn->part is the local mover's scratch and
// free here.
static void contract_level(Ctx& c, const DGraph& g, const i32* map, i64 nc, OwnedGraph& og, Renum* rn) {
    if (!rn) {
        contract(c, g, map, nc, og);
    } else {
        renum_gather(c, *rn, map, rn->part);
        contract(c, rn->g, rn->part, nc, og);
    }
    narrow_weights(c, g, og);
}

// louvain_part<constrained> (src/clustering_methods.cpp:139-222)
static void louvain_part(Ctx& c, MoveScratch& ms, const DGraph& top, RefineData& rfd, bool constrained,
                         i32* constraint, i32* part_out, RunStats& rs, LevelPlan* top_plan, Renum* rn) {
    if (constrained) {  // cluster_data::reset (interface/cluster_data.hpp:49-55)
        copy_dev(c, rfd.vol, top.vw, top.n * sizeof(i64));
        rfd.uncut = 0;
        rfd.label_count = (i32)top.n;
        update_objective(c, rfd);
    }
    std::vector<Level> levels;
    std::vector<Buf<i32>> parts;
    std::vector<Buf<i32>> cons_store;
    levels.push_back(top_level(top));
    if (top_plan) levels[0].plan.reset();
    i32* cons = constraint;
    bool drop = constrained;
    while (true) {
        Level& L = levels.back();
        LevelPlan& lp = (levels.size() == 1 && top_plan) ? *top_plan : *L.plan;
        const DGraph g = L.g;
        Buf<i32> part(c, std::max<i64>(g.n, 1));
        iota_i32(c, part, g.n);
        c.mem_note("louvain: level start");
        move_level(c, ms, lp, g, part, rfd, true, constrained, cons, true, rs, levels.size() == 1 ? rn : nullptr);
        c.mem_note("louvain: after local move");
        if (constrained && rfd.label_count == g.n) local_move_strict(c, ms, lp, g, part, rfd, true, constrained, cons, &rs.mv);
        const i32* pp = part.p;
        parts.push_back(std::move(part));
        if (rfd.label_count < g.n) {
            const i64 nc = rfd.label_count;
            std::unique_ptr<OwnedGraph> og(new OwnedGraph());
            auto t0 = std::chrono::steady_clock::now();
            contract_level(c, g, pp, nc, *og, levels.size() == 1 ? rn : nullptr);
            c.sync();
            rs.contract_s += since(t0);
            og->vw = Buf<i64>(c, std::max<i64>(nc, 1));
            copy_dev(c, og->vw, rfd.vol, nc * sizeof(i64));
            if (constrained) {
                Buf<i32> nxt(c, std::max<i64>(nc, 1));
                downsample_max(c, g.n, pp, nc, cons, nxt);
                cons = nxt.p;
                cons_store.push_back(std::move(nxt));
            }
            levels.push_back(owned_level(std::move(og)));
        } else if (drop) {
            fill_i32(c, cons, g.n, 0);
            parts.pop_back();
            drop = false;
        } else {
            break;
        }
    }
    rs.levels = (i64)levels.size();  // setLevelCount / setTotalNnz of the last call
    rs.total_nnz = 0;
    for (auto& l : levels) rs.total_nnz += l.g.nnz;
    if (levels.size() > 1) {
        levels.pop_back();
        parts.pop_back();
    }
    for (int i = (int)levels.size() - 2; i >= 0; i--) {
        const DGraph g = levels[i].g;
        LevelPlan& lp = (i == 0 && top_plan) ? *top_plan : *levels[i].plan;
        project(c, g.n, parts[i + 1], parts[i]);
        release_level(levels[i + 1], parts[i + 1]);  // done with the coarser level
        move_level(c, ms, lp, g, parts[i], rfd, false, false, nullptr, true, rs, i == 0 ? rn : nullptr);
    }
    copy_dev(c, part_out, parts[0].p, top.n * sizeof(i32));
    c.sync();
}

// leiden_part<plus, improve> (src/clustering_methods.cpp:47-137)
static void leiden_part(Ctx& c, MoveScratch& ms, const DGraph& top, RefineData& rfd, bool plus, bool improve,
                        const i32* input, i32* part_out, RunStats& rs, LevelPlan* top_plan, Renum* rn) {
    std::vector<Level> levels;
    std::vector<Buf<i32>> parts;
    levels.push_back(top_level(top));
    Buf<i32> part(c, std::max<i64>(top.n, 1));
    if (!improve) iota_i32(c, part, top.n);
    else copy_dev(c, part, input, top.n * sizeof(i32));
    int stalled = 0;  // consecutive refinements that contracted nothing
    while (true) {
        const size_t li = levels.size() - 1;
        const DGraph g = levels[li].g;
        LevelPlan& lp = (li == 0 && top_plan) ? *top_plan : *levels[li].plan;
        const double old_obj = rfd.obj;
        const bool init = !improve && levels.size() == 1;
        Renum* r0 = li == 0 ? rn : nullptr;
        c.mem_note("leiden: level start");
        move_level(c, ms, lp, g, part, rfd, init, false, nullptr, true, rs, r0);
        c.mem_note("leiden: after local move");
        if (g.nnz > 100000 && old_obj == rfd.obj) move_level(c, ms, lp, g, part, rfd, init, false, nullptr, false, rs, r0);
        if (old_obj == rfd.obj) local_move_strict(c, ms, lp, g, part, rfd, init, false, nullptr, &rs.mv);
        if (rfd.label_count == g.n) {
            parts.push_back(std::move(part));
            break;
        }
        Buf<i32> louv = std::move(part);
        Buf<i64> pv(c, std::max<i64>(g.n, 1));
        own_weight(c, g, louv, pv);
        const bool uni = levels.size() == 1 && g.wk == W_UNIT;
        Buf<i32> cmap(c, std::max<i64>(g.n, 1));
        auto t0 = std::chrono::steady_clock::now();
        const Orders* ord = &lp.ord;
        if (r0) {  // the local mover ran on the renumbered copy's plan
            if (!r0->ord_built) {
                make_orders(c, g, r0->ord);
                r0->ord_built = true;
            }
            ord = &r0->ord;
        }
        const i64 nc = leiden_refine(c, g, *ord, louv, rfd.vol, rfd.lambda, pv, uni, cmap);
        c.mem_note("leiden: after refinement");
        c.sync();
        rs.refine_s += since(t0);
        const i32* cm = cmap.p;
        parts.push_back(std::move(cmap));
        if (nc < g.n) {
            std::unique_ptr<OwnedGraph> og(new OwnedGraph());
            auto t1 = std::chrono::steady_clock::now();
            contract_level(c, g, cm, nc, *og, r0);
            c.sync();
            rs.contract_s += since(t1);
            og->vw = Buf<i64>(c, std::max<i64>(nc, 1));
            coarsen_vertex_weights(c, g, cm, nc, og->vw);
            part = Buf<i32>(c, std::max<i64>(nc, 1));
            k_downsample<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.n, cm, louv, part);
            PGGC_LAUNCH_CHECK();
            levels.push_back(owned_level(std::move(og)));
            stalled = 0;
        } else {
            part = std::move(louv);
            // The reference retries the same level with fresh seeds until the
            // refinement contracts (:105-110), which never happens when no
            // neighbour passes the refinement tests; stop after LEIDEN_MAX_STALLS
            // tries at the current clustering (the map just pushed is the identity).
            if (++stalled >= LEIDEN_MAX_STALLS) {
                parts.pop_back();
                parts.push_back(std::move(part));
                break;
            }
            Level same;
            same.g = g;
            same.plan.reset(new LevelPlan());
            levels.push_back(std::move(same));
        }
    }
    rs.levels = (i64)levels.size();  // setLevelCount / setTotalNnz of the last call
    rs.total_nnz = 0;
    for (auto& l : levels) rs.total_nnz += l.g.nnz;
    for (int i = (int)levels.size() - 2; i >= 0; i--) {
        const DGraph g = levels[i].g;
        LevelPlan& lp = (i == 0 && top_plan) ? *top_plan : *levels[i].plan;
        project(c, g.n, parts[i + 1], parts[i]);
        release_level(levels[i + 1], parts[i + 1]);  // done with the coarser level
        if (plus) move_level(c, ms, lp, g, parts[i], rfd, false, false, nullptr, true, rs, i == 0 ? rn : nullptr);
    }
    copy_dev(c, part_out, parts[0].p, top.n * sizeof(i32));
    c.sync();
}

// get_objective (app/objective_helpers.hpp:9-22, interface/cluster_data.hpp)
static void state_init(Ctx& c, const DGraph& g, int objective, double mult, RefineData& rd) {
    copy_dev(c, rd.vol, g.vw, g.n * sizeof(i64));
    rd.top_nnz = g.nnz;
    rd.label_count = (i32)g.n;
    rd.uncut = 0;
    rd.lambda = 1.0;
    rd.scale = 1.0;
    rd.offset = 0.0;
    Buf<i64> tmp(c, 2);
    if (objective == PGGC_OBJ_MOD || objective == PGGC_OBJ_WMOD) {
        i64 gd;
        if (objective == PGGC_OBJ_MOD || g.wk == W_UNIT) gd = g.nnz;
        else {
            sum_i64(c, g.vw, g.n, tmp);
            gd = c.read1(tmp);
        }
        double inv = 1.0 / (double)gd;
        rd.lambda = mult * inv;
        rd.scale = inv;
    } else if (objective == PGGC_OBJ_CPM) {
        rd.lambda = mult * 0.5;
        rd.offset = rd.lambda * (double)g.n;
    } else if (objective == PGGC_OBJ_NLCC) {
        i64 gd = g.nnz;
        if (g.wk != W_UNIT) {
            Buf<i64> wd(c, std::max<i64>(g.n, 1));
            PGGC_WDISPATCH(g, (k_wdeg<<<grid_items(g.n, 4), 256, 0, c.stream>>>(g.rp, W, g.n, wd)));
            PGGC_LAUNCH_CHECK();
            sum_i64(c, wd, g.n, tmp);
            gd = c.read1(tmp);
        }
        sum_i64(c, g.vw, g.n, tmp);
        u64 vt = (u64)c.read1(tmp);
        double denom = (double)(vt * vt);
        rd.lambda = mult * (double)gd / denom;
        rd.scale = 1.0 / (double)gd;
    } else if (objective == PGGC_OBJ_LCC) {
        rd.lambda = mult;
    } else {
        throw Error(-1, "unknown objective");
    }
    update_objective(c, rd);
}

}  // namespace pggc

// =================================================================== C ABI
static thread_local std::string g_err;

// shared with the host-only file readers (io.cpp)
void set_last_error(const std::string& msg) { g_err = msg; }

template <class F>
static int guard(F f) {
    try {
        f();
        return PGGC_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PGGC_EINTERNAL;
    }
}

static DGraph view_of(const pggc_graph* g) {
    if (!g || g->n < 0 || g->nnz < 0 || (g->n > 0 && !g->row_ptr) || (g->nnz > 0 && !g->col))
        throw Error(PGGC_EINVAL, "invalid graph");
    if (g->w && g->w64) throw Error(PGGC_EINVAL, "graph has both int32 and int64 weights");
    if (g->n >= 0x7fffffff) throw Error(PGGC_EINVAL, "vertex count exceeds int32 ids");
    DGraph d;
    d.n = g->n;
    d.nnz = g->nnz;
    d.rp = g->row_ptr;
    d.col = g->col;
    d.vw = g->vtx_w;
    if (g->w) {
        d.w = g->w;
        d.wk = W_I32;
    } else if (g->w64) {
        d.w = g->w64;
        d.wk = W_I64;
    }
    return d;
}

static RefineData rd_of(const pggc_state* s) {
    if (!s || !s->vol) throw Error(PGGC_EINVAL, "invalid state");
    RefineData r;
    r.vol = s->vol;
    r.cap = s->vol_cap;
    r.uncut = s->uncut;
    r.top_nnz = s->top_nnz;
    r.obj = s->obj;
    r.label_count = s->label_count;
    r.lambda = s->lambda;
    r.scale = s->scale;
    r.offset = s->offset;
    return r;
}
static void rd_store(const RefineData& r, pggc_state* s) {
    s->uncut = r.uncut;
    s->top_nnz = r.top_nnz;
    s->obj = r.obj;
    s->label_count = r.label_count;
    s->lambda = r.lambda;
    s->scale = r.scale;
    s->offset = r.offset;
}

#define CTX_CHECK(ctx) \
    if (!(ctx)) throw Error(PGGC_EINVAL, "null context")

extern "C" {

int pggc_version(void) { return 100; }
const char* pggc_last_error(void) { return g_err.c_str(); }

int pggc_ctx_create(pggc_ctx** out, void* stream) {
    return guard([&] {
        if (!out) throw Error(PGGC_EINVAL, "null output");
        std::unique_ptr<pggc_ctx> x(new pggc_ctx());
        x->c.stream = (hipStream_t)stream;
        PGGC_HIP(hipGetDevice(&x->c.device));
        // coherent (fine-grained) host memory: the readback kernel's stores
        // reach the host without a cache writeback the host would have to wait for
        PGGC_HIP(hipHostMalloc((void**)&x->c.pinned, 64 * sizeof(i64), hipHostMallocCoherent));
        PGGC_HIP(hipHostGetDevicePointer((void**)&x->c.pinned_dev, x->c.pinned, 0));
        noalloc_init(x->c.device);  // resolves the evaluator scratch of this device (once per device)
        x->ms.gnn = &x->gnn;
        *out = x.release();
    });
}

int pggc_ctx_destroy(pggc_ctx* ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->c.stream);
        for (pggc_ctx* w : ctx->workers) (void)pggc_ctx_destroy(w);
        ctx->workers.clear();
        ctx->gnn = Gnn();
        ctx->ms = MoveScratch();
        for (auto& kv : ctx->c.live) (void)hipFree(kv.first);
        ctx->c.live.clear();
        ctx->c.trim();
        if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
        if (ctx->c.stage) (void)hipHostFree(ctx->c.stage);
        ctx->c.kt_free();
        delete ctx;
    });
}

int pggc_ctx_set_seed(pggc_ctx* ctx, uint64_t seed, uint64_t draws) {
    return guard([&] {
        CTX_CHECK(ctx);
        ctx->c.seed = seed;
        ctx->c.draws = draws;
    });
}
int pggc_ctx_set_verbose(pggc_ctx* ctx, int level) {
    return guard([&] {
        CTX_CHECK(ctx);
        ctx->c.verbose = level;
    });
}
int pggc_ctx_set_progress(pggc_ctx* ctx, pggc_progress_fn fn, void* user) {
    return guard([&] {
        CTX_CHECK(ctx);
        ctx->c.progress = fn;
        ctx->c.progress_user = user;
    });
}

// ------------------------------------------------------ active-set predictor
// Deterministic initial parameters: uniform(-a, a) with a = sqrt(6 / (fan in
// + fan out)) from splitmix64, zero hidden biases, output bias at the log-odds
// of a 10% move rate.
static void gnn_initial_params(float* p) {
    u64 st = 0x5eed5eedULL;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ULL;
        u64 z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const double a1 = std::sqrt(6.0 / (GNN_IN + GNN_HID)), a2 = std::sqrt(6.0 / (GNN_HID + 1));
    for (int k = 0; k < GNN_IN * GNN_HID; k++) p[GP_W1 + k] = (float)(a1 * uni());
    for (int k = 0; k < GNN_HID; k++) p[GP_B1 + k] = 0.0f;
    for (int k = 0; k < GNN_HID; k++) p[GP_W2S + k] = (float)(a2 * uni());
    for (int k = 0; k < GNN_HID; k++) p[GP_W2N + k] = (float)(a2 * uni());
    p[GP_B2] = (float)std::log(0.1 / 0.9);
}

int pggc_gnn_configure(pggc_ctx* ctx, int mode, double tau, int64_t min_nnz, int full_every, double lr,
                       double pos_weight) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (mode < PGGC_GNN_OFF || mode > PGGC_GNN_PRUNE)
            throw Error(PGGC_EINVAL, "predictor mode must be PGGC_GNN_OFF, PGGC_GNN_TRAIN or PGGC_GNN_PRUNE");
        if (min_nnz < 0 || full_every < 0 || !(lr > 0.0) || !(pos_weight > 0.0) || !std::isfinite(tau))
            throw Error(PGGC_EINVAL, "invalid predictor configuration");
        Gnn& G = ctx->gnn;
        G.init(ctx->c);
        if (mode == PGGC_GNN_PRUNE && !G.have_params)
            throw Error(PGGC_EINVAL, "pruning needs trained predictor parameters (pggc_gnn_set_params)");
        if (mode == PGGC_GNN_TRAIN && !G.have_params) {
            std::vector<float> p(GNN_P);
            gnn_initial_params(p.data());
            G.set_params(ctx->c, p.data());
        }
        G.cfg.mode = mode;
        G.cfg.tau = (float)tau;
        G.cfg.min_nnz = min_nnz;
        G.cfg.full_every = full_every;
        G.cfg.lr = (float)lr;
        G.cfg.pos_weight = (float)pos_weight;
    });
}

int pggc_gnn_set_params(pggc_ctx* ctx, const float* params, int count) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!params || count != GNN_P) throw Error(PGGC_EINVAL, "expected PGGC_GNN_PARAMS predictor parameters");
        for (int k = 0; k < count; k++)
            if (!std::isfinite(params[k])) throw Error(PGGC_EINVAL, "predictor parameters must be finite");
        ctx->gnn.set_params(ctx->c, params);
    });
}

int pggc_gnn_get_params(pggc_ctx* ctx, float* params, int count) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!params || count != GNN_P) throw Error(PGGC_EINVAL, "expected PGGC_GNN_PARAMS predictor parameters");
        if (!ctx->gnn.have_params) throw Error(PGGC_EINVAL, "no predictor parameters set");
        PGGC_HIP(hipMemcpyAsync(params, ctx->gnn.prm.p, GNN_P * sizeof(float), hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}

int pggc_gnn_stats(pggc_ctx* ctx, int64_t* counts, double* loss_sum) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!counts) throw Error(PGGC_EINVAL, "null output");
        Gnn& G = ctx->gnn;
        G.init(ctx->c);
        i64 d[GC_COUNT];
        double l = 0.0;
        PGGC_HIP(hipMemcpyAsync(d, G.dcnt.p, sizeof(d), hipMemcpyDeviceToHost, ctx->c.stream));
        PGGC_HIP(hipMemcpyAsync(&l, G.dloss.p, sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
        const int64_t out[9] = {G.sweeps, d[GC_ACTIVE], d[GC_PRUNED], G.fallbacks, G.train_steps,
                                d[GC_TP],  d[GC_FP],     d[GC_FN],     d[GC_TN]};
        std::memcpy(counts, out, sizeof(out));
        if (loss_sum) *loss_sum = l;
    });
}

int pggc_gnn_reset_stats(pggc_ctx* ctx) {
    return guard([&] {
        CTX_CHECK(ctx);
        ctx->gnn.init(ctx->c);
        ctx->gnn.reset_counters(ctx->c);
        ctx->c.sync();
    });
}

int pggc_gnn_eval(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* part, const int32_t* dest, const float* gain,
                  const int32_t* stamp, int32_t cur, int first, const int64_t* vol, const int32_t* labels,
                  uint16_t* rows_out, float* logit_out, double* loss_out, float* grad_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        if (!g.vw) throw Error(PGGC_EINVAL, "vertex weights required");
        if (g.n > 0 && (!part || !dest || !gain || !stamp || !vol))
            throw Error(PGGC_EINVAL, "part, dest, gain, stamp and vol are required");
        if ((loss_out || grad_out) && !labels) throw Error(PGGC_EINVAL, "loss and gradient need labels");
        Ctx& c = ctx->c;
        Gnn& G = ctx->gnn;
        if (!G.have_params) throw Error(PGGC_EINVAL, "no predictor parameters set");
        const i64 n = g.n;
        LevelPlan lp;
        plan_level(c, g, lp);
        G.ensure(c, n, true);
        Buf<i32> d2(c, std::max<i64>(n, 1));
        if (n > 0) {
            copy_dev(c, d2.p, dest, n * sizeof(i32));
            copy_dev(c, G.stamp.p, stamp, n * sizeof(i32));
            fill_i32(c, (i32*)G.logit.p, n, 0);
        }
        const i64 sw = G.sweeps;
        gnn_predict(c, G, lp, g, part, d2, gain, vol, cur, first != 0, false, true);
        G.sweeps = sw;
        if (rows_out && n > 0)
            copy_dev(c, rows_out, G.in.p, n * GNN_IN * sizeof(u16));
        if (logit_out && n > 0)
            copy_dev(c, logit_out, G.logit.p, n * sizeof(float));
        if (labels) {
            double l0 = 0.0, l1 = 0.0;
            PGGC_HIP(hipMemcpyAsync(&l0, G.dloss.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            c.sync();
            gnn_backward(c, G, lp, g, labels, cur, false);
            PGGC_HIP(hipMemcpyAsync(&l1, G.dloss.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            if (grad_out) PGGC_HIP(hipMemcpyAsync(grad_out, G.grad.p, GNN_P * sizeof(float), hipMemcpyDeviceToHost, c.stream));
            c.sync();
            if (loss_out) *loss_out = l1 - l0;
        }
        c.sync();
    });
}

// candidate values of the execution-shape knobs (what the kernels are
// instantiated for); the other knobs take any value >= 0
static const std::vector<int64_t>* knob_candidates(int knob) {
    static const std::vector<int64_t> threads{64, 128, 256}, spe{2, 4}, lds{2048, 4096}, edges{512, 1024, 2048},
        slots{2048, 4096, 8192}, div{16, 32, 64};
    switch (knob) {
        case PGGC_KNOB_TABLE_THREADS: return &threads;
        case PGGC_KNOB_TABLE_SLOTS_PER_EDGE: return &spe;
        case PGGC_KNOB_HUB_LDS_MAX_DEGREE: return &lds;
        case PGGC_KNOB_HUB_EDGES_PER_ITEM: return &edges;
        case PGGC_KNOB_HUB_SLOTS_PER_ITEM: return &slots;
        case PGGC_KNOB_HUB_UPDATE_DIV: return &div;  // the autotuner's range (any value >= 0 is accepted)
        default: return nullptr;
    }
}

static void set_knob(Ctx& c, int knob, int64_t value) {
    if (value < 0) throw Error(PGGC_EINVAL, "knob value must be >= 0");
    const std::vector<int64_t>* cand = knob_candidates(knob);
    const bool shape = knob == PGGC_KNOB_TABLE_THREADS || knob == PGGC_KNOB_TABLE_SLOTS_PER_EDGE ||
                       knob == PGGC_KNOB_HUB_LDS_MAX_DEGREE || knob == PGGC_KNOB_HUB_EDGES_PER_ITEM ||
                       knob == PGGC_KNOB_HUB_SLOTS_PER_ITEM;
    if (shape && std::find(cand->begin(), cand->end(), value) == cand->end())
        throw Error(PGGC_EINVAL, "knob value is not one the kernels are built for");
    switch (knob) {
        case PGGC_KNOB_HUB_UPDATE_DIV: c.hub_update_div = value; break;
        case PGGC_KNOB_GNN_MAX_SWEEPS: c.gnn_max_sweeps = value; break;
        case PGGC_KNOB_TABLE_THREADS: c.table_threads = (int)value; break;
        case PGGC_KNOB_TABLE_SLOTS_PER_EDGE: c.table_spe = (int)value; break;
        case PGGC_KNOB_HUB_LDS_MAX_DEGREE: c.hub_lds_max = value; break;
        case PGGC_KNOB_HUB_EDGES_PER_ITEM: c.hub_edges_item = value; break;
        case PGGC_KNOB_HUB_SLOTS_PER_ITEM: c.hub_slots_item = value; break;
        case PGGC_KNOB_RAW_SWEEPS: c.raw_sweeps = value ? 1 : 0; break;
        case PGGC_KNOB_HUB_BUILD: c.hub_build = (int)std::min<int64_t>(value, 2); break;
        case PGGC_KNOB_RENUMBER: c.renumber = (int)std::min<int64_t>(value, 2); break;
        case PGGC_KNOB_CONTRACT_BATCH: c.contract_batch = std::max<int64_t>(value, 1); break;
        default: throw Error(PGGC_EINVAL, "unknown knob");
    }
    if (shape) c.knob_epoch++;  // level plans depend on these
}

static int64_t get_knob(const Ctx& c, int knob) {
    switch (knob) {
        case PGGC_KNOB_HUB_UPDATE_DIV: return c.hub_update_div;
        case PGGC_KNOB_GNN_MAX_SWEEPS: return c.gnn_max_sweeps;
        case PGGC_KNOB_TABLE_THREADS: return c.table_threads;
        case PGGC_KNOB_TABLE_SLOTS_PER_EDGE: return c.table_spe;
        case PGGC_KNOB_HUB_LDS_MAX_DEGREE: return c.hub_lds_max;
        case PGGC_KNOB_HUB_EDGES_PER_ITEM: return c.hub_edges_item;
        case PGGC_KNOB_HUB_SLOTS_PER_ITEM: return c.hub_slots_item;
        case PGGC_KNOB_RAW_SWEEPS: return c.raw_sweeps;
        case PGGC_KNOB_HUB_BUILD: return c.hub_build;
        case PGGC_KNOB_RENUMBER: return c.renumber;
        case PGGC_KNOB_CONTRACT_BATCH: return c.contract_batch;
        default: throw Error(PGGC_EINVAL, "unknown knob");
    }
}

static void copy_knobs(Ctx& dst, const Ctx& src) {
    for (int k = 0; k < PGGC_KNOB_COUNT; k++) {
        const int64_t v = get_knob(src, k);
        if (get_knob(dst, k) != v) set_knob(dst, k, v);
    }
}

int pggc_ctx_set_knob(pggc_ctx* ctx, int knob, int64_t value) {
    return guard([&] {
        CTX_CHECK(ctx);
        set_knob(ctx->c, knob, value);
    });
}

int pggc_ctx_get_knob(pggc_ctx* ctx, int knob, int64_t* value) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!value) throw Error(PGGC_EINVAL, "null output");
        *value = get_knob(ctx->c, knob);
    });
}

// Coordinate descent over the execution-shape knobs: each knob in turn
// takes every candidate value, one pLouvain clustering of g (louvain_part:
// every level down and up, the unit of the ex_iters loop) is timed, and the
// fastest value is kept.  A first untimed run per value builds the plans.
int pggc_autotune(pggc_ctx* ctx, const pggc_graph* gg, int objective, int reps, int64_t* report, int cap,
                  int* nrows) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (reps < 1) throw Error(PGGC_EINVAL, "reps must be >= 1");
        if (ctx->gnn.cfg.mode != GNN_OFF) throw Error(PGGC_EINVAL, "turn the active-set predictor off to tune");
        DGraph g = view_of(gg);
        if (!g.vw) throw Error(PGGC_EINVAL, "vertex weights required");
        Ctx& c = ctx->c;
        const i64 n = g.n;
        Buf<i64> vol(c, std::max<i64>(n, 1));
        Buf<i32> part(c, std::max<i64>(n, 1));
        int rows = 0;
        const int order[] = {PGGC_KNOB_HUB_LDS_MAX_DEGREE, PGGC_KNOB_TABLE_THREADS, PGGC_KNOB_TABLE_SLOTS_PER_EDGE,
                             PGGC_KNOB_HUB_EDGES_PER_ITEM, PGGC_KNOB_HUB_SLOTS_PER_ITEM, PGGC_KNOB_HUB_UPDATE_DIV};
        for (int knob : order) {
            const std::vector<int64_t>& cand = *knob_candidates(knob);
            int64_t best_v = get_knob(c, knob);
            double best_t = std::numeric_limits<double>::infinity();
            for (int64_t v : cand) {
                set_knob(c, knob, v);
                LevelPlan top;
                Renum rtop;
                double t_min = std::numeric_limits<double>::infinity();
                for (int r = 0; r <= reps; r++) {
                    RefineData rd;
                    rd.vol = vol;
                    rd.cap = std::max<i64>(n, 1);
                    state_init(c, g, objective, 1.0, rd);
                    RunStats rs;
                    c.sync();
                    const auto t0 = std::chrono::steady_clock::now();
                    louvain_part(c, ctx->ms, g, rd, false, nullptr, part, rs, &top, top_renum(c, ctx->ms, g, rtop, true));
                    c.sync();
                    if (r > 0) t_min = std::min(t_min, since(t0));
                }
                if (report && rows < cap) {
                    report[3 * rows] = knob;
                    report[3 * rows + 1] = v;
                    report[3 * rows + 2] = (int64_t)(t_min * 1e6);
                }
                rows++;
                if (t_min < best_t) {
                    best_t = t_min;
                    best_v = v;
                }
            }
            set_knob(c, knob, best_v);
        }
        if (nrows) *nrows = rows;
    });
}

int pggc_ctx_set_sweep_trace(pggc_ctx* ctx, pggc_sweep_fn fn, void* user) {
    return guard([&] {
        CTX_CHECK(ctx);
        ctx->c.sweep_trace = fn;
        ctx->c.sweep_user = user;
    });
}

int pggc_ctx_last_times(pggc_ctx* ctx, double* contract_s, double* refine_s) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (contract_s) *contract_s = ctx->c.last_contract_s;
        if (refine_s) *refine_s = ctx->c.last_refine_s;
    });
}

int pggc_ctx_set_timing(pggc_ctx* ctx, int on) {
    return guard([&] {
        CTX_CHECK(ctx);
        ctx->c.kt_resolve();
        KTimer& k = ctx->c.kt;
        k.on = on != 0;
        for (int i = 0; i < KS_COUNT; i++) {
            k.launches[i] = 0;
            k.ms[i] = 0.0;
            k.bytes[i] = 0.0;
        }
    });
}

int pggc_ctx_kernel_stats(pggc_ctx* ctx, int slot, int64_t* launches, double* ms, double* bytes) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (slot < 0 || slot >= KS_COUNT) throw Error(PGGC_EINVAL, "unknown timing slot");
        ctx->c.kt_resolve();
        const KTimer& k = ctx->c.kt;
        if (launches) *launches = k.launches[slot];
        if (ms) *ms = k.ms[slot];
        if (bytes) *bytes = k.bytes[slot];
    });
}

int pggc_ctx_get_draws(pggc_ctx* ctx, uint64_t* draws) {
    return guard([&] {
        CTX_CHECK(ctx);
        *draws = ctx->c.draws;
    });
}
int pggc_ctx_memory(pggc_ctx* ctx, int64_t* live, int64_t* cached) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (live) *live = (int64_t)ctx->c.bytes_live;
        if (cached) *cached = (int64_t)ctx->c.bytes_cached;
    });
}

int pggc_vertex_weights(pggc_ctx* ctx, const pggc_graph* gg, int objective, int64_t* out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        Ctx& c = ctx->c;
        if (g.n == 0) return;
        if (objective == PGGC_OBJ_MOD || (objective == PGGC_OBJ_WMOD && g.wk == W_UNIT)) {
            k_deg<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.rp, g.n, out);
        } else if (objective == PGGC_OBJ_WMOD) {
            PGGC_WDISPATCH(g, (k_wdeg<<<grid_items(g.n, 4), 256, 0, c.stream>>>(g.rp, W, g.n, out)));
        } else {
            fill_i64(c, out, g.n, 1);
        }
        PGGC_LAUNCH_CHECK();
        c.sync();
    });
}

int pggc_state_init(pggc_ctx* ctx, const pggc_graph* gg, int objective, double mult, pggc_state* st) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        RefineData rd = rd_of(st);
        if (rd.cap < g.n) throw Error(PGGC_ECAPACITY, "state volume capacity below n");
        state_init(ctx->c, g, objective, mult, rd);
        rd_store(rd, st);
    });
}

int pggc_orderings(pggc_ctx* ctx, const pggc_graph* gg, int32_t* order1, int32_t* order2, int64_t* offs) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        Ctx& c = ctx->c;
        Orders o;
        make_orders(c, g, o);
        if (g.n > 0) {
            copy_dev(c, order1, o.o1, g.n * sizeof(i32));
            copy_dev(c, order2, o.o2, g.n * sizeof(i32));
        }
        c.sync();
        offs[0] = o.off_large;
        offs[1] = o.off_large2;
        offs[2] = o.off_mid;
        offs[3] = o.off_mid2;
    });
}

int pggc_evaluate(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* labels, int64_t L, int64_t* vol_out,
                  int64_t* uncut_out, uint64_t* sq_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        Ctx& c = ctx->c;
        if (!g.vw) throw Error(PGGC_EINVAL, "vertex weights required");
        Buf<i64> scal(c, 4);
        evaluate(c, g, labels, L, vol_out, scal);
        sum_sq_u128(c, vol_out, L, scal.p + 1);
        const i64* h = c.read(scal, 3);
        *uncut_out = h[0];
        sq_out[0] = (u64)h[1];
        sq_out[1] = (u64)h[2];
    });
}

int pggc_evaluate_noalloc(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* labels, int64_t L, int64_t* vol_out,
                          int64_t* uncut_out, uint64_t* sq_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!uncut_out || !sq_out) throw Error(PGGC_EINVAL, "uncut_out and sq_out are required");
        DGraph g = view_of(gg);
        evaluate_noalloc(ctx->c, g, labels, L, vol_out, uncut_out, sq_out);
    });
}

int pggc_evaluate_noalloc_limits(int64_t* max_nnz, int64_t* scratch_bytes) {
    return guard([&] {
        if (max_nnz) *max_nnz = noalloc_max_nnz();
        if (scratch_bytes) *scratch_bytes = noalloc_scratch_bytes();
    });
}

int pggc_own_weight(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* labels, int64_t* out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        own_weight(ctx->c, g, labels, out);
        ctx->c.sync();
    });
}

int pggc_contract(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* vcmap, int64_t nc, int64_t* out_rp,
                  int32_t* out_col, int64_t* out_w, int64_t cap, int64_t* nnz_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        if (nc < 0) throw Error(PGGC_EINVAL, "negative cluster count");
        Ctx& c = ctx->c;
        OwnedGraph og;
        contract(c, g, vcmap, nc, og);
        *nnz_out = og.nnz;
        copy_dev(c, out_rp, og.rp, (nc + 1) * sizeof(i64));
        if (og.nnz > cap) {
            c.sync();
            throw Error(PGGC_ECAPACITY, "coarse nnz exceeds output capacity");
        }
        if (og.nnz > 0) {
            copy_dev(c, out_col, og.col, og.nnz * sizeof(i32));
            copy_dev(c, out_w, og.w, og.nnz * sizeof(i64));
        }
        c.sync();
    });
}

int pggc_sweep(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* part, const int64_t* vol, int64_t L,
               double lambda, float ratio, int raw, const int32_t* constraint, int32_t* dest, float* gain,
               int32_t* moves_out, int64_t* nmoves, int64_t* ncand) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        LevelPlan lp;
        *nmoves = sweep_once(ctx->c, ctx->ms, lp, g, part, vol, L, lambda, ratio, raw != 0, constraint, dest, gain,
                             moves_out, ncand);
    });
}

int pggc_local_move(pggc_ctx* ctx, const pggc_graph* gg, int32_t* part, pggc_state* st, int is_initial, int strict,
                    int SA) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        RefineData rd = rd_of(st);
        LevelPlan lp;
        if (strict) local_move_strict(ctx->c, ctx->ms, lp, g, part, rd, is_initial != 0, false, nullptr, nullptr);
        else local_move(ctx->c, ctx->ms, lp, g, part, rd, is_initial != 0, false, nullptr, SA != 0, nullptr);
        ctx->c.sync();
        rd_store(rd, st);
    });
}

int pggc_leiden_refine(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* louv, const int64_t* vol, int64_t L,
                       double lambda, int uniform, uint64_t seed, uint64_t draw0, int32_t* vcmap_out,
                       int64_t* nc_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        Ctx& c = ctx->c;
        (void)L;
        u64 s0 = c.seed, d0 = c.draws;
        c.seed = seed;
        c.draws = draw0;
        Orders o;
        make_orders(c, g, o);
        Buf<i64> pv(c, std::max<i64>(g.n, 1));
        own_weight(c, g, louv, pv);
        *nc_out = leiden_refine(c, g, o, louv, vol, lambda, pv, uniform != 0, vcmap_out);
        c.sync();
        c.seed = s0;
        c.draws = d0;
    });
}

int pggc_louvain(pggc_ctx* ctx, const pggc_graph* gg, pggc_state* st, int constrained, int32_t* constraint,
                 int32_t* part_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        RefineData rd = rd_of(st);
        RunStats rs;
        Renum rtop;
        louvain_part(ctx->c, ctx->ms, g, rd, constrained != 0, constraint, part_out, rs, nullptr,
                     top_renum(ctx->c, ctx->ms, g, rtop, false));
        rd_store(rd, st);
    });
}

int pggc_leiden(pggc_ctx* ctx, const pggc_graph* gg, pggc_state* st, int plus, int improve, const int32_t* input,
                int32_t* part_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        RefineData rd = rd_of(st);
        RunStats rs;
        Renum rtop;
        leiden_part(ctx->c, ctx->ms, g, rd, plus != 0, improve != 0, input, part_out, rs, nullptr,
                    top_renum(ctx->c, ctx->ms, g, rtop, false));
        rd_store(rd, st);
    });
}

int pggc_cluster(pggc_ctx* ctx, const pggc_graph* gg, int algo, int objective, double mult, int ex_iters,
                 uint64_t seed, int32_t* part_out, double* objective_out, int64_t* info) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        if (!g.vw) throw Error(PGGC_EINVAL, "vertex weights required");
        // Mod and CPM treat edge weights as unit (app/objective_helpers.hpp:24-29)
        if (objective == PGGC_OBJ_MOD || objective == PGGC_OBJ_CPM) {
            g.w = nullptr;
            g.wk = W_UNIT;
        }
        Ctx& c = ctx->c;
        c.seed = seed;
        c.draws = 0;
        Buf<i64> vol(c, g.n + 1);
        RefineData rd;
        rd.vol = vol;
        rd.cap = g.n;
        state_init(c, g, objective, mult, rd);
        RunStats rs;
        LevelPlan top;
        Renum rtop;
        Renum* rn = top_renum(c, ctx->ms, g, rtop, ex_iters != 0);
        // per-iteration line as printed by the reference (app/cluster_main.cpp:137,156,
        // interface/cluster_data.hpp:98-103)
        auto t_iter = std::chrono::steady_clock::now();
        i64 iter_no = 0;
        auto report = [&]() {
            c.sync();
            auto now = std::chrono::steady_clock::now();
            double secs = std::chrono::duration<double>(now - t_iter).count();
            t_iter = now;
            const i64 cut = (rd.top_nnz - rd.uncut) / 2;
            if (c.progress) c.progress(c.progress_user, iter_no, secs, cut, rd.objective(), rd.label_count);
            iter_no++;
            if (!c.verbose) return;
            fprintf(stderr, "Cluster time: %g; Cut: %lld; Objective: %.9g; Labels: %d; sweeps so far: %lld\n", secs,
                    (long long)cut, rd.objective(), rd.label_count, (long long)rs.mv.sweeps);
            fflush(stderr);
        };
        if (algo == PGGC_LOUVAIN) louvain_part(c, ctx->ms, g, rd, false, nullptr, part_out, rs, &top, rn);
        else leiden_part(c, ctx->ms, g, rd, algo == PGGC_LEIDEN_PLUS, false, nullptr, part_out, rs, &top, rn);
        report();
        double last = -std::numeric_limits<double>::infinity();
        int counter = 0;
        i64 outer = 1;
        Buf<i32> cons(c, g.n + 1);
        while (true) {  // iteration_termination_condition (app/cluster_main.cpp:101-114)
            bool term = true;
            if (ex_iters < 0) {
                if (rd.objective() > last) term = false;
                last = rd.objective();
            } else if (counter++ < ex_iters) {
                term = false;
            }
            if (term) break;
            copy_dev(c, cons, part_out, g.n * sizeof(i32));
            if (algo == PGGC_LOUVAIN) louvain_part(c, ctx->ms, g, rd, true, cons, part_out, rs, &top, rn);
            else leiden_part(c, ctx->ms, g, rd, algo == PGGC_LEIDEN_PLUS, true, cons, part_out, rs, &top, rn);
            outer++;
            report();
        }
        c.sync();
        c.last_contract_s = rs.contract_s;
        c.last_refine_s = rs.refine_s;
        *objective_out = rd.objective();
        if (info) {
            info[0] = rd.label_count;
            info[1] = rd.uncut;
            info[2] = rs.levels;
            info[3] = rs.mv.sweeps;
            info[4] = rs.mv.moves;
            info[5] = outer;
            info[6] = rs.mv.cands;
            info[7] = rs.total_nnz;
        }
    });
}

// Many independent clusterings in one call: worker contexts, each with a
// host thread of its own, take instances from a shared counter.  Each
// instance runs exactly as pggc_cluster would run it alone.
int pggc_cluster_batch(pggc_ctx* ctx, int count, const pggc_graph* graphs, int algo, int objective, double mult,
                       int ex_iters, const uint64_t* seeds, int32_t* const* parts_out, double* objectives_out,
                       int64_t* info_out, int streams) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (count < 0 || streams < 1) throw Error(PGGC_EINVAL, "count must be >= 0 and streams >= 1");
        if (count > 0 && (!graphs || !parts_out || !objectives_out))
            throw Error(PGGC_EINVAL, "graphs, parts_out and objectives_out are required");
        if (ctx->gnn.cfg.mode == GNN_TRAIN) throw Error(PGGC_EINVAL, "predictor training is not available in batches");
        if (count == 0) return;
        if (algo == PGGC_LOUVAIN && ex_iters == 0 && ctx->gnn.cfg.mode == GNN_OFF) {
            // one disjoint union: every kernel runs once over all instances
            Ctx& c = ctx->c;
            c.sync();
            std::vector<DGraph> gs(count);
            std::vector<RefineData> init(count);
            std::vector<i32*> outs(count);
            i64 nmax = 1;
            for (int k = 0; k < count; k++) {
                gs[k] = view_of(&graphs[k]);
                if (!gs[k].vw) throw Error(PGGC_EINVAL, "instance " + std::to_string(k) + ": vertex weights required");
                if (gs[k].n > 0 && !parts_out[k])
                    throw Error(PGGC_EINVAL, "instance " + std::to_string(k) + ": null output clustering");
                if (objective == PGGC_OBJ_MOD || objective == PGGC_OBJ_CPM) {  // unit edge weights
                    gs[k].w = nullptr;
                    gs[k].wk = W_UNIT;
                }
                outs[k] = parts_out[k];
                nmax = std::max(nmax, gs[k].n);
            }
            {
                Buf<i64> scratch(c, nmax + 1);
                for (int k = 0; k < count; k++) {
                    RefineData rd;
                    rd.vol = scratch;
                    rd.cap = std::max<i64>(gs[k].n, 1);
                    state_init(c, gs[k], objective, mult, rd);
                    init[k] = rd;
                }
            }
            std::vector<RefineData> fin;
            std::vector<BatchInfo> bi;
            louvain_batch(c, ctx->ms, gs, init, outs, fin, bi);
            for (int k = 0; k < count; k++) {
                objectives_out[k] = fin[k].objective();
                if (info_out) {
                    int64_t* o = info_out + 8 * (int64_t)k;
                    o[0] = fin[k].label_count;
                    o[1] = fin[k].uncut;
                    o[2] = bi[k].levels;
                    o[3] = bi[k].mv.sweeps;
                    o[4] = bi[k].mv.moves;
                    o[5] = 1;
                    o[6] = bi[k].mv.cands;
                    o[7] = bi[k].total_nnz;
                }
            }
            return;
        }
        const int S = std::min(streams, count);
        if (S == 0) return;
        // inputs may still be in flight on the caller's stream; the workers
        // run on streams of their own
        ctx->c.sync();
        // Workers share the caller's stream: kernels run in submission order
        // while the host threads overlap their launches and readbacks.  (With
        // a stream per worker, concurrently executing instances gave
        // different clusterings from their single runs on gfx950.)
        while ((int)ctx->workers.size() < S) {
            pggc_ctx* w = nullptr;
            const int rc = pggc_ctx_create(&w, ctx->c.stream);
            if (rc != PGGC_OK) throw Error(rc, g_err);
            ctx->workers.push_back(w);
        }
        // the caller's knobs and predictor settings apply to every instance
        std::vector<float> prm;
        const bool prune = ctx->gnn.cfg.mode == GNN_PRUNE;
        if (prune) {
            prm.resize(GNN_P);
            PGGC_HIP(hipMemcpyAsync(prm.data(), ctx->gnn.prm.p, GNN_P * sizeof(float), hipMemcpyDeviceToHost,
                                    ctx->c.stream));
            ctx->c.sync();
        }
        for (int k = 0; k < S; k++) {
            pggc_ctx* w = ctx->workers[k];
            copy_knobs(w->c, ctx->c);
            if (prune) {
                w->gnn.set_params(w->c, prm.data());
                w->gnn.cfg = ctx->gnn.cfg;
            } else {
                w->gnn.cfg.mode = GNN_OFF;
            }
        }
        std::atomic<int> next{0};
        std::mutex mu;
        int first_rc = PGGC_OK;
        std::string first_msg;
        auto work = [&](int k) {
            if (hipSetDevice(ctx->c.device) != hipSuccess) {
                std::lock_guard<std::mutex> l(mu);
                if (first_rc == PGGC_OK) {
                    first_rc = PGGC_EHIP;
                    first_msg = "hipSetDevice failed in a batch worker";
                }
                return;
            }
            pggc_ctx* w = ctx->workers[k];
            while (true) {
                {
                    std::lock_guard<std::mutex> l(mu);
                    if (first_rc != PGGC_OK) return;
                }
                const int x = next.fetch_add(1);
                if (x >= count) return;
                const int rc = pggc_cluster(w, &graphs[x], algo, objective, mult, ex_iters,
                                            seeds ? seeds[x] : (uint64_t)x, parts_out[x], &objectives_out[x],
                                            info_out ? info_out + 8 * (int64_t)x : nullptr);
                if (rc != PGGC_OK) {
                    std::lock_guard<std::mutex> l(mu);
                    if (first_rc == PGGC_OK) {
                        first_rc = rc;
                        first_msg = "instance " + std::to_string(x) + ": " + g_err;
                    }
                    return;
                }
            }
        };
        std::vector<std::thread> th;
        for (int k = 0; k < S; k++) th.emplace_back(work, k);
        for (auto& t : th) t.join();
        if (first_rc != PGGC_OK) throw Error(first_rc, first_msg);
    });
}

int pggc_cut_signature(pggc_ctx* ctx, const pggc_graph* gg, const int32_t* part, uint32_t* sig_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        DGraph g = view_of(gg);
        if (!part || (!sig_out && g.nnz > 0)) throw Error(PGGC_EINVAL, "null array");
        cut_signature(ctx->c, g, part, sig_out);
        ctx->c.sync();
    });
}

int pggc_signature_diff(pggc_ctx* ctx, const uint32_t* a, const uint32_t* b, int64_t words, int64_t* diff_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (words < 0 || !diff_out || (words > 0 && (!a || !b))) throw Error(PGGC_EINVAL, "bad signature arguments");
        *diff_out = signature_diff(ctx->c, a, b, words);
    });
}

int pggc_intersect(pggc_ctx* ctx, int64_t n, const int32_t* c1, const int32_t* c2, int64_t l2, int32_t* out,
                   int64_t* labels_out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (n < 0 || l2 < 1 || (n > 0 && (!c1 || !c2 || !out))) throw Error(PGGC_EINVAL, "bad intersection arguments");
        const i64 L = intersect_clusterings(ctx->c, n, c1, c2, l2, out);
        if (labels_out) *labels_out = L;
    });
}

int pggc_graph_from_edges(pggc_ctx* ctx, int64_t n, int64_t m, const int32_t* src, const int32_t* dst,
                          const int32_t* w, pggc_devgraph** out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!out) throw Error(PGGC_EINVAL, "null output");
        std::unique_ptr<pggc_devgraph> h(new pggc_devgraph());
        try {
            edges_csr(ctx->c, n, m, src, dst, w, h->g);
        } catch (...) {
            h->g.release();
            throw;
        }
        *out = h.release();
    });
}

int pggc_rmat(pggc_ctx* ctx, int scale, int edge_factor, double a, double b, double c, uint64_t seed, int32_t wmin,
              int32_t wmax, pggc_devgraph** out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!out) throw Error(PGGC_EINVAL, "null output");
        std::unique_ptr<pggc_devgraph> h(new pggc_devgraph());
        try {
            rmat_csr(ctx->c, scale, edge_factor, a, b, c, seed, wmin, wmax, h->g);
        } catch (...) {
            h->g.release();
            throw;
        }
        *out = h.release();
    });
}

int pggc_lfr(pggc_ctx* ctx, int64_t n, double tau1, double kmean, int32_t kmax, double tau2, int32_t cmin,
             int32_t cmax, double mu, int32_t wmin, int32_t wmax, uint64_t seed, int32_t* gt_out,
             pggc_devgraph** out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!out) throw Error(PGGC_EINVAL, "null output");
        LfrParams p;
        p.n = n;
        p.tau1 = tau1;
        p.kmean = kmean;
        p.kmax = kmax;
        p.tau2 = tau2;
        p.cmin = cmin;
        p.cmax = cmax;
        p.mu = mu;
        p.wmin = wmin;
        p.wmax = wmax;
        p.seed = seed;
        std::unique_ptr<pggc_devgraph> h(new pggc_devgraph());
        try {
            lfr_csr(ctx->c, p, h->g, gt_out);
        } catch (...) {
            h->g.release();
            throw;
        }
        *out = h.release();
    });
}

int pggc_devgraph_view(const pggc_devgraph* g, pggc_graph* view) {
    return guard([&] {
        if (!g || !view) throw Error(PGGC_EINVAL, "null argument");
        view->n = g->g.n;
        view->nnz = g->g.nnz;
        view->row_ptr = g->g.rp;
        view->col = g->g.col;
        view->w = g->g.w;
        view->w64 = nullptr;
        view->vtx_w = nullptr;
    });
}

int pggc_devgraph_copy(pggc_ctx* ctx, const pggc_devgraph* g, int64_t* row_ptr, int32_t* col, int32_t* w) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (!g || !row_ptr || (g->g.nnz > 0 && !col) || (g->g.w && g->g.nnz > 0 && !w))
            throw Error(PGGC_EINVAL, "null argument");
        copy_dev(ctx->c, row_ptr, g->g.rp, (size_t)(g->g.n + 1) * sizeof(i64));
        copy_dev(ctx->c, col, g->g.col, (size_t)g->g.nnz * sizeof(i32));
        if (g->g.w) copy_dev(ctx->c, w, g->g.w, (size_t)g->g.nnz * sizeof(i32));
    });
}

int pggc_devgraph_free(pggc_devgraph* g) {
    return guard([&] {
        if (!g) return;
        g->g.release();
        delete g;
    });
}

int pggc_random_labels(pggc_ctx* ctx, int64_t n, int64_t L, uint64_t seed, int32_t* out) {
    return guard([&] {
        CTX_CHECK(ctx);
        if (n > 0 && !out) throw Error(PGGC_EINVAL, "null output");
        random_labels(ctx->c, n, L, seed, out);
    });
}

}  // extern "C"
