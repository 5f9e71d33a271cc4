// Batched pLouvain: many independent graphs clustered as one disjoint union
// graph, so every kernel of the multilevel run (ordering, scoring,
// afterburner, moves, contraction) is one launch over all instances.  The
// reference runs its trials one after another (app/cluster_main.cpp:189-204);
// here the run of each instance still follows louvain_part
// (src/clustering_methods.cpp:139-222) step for step -- the per-instance
// decisions are taken on the host from per-instance counts -- so each
// instance ends exactly as it would alone.
//
// Layout: instance k owns the vertex range [voff[k], voff[k+1]) of every
// level and the label range [loff[k], loff[k] + n0[k]) (n0 = its top-level
// size) at every level, so its labels are its single-run labels plus loff[k]
// and projections between levels need no remapping.  An instance that stops
// contracting stays in the deeper union levels as an unchanged copy that no
// local move touches.
#include <algorithm>
#include <memory>

#include "batch.h"
#include "device.h"

namespace pggc {

// union rows of instance k: rp_u[voff + i] = rp_k[i] + eoff, columns + voff
__global__ void k_union_rows(i64 n, const i64* __restrict__ rp, i64 eoff, i64* __restrict__ out) {
    const i64 i = thread_item();
    if (i >= n) return;
    out[i] = rp[i] - rp[0] + eoff;
}
__global__ void k_union_cols(i64 m, const i32* __restrict__ col, i32 voff, i32* __restrict__ out) {
    const i64 j = thread_item();
    if (j >= m) return;
    out[j] = col[j] + voff;
}
__global__ void k_union_w32(i64 m, const i32* __restrict__ w, i64* __restrict__ out) {
    const i64 j = thread_item();
    if (j >= m) return;
    out[j] = w[j];
}
__global__ void k_set1(i64* p, i64 v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
// initial labels of a level: the single-run labels (vertex ids within the
// instance) plus the instance's label offset
__global__ void k_batch_init_labels(i64 n, const i32* __restrict__ inst, const i32* __restrict__ loc,
                                    const i64* __restrict__ loff, i32* __restrict__ part) {
    const i64 v = thread_item();
    if (v >= n) return;
    part[v] = (i32)(loff[inst[v]] + loc[v]);
}
// contraction map: an instance that contracts maps label l to coff + (l -
// loff); any other keeps its vertices (coff + local id)
__global__ void k_batch_vcmap(i64 n, const i32* __restrict__ inst, const i32* __restrict__ loc,
                              const i32* __restrict__ labels, const i32* __restrict__ contracts,
                              const i64* __restrict__ loff, const i64* __restrict__ coff, i32* __restrict__ vcmap) {
    const i64 v = thread_item();
    if (v >= n) return;
    const i32 k = inst[v];
    vcmap[v] = (i32)(coff[k] + (contracts[k] ? (i64)labels[v] - loff[k] : (i64)loc[v]));
}
// coarse vertex weights: the volumes of a contracting instance's clusters,
// else the vertex weights of the level
__global__ void k_batch_coarse_vw(i64 nc, const i32* __restrict__ cinst, const i32* __restrict__ cloc,
                                  const i32* __restrict__ contracts, const i64* __restrict__ loff,
                                  const i64* __restrict__ voff, const i64* __restrict__ vol,
                                  const i64* __restrict__ vw, i64* __restrict__ out) {
    const i64 u = thread_item();
    if (u >= nc) return;
    const i32 k = cinst[u];
    out[u] = contracts[k] ? vol[loff[k] + cloc[u]] : vw[voff[k] + cloc[u]];
}
// labels[v] = coarse labels[vcmap[v]] for vertices of masked instances
__global__ void k_batch_project(i64 n, const i32* __restrict__ inst, const i32* __restrict__ mask,
                                const i32* __restrict__ vcmap, const i32* __restrict__ coarse, i32* __restrict__ labels) {
    const i64 v = thread_item();
    if (v >= n) return;
    if (mask[inst[v]]) labels[v] = coarse[vcmap[v]];
}
// result of instance k in its own ids
__global__ void k_batch_out(i64 n, const i32* __restrict__ labels, i64 loff, i32* __restrict__ out) {
    const i64 i = thread_item();
    if (i >= n) return;
    out[i] = (i32)(labels[i] - loff);
}

template <class T>
static void to_device(Ctx& c, Buf<T>& dst, const std::vector<T>& src) {
    dst = Buf<T>(c, std::max<size_t>(src.size(), 1));
    copy_h2d(c, dst.p, src.data(), src.size() * sizeof(T));
}

__global__ void k_at_offsets(i64 K1, const i64* __restrict__ a, const i64* __restrict__ idx, i64* __restrict__ out) {
    const i64 k = thread_item();
    if (k >= K1) return;
    out[k] = a[idx[k]];
}

// rp at every instance boundary: the instances' edge offsets
static std::vector<i64> read_offsets(Ctx& c, const i64* rp, const BatchLevel& bl) {
    const i64 K1 = bl.K + 1;
    Buf<i64> d(c, K1);
    k_at_offsets<<<grid_items(K1, 256), 256, 0, c.stream>>>(K1, rp, bl.dvoff, d);
    PGGC_LAUNCH_CHECK();
    std::vector<i64> e(K1);
    PGGC_HIP(hipMemcpyAsync(e.data(), d.p, K1 * sizeof(i64), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return e;
}

namespace {
struct BLevel {
    DGraph g;
    std::unique_ptr<OwnedGraph> own;
    LevelPlan plan;
    BatchLevel bl;
    Buf<i32> labels;  // labels after the level's local move on the way down
    Buf<i32> vcmap;   // vertex -> vertex of the next level
    std::vector<i64> eoff;  // instance edge offsets
};
}  // namespace

void louvain_batch(Ctx& c, MoveScratch& ms, const std::vector<DGraph>& gs, const std::vector<RefineData>& init,
                   const std::vector<i32*>& outs, std::vector<RefineData>& final_rd, std::vector<BatchInfo>& info) {
    const i64 K = (i64)gs.size();
    if (K == 0) return;
    // ---- top level: the union graph
    bool weighted = false;
    for (const DGraph& g : gs) weighted = weighted || g.wk != W_UNIT;
    std::vector<i64> voff(K + 1, 0), eoff(K + 1, 0);
    for (i64 k = 0; k < K; k++) {
        voff[k + 1] = voff[k] + gs[k].n;
        eoff[k + 1] = eoff[k] + gs[k].nnz;
    }
    const i64 N = voff[K], M = eoff[K];
    if (N >= 0x7fffffff) throw Error(-1, "batch: the union of the graphs exceeds int32 vertex ids");
    std::unique_ptr<OwnedGraph> top(new OwnedGraph());
    top->n = N;
    top->nnz = M;
    top->rp = Buf<i64>(c, N + 1);
    top->col = Buf<i32>(c, std::max<i64>(M, 1));
    top->vw = Buf<i64>(c, std::max<i64>(N, 1));
    if (weighted) top->w = Buf<i64>(c, std::max<i64>(M, 1));
    for (i64 k = 0; k < K; k++) {
        const DGraph& g = gs[k];
        if (g.n > 0) {
            k_union_rows<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.n, g.rp, eoff[k], top->rp.p + voff[k]);
            copy_dev(c, top->vw.p + voff[k], g.vw, g.n * sizeof(i64));
        }
        if (g.nnz > 0) {
            k_union_cols<<<grid_items(g.nnz, 256), 256, 0, c.stream>>>(g.nnz, g.col, (i32)voff[k],
                                                                       top->col.p + eoff[k]);
            if (weighted) {
                if (g.wk == W_I32)
                    k_union_w32<<<grid_items(g.nnz, 256), 256, 0, c.stream>>>(g.nnz, (const i32*)g.w,
                                                                              top->w.p + eoff[k]);
                else if (g.wk == W_I64)
                    copy_dev(c, top->w.p + eoff[k], g.w, g.nnz * sizeof(i64));
                else
                    fill_i64(c, top->w.p + eoff[k], g.nnz, 1);
            }
        }
        PGGC_LAUNCH_CHECK();
    }
    k_set1<<<1, 64, 0, c.stream>>>(top->rp.p + N, M);
    PGGC_LAUNCH_CHECK();
    DGraph g0 = top->view();
    if (!weighted) {
        g0.w = nullptr;
        g0.wk = W_UNIT;
    }
    // ---- per-instance state (cluster_data per instance, one label space)
    BatchState st;
    st.K = K;
    st.loff = voff;
    to_device(c, st.dloff, st.loff);
    st.vol = Buf<i64>(c, std::max<i64>(N, 1));
    copy_dev(c, st.vol, g0.vw, N * sizeof(i64));
    st.rd = init;
    std::vector<float> lamf(K);
    for (i64 k = 0; k < K; k++) lamf[k] = (float)init[k].lambda;
    to_device(c, st.dlam, lamf);
    std::vector<MoveStats> stats(K);
    std::vector<i64> last(K, -1);  // level where an instance stopped contracting
    std::vector<std::vector<i64>> lvl_nnz;
    // ---- down: local moves and contraction until no instance contracts
    std::vector<std::unique_ptr<BLevel>> L;
    L.emplace_back(new BLevel());
    L[0]->g = g0;
    L[0]->own = std::move(top);
    L[0]->bl.build(c, voff);
    st.linst = Buf<i32>(c, std::max<i64>(N, 1));
    copy_dev(c, st.linst, L[0]->bl.inst, N * sizeof(i32));
    while (true) {
        BLevel& lv = *L.back();
        const i64 l = (i64)L.size() - 1;
        const DGraph g = lv.g;
        const std::vector<i64>& vo = lv.bl.voff;
        lv.eoff = read_offsets(c, g.rp, lv.bl);
        std::vector<char> act(K);
        for (i64 k = 0; k < K; k++) act[k] = last[k] < 0;
        lv.labels = Buf<i32>(c, std::max<i64>(g.n, 1));
        if (g.n > 0) {
            k_batch_init_labels<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.n, lv.bl.inst, lv.bl.loc, st.dloff,
                                                                           lv.labels);
            PGGC_LAUNCH_CHECK();
        }
        local_move_batch(c, ms, lv.plan, g, lv.bl, lv.labels, st, true, act, stats);
        std::vector<i32> contracts(K, 0);
        std::vector<i64> coff(K + 1, 0);
        bool any = false;
        for (i64 k = 0; k < K; k++) {
            const i64 nk = vo[k + 1] - vo[k];
            if (act[k]) {
                if (st.rd[k].label_count < nk) {
                    contracts[k] = 1;
                    any = true;
                } else {
                    last[k] = l;
                }
            }
            coff[k + 1] = coff[k] + (contracts[k] ? st.rd[k].label_count : nk);
        }
        if (!any) break;
        Buf<i32> dcon;
        Buf<i64> dcoff;
        to_device(c, dcon, contracts);
        to_device(c, dcoff, coff);
        const i64 nc = coff[K];
        lv.vcmap = Buf<i32>(c, std::max<i64>(g.n, 1));
        if (g.n > 0) {
            k_batch_vcmap<<<grid_items(g.n, 256), 256, 0, c.stream>>>(g.n, lv.bl.inst, lv.bl.loc, lv.labels, dcon,
                                                                     st.dloff, dcoff, lv.vcmap);
            PGGC_LAUNCH_CHECK();
        }
        std::unique_ptr<BLevel> nx(new BLevel());
        std::unique_ptr<OwnedGraph> og(new OwnedGraph());
        contract(c, g, lv.vcmap, nc, *og);
        narrow_weights(c, g, *og);
        nx->bl.build(c, coff);
        og->vw = Buf<i64>(c, std::max<i64>(nc, 1));
        if (nc > 0) {
            k_batch_coarse_vw<<<grid_items(nc, 256), 256, 0, c.stream>>>(nc, nx->bl.inst, nx->bl.loc, dcon, st.dloff,
                                                                         lv.bl.dvoff, st.vol, g.vw, og->vw);
            PGGC_LAUNCH_CHECK();
        }
        nx->g = og->view();
        nx->own = std::move(og);
        L.push_back(std::move(nx));
    }
    // ---- up: project and refine the instances with levels left below
    const i64 D = (i64)L.size();
    for (i64 l = D - 2; l >= 0; l--) {
        BLevel& lv = *L[l];
        std::vector<char> act(K);
        std::vector<i32> hm(K);
        bool any = false;
        for (i64 k = 0; k < K; k++) {
            act[k] = l <= last[k] - 2;
            hm[k] = act[k] ? 1 : 0;
            any = any || act[k];
        }
        if (!any) continue;
        Buf<i32> mask;
        to_device(c, mask, hm);
        if (lv.g.n > 0) {
            k_batch_project<<<grid_items(lv.g.n, 256), 256, 0, c.stream>>>(lv.g.n, lv.bl.inst, mask, lv.vcmap,
                                                                          L[l + 1]->labels, lv.labels);
            PGGC_LAUNCH_CHECK();
        }
        local_move_batch(c, ms, lv.plan, lv.g, lv.bl, lv.labels, st, false, act, stats);
    }
    // ---- results
    for (i64 k = 0; k < K; k++) {
        if (gs[k].n > 0) {
            k_batch_out<<<grid_items(gs[k].n, 256), 256, 0, c.stream>>>(gs[k].n, L[0]->labels.p + voff[k], voff[k],
                                                                       outs[k]);
            PGGC_LAUNCH_CHECK();
        }
    }
    final_rd = st.rd;
    info.assign(K, BatchInfo());
    for (i64 k = 0; k < K; k++) {
        info[k].levels = last[k] + 1;
        info[k].mv = stats[k];
        for (i64 l = 0; l <= last[k] && l < D; l++) info[k].total_nnz += L[l]->eoff[k + 1] - L[l]->eoff[k];
    }
    c.sync();
}

}  // namespace pggc
