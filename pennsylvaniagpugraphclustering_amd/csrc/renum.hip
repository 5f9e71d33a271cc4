// Internal degree order of the top level for the local mover.
#include "renum.h"

#include "device.h"

namespace pggc {

// sort key of vertex v: descending degree (the radix sort is stable, so ties
// keep ascending ids)
__global__ void k_renum_keys(const i64* __restrict__ rp, i64 n, u64* __restrict__ key, i32* __restrict__ id) {
    const i64 v = thread_item();
    if (v >= n) return;
    key[v] = (u64)(0x7fffffff - (rp[v + 1] - rp[v]));
    id[v] = (i32)v;
}

// inv[ids[i']] = i'; deg[i'] = degree of ids[i']; vw' = vw[ids]
__global__ void k_renum_index(const i64* __restrict__ rp, const i32* __restrict__ ids, const i64* __restrict__ vw,
                              i64 n, i32* __restrict__ inv, i64* __restrict__ deg, i64* __restrict__ nvw) {
    const i64 x = thread_item();
    if (x >= n) return;
    const i32 v = ids[x];
    inv[v] = (i32)x;
    deg[x] = rp[v + 1] - rp[v];
    nvw[x] = vw[v];
}

// Row i' of the renumbered graph is row ids[i'] with its columns mapped
// through inv, entries in their original order.  A thread copies RN_EPT
// consecutive entries and walks their rows with a monotone row pointer.
constexpr int RN_EPT = 8;
template <class T>
__global__ void __launch_bounds__(256) k_renum_rows(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                    const T* __restrict__ w, const i64* __restrict__ nrp,
                                                    const i32* __restrict__ ids, const i32* __restrict__ inv, i64 n,
                                                    i64 nnz, i32* __restrict__ ncol, T* __restrict__ nw) {
    const i64 e0 = thread_item() * RN_EPT;
    if (e0 >= nnz) return;
    i64 r = upper_row(nrp, 0, n, e0);
    i64 nxt = nrp[r + 1];
    i64 base = rp[ids[r]] - nrp[r];
#pragma unroll
    for (int q = 0; q < RN_EPT; q++) {
        const i64 e = e0 + q;
        if (e >= nnz) break;
        if (e >= nxt) {
            r = advance_row(nrp, r, n, e, nxt);
            base = rp[ids[r]] - nrp[r];
        }
        const i64 j = base + e;
        ncol[e] = inv[col[j]];
        if (w) nw[e] = w[j];
    }
}

__global__ void k_renum_gather(i64 n, const i32* __restrict__ ids, const i32* __restrict__ part,
                               i32* __restrict__ out) {
    const i64 x = thread_item();
    if (x >= n) return;
    out[x] = part[ids[x]];
}
__global__ void k_renum_scatter(i64 n, const i32* __restrict__ ids, const i32* __restrict__ in,
                                i32* __restrict__ part) {
    const i64 x = thread_item();
    if (x >= n) return;
    part[ids[x]] = in[x];
}

void renum_build(Ctx& c, const DGraph& top, Renum& r) {
    if (r.built_for(top)) return;
    const i64 n = top.n, nnz = top.nnz;
    r.key = top.rp;
    r.n = n;
    r.plan = LevelPlan();
    r.ord_built = false;
    r.ids = Buf<i32>(c, std::max<i64>(n, 1));
    r.part = Buf<i32>(c, std::max<i64>(n, 1));
    r.rp = Buf<i64>(c, n + 1);
    r.vw = Buf<i64>(c, std::max<i64>(n, 1));
    r.col = Buf<i32>(c, std::max<i64>(nnz, 1));
    r.w32 = Buf<i32>();
    r.w64 = Buf<i64>();
    Buf<i32> inv(c, std::max<i64>(n, 1));
    Buf<i64> deg(c, std::max<i64>(n, 1));
    if (n > 0) {
        Buf<u64> key(c, n);
        k_renum_keys<<<grid_items(n, 256), 256, 0, c.stream>>>(top.rp, n, key, r.ids);
        PGGC_LAUNCH_CHECK();
        radix_sort_pairs(c, key, r.ids, n, 31);
        k_renum_index<<<grid_items(n, 256), 256, 0, c.stream>>>(top.rp, r.ids, top.vw, n, inv, deg, r.vw);
        PGGC_LAUNCH_CHECK();
    }
    scan_i64(c, deg, r.rp, n);
    DGraph g;
    g.n = n;
    g.nnz = nnz;
    g.rp = r.rp;
    g.col = r.col;
    g.wk = top.wk;
    g.vw = r.vw;
    if (nnz > 0) {
        const dim3 grid = grid_items(ceil_div(nnz, RN_EPT), 256);
        if (top.wk == W_I32) {
            r.w32 = Buf<i32>(c, nnz);
            k_renum_rows<i32><<<grid, 256, 0, c.stream>>>(top.rp, top.col, (const i32*)top.w, r.rp, r.ids, inv, n,
                                                          nnz, r.col, r.w32);
            g.w = r.w32.p;
        } else if (top.wk == W_I64) {
            r.w64 = Buf<i64>(c, nnz);
            k_renum_rows<i64><<<grid, 256, 0, c.stream>>>(top.rp, top.col, (const i64*)top.w, r.rp, r.ids, inv, n,
                                                          nnz, r.col, r.w64);
            g.w = r.w64.p;
        } else {
            k_renum_rows<i32><<<grid, 256, 0, c.stream>>>(top.rp, top.col, nullptr, r.rp, r.ids, inv, n, nnz, r.col,
                                                          nullptr);
        }
        PGGC_LAUNCH_CHECK();
    }
    r.g = g;
}

void renum_gather(Ctx& c, const Renum& r, const i32* in, i32* out) {
    if (r.n == 0) return;
    k_renum_gather<<<grid_items(r.n, 256), 256, 0, c.stream>>>(r.n, r.ids, in, out);
    PGGC_LAUNCH_CHECK();
}

void local_move_renum(Ctx& c, MoveScratch& s, Renum& r, i32* part, RefineData& best, bool is_initial,
                      bool constrained, i32* cons, bool SA, MoveStats* st) {
    const i64 n = r.n;
    renum_gather(c, r, part, r.part);
    local_move(c, s, r.plan, r.g, r.part, best, is_initial, constrained, cons, SA, st, r.ids);
    if (n > 0) {
        k_renum_scatter<<<grid_items(n, 256), 256, 0, c.stream>>>(n, r.ids, r.part, part);
        PGGC_LAUNCH_CHECK();
    }
}

}  // namespace pggc
