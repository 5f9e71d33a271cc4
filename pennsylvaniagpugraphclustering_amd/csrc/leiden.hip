// Leiden refinement (src/leidenR.cpp): gamma-well-connected flags, seeded
// neighbour choice, pointer-chasing forests, gamma-connectivity splitting
// inside each tree (radix-sorted by (tree, order, id)), contiguous relabel.
#include <algorithm>

#include "device.h"
#include "leiden.h"

namespace pggc {

// :281-287
__global__ void k_well(const i64* vw, const i32* louv, const i64* vol, const i64* pv, i64 n, float gamma,
                       i32* well) {
    const i64 i = thread_item();
    if (i >= n) return;
    i64 wd = vw[i];
    well[i] = ((float)pv[i] >= gamma * (float)wd * (float)(vol[louv[i]] - wd)) ? 1 : 0;
}

// :296-393 (unit weights): first satisfactory neighbour cyclically from a
// hashed start; 1/1000 of the vertices keep themselves.
// list: the vertices of degree < 128 (order1 prefix); larger degrees take
// k_hn_uniform_wave
__global__ void k_hn_uniform(const i64* rp, const i32* col, const i64* vw, const i32* louv, const i32* well,
                             const i32* list, i64 m, float gamma, i32 seed, i32* hn) {
    const i64 x = thread_item();
    if (x >= m) return;
    const i32 i = list[x];
    const i64 s = rp[i], e = rp[i + 1], width = e - s;
    const float multi = gamma * (float)vw[i];
    const u32 h = murmur3_i32((i32)((u32)seed + (u32)i));
    const i32 jx32 = (i32)h;
    if (jx32 % 1000 == 0 || well[i] == 0 || width == 0) {
        hn[i] = (i32)i;
        return;
    }
    const i64 jx = s + (i64)(h % (u32)width);
    const i32 ci = louv[i];
    i32 pick = (i32)i;
    bool found = false;
    for (i64 j = jx; j < e && !found; j++) {
        i32 v = col[j];
        if (ci != louv[v] || well[v] == 0) continue;
        if (1.0f >= multi * (float)vw[v]) { pick = v; found = true; }
    }
    for (i64 j = s; j < jx && !found; j++) {
        i32 v = col[j];
        if (ci != louv[v] || well[v] == 0) continue;
        if (1.0f >= multi * (float)vw[v]) { pick = v; found = true; }
    }
    hn[i] = pick;
}

// the same choice with one wave per vertex: the lanes test 64 consecutive
// positions of the cyclic scan from the hashed start at a time, and the
// first passing position (lowest lane of the first chunk with one) is the
// serial scan's pick
__global__ void __launch_bounds__(256) k_hn_uniform_wave(const i64* rp, const i32* col, const i64* vw,
                                                         const i32* louv, const i32* well, const i32* list, i64 m,
                                                         float gamma, i32 seed, i32* hn) {
    const int lane = threadIdx.x & 63;
    const i64 x = block_id() * 4 + (threadIdx.x >> 6);
    if (x >= m) return;
    const i32 i = list[x];
    const i64 s = rp[i], e = rp[i + 1], width = e - s;
    const float multi = gamma * (float)vw[i];
    const u32 h = murmur3_i32((i32)((u32)seed + (u32)i));
    const i32 jx32 = (i32)h;
    if (jx32 % 1000 == 0 || well[i] == 0 || width == 0) {
        if (lane == 0) hn[i] = (i32)i;
        return;
    }
    const i64 off = (i64)(h % (u32)width);
    const i32 ci = louv[i];
    for (i64 base = 0; base < width; base += 64) {  // wave-uniform loop control
        const i64 t = base + lane;
        i32 ok = 0;
        i32 v = -1;
        if (t < width) {
            const i64 q = off + t;
            v = col[s + (q < width ? q : q - width)];
            ok = ci == louv[v] && well[v] != 0 && 1.0f >= multi * (float)vw[v];
        }
        xlane_fence(ok);  // the values below cross lanes
        xlane_fence(v);
        const u64 bal = __ballot(ok != 0);
        if (bal) {
            const int first = __ffsll((unsigned long long)bal) - 1;
            const i32 pick = __shfl(v, first);
            if (lane == 0) hn[i] = pick;
            return;
        }
    }
    if (lane == 0) hn[i] = (i32)i;
}

// :395-474 (weighted): argmax over neighbours of rand * (w - gamma w_i w_v)^4
// with (.)^4 evaluated as ((double)val^2)^2.  Degree < 128 keeps the last
// maximum (serial path); larger degrees keep the first (MaxFirstLoc).
template <class WT>
__device__ __forceinline__ bool hn_sample(const i32* col, WT W, const i64* vw, const i32* louv, const i32* well, i64 j,
                                          i32 ci, float multi, i32 seed, float& smp) {
    i32 v = col[j];
    if (ci != louv[v] || well[v] == 0) return false;
    float val = fmaf(-multi, (float)vw[v], (float)W(j));
    if (!(val >= 0.0f)) return false;
    float r = (float)murmur3_i32((i32)((u32)seed + (u32)j)) / (float)ORD_MAX;
    double v2 = (double)val * (double)val;
    smp = (float)((double)r * (v2 * v2));
    return true;
}

template <class WT>
__global__ void k_hn_weighted_small(const i64* rp, const i32* col, WT W, const i64* vw, const i32* louv,
                                    const i32* well, const i32* list, i64 m, float gamma, i32 seed, i32* hn) {
    const i64 x = thread_item();
    if (x >= m) return;
    const i32 i = list[x];
    const i64 s = rp[i], e = rp[i + 1];
    const float multi = gamma * (float)vw[i];
    if (murmur3_i32((i32)((u32)seed + (u32)i)) % 10000u == 0 || well[i] == 0) {
        hn[i] = i;
        return;
    }
    const i32 ci = louv[i];
    float m_ = 0.0f;
    i64 am = e;
    for (i64 j = s; j < e; j++) {
        float smp;
        if (hn_sample(col, W, vw, louv, well, j, ci, multi, seed, smp) && smp >= m_) {
            m_ = smp;
            am = j;
        }
    }
    hn[i] = (am < e) ? col[am] : i;
}

template <class WT>
__global__ void __launch_bounds__(256) k_hn_weighted_large(const i64* rp, const i32* col, WT W, const i64* vw,
                                                           const i32* louv, const i32* well, const i32* list, i64 m,
                                                           float gamma, i32 seed, i32* hn) {
    const int lane = threadIdx.x & 63;
    const i64 x = block_id() * 4 + (threadIdx.x >> 6);
    if (x >= m) return;
    const i32 i = list[x];
    const i64 s = rp[i], e = rp[i + 1];
    const float multi = gamma * (float)vw[i];
    if (murmur3_i32((i32)((u32)seed + (u32)i)) % 10000u == 0 || well[i] == 0) {
        if (lane == 0) hn[i] = i;
        return;
    }
    const i32 ci = louv[i];
    u64 best = 0;
    for (i64 base = s; base < e; base += 64) {
        const i64 j = base + lane;
        if (j >= e) continue;
        float smp;
        if (hn_sample(col, W, vw, louv, well, j, ci, multi, seed, smp)) {
            u64 k = ((u64)ord_f32(smp) << 32) | (u64)(0xFFFFFFFFu - (u32)(j - s));
            best = k > best ? k : best;
        }
    }
    best = wave_max_u64(best);
    if (lane == 0) hn[i] = best ? col[s + (i64)(0xFFFFFFFFu - (u32)(best & 0xffffffffull))] : i;
}

// :482-494
__global__ void k_order(i64 n, i32 seed, i32* order) {
    const i64 i = thread_item();
    if (i >= n) return;
    i32 o = (i32)murmur3_i32((i32)((u32)i + (u32)seed));
    i32 r = o % SPLIT;
    order[i] = r < 0 ? -r : r;
}
__global__ void k_unselect(i64 n, const i32* order, i32* hn) {
    const i64 i = thread_item();
    if (i >= n) return;
    if (order[i] == order[hn[i]]) hn[i] = (i32)i;
}

// find_trees (:212-260)
__global__ void k_trees1(i64 n, const i32* hn, const i32* order, i32* vcmap) {
    const i64 i = thread_item();
    if (i >= n) return;
    i32 now = (i32)i, then = hn[i];
    if (now == then) {
        vcmap[i] = (i32)i;
        return;
    }
    while (order[now] > order[then]) {
        now = then;
        then = hn[then];
    }
    if (now != i) {
        vcmap[i] = now;
        vcmap[now] = now;
    }
}
__global__ void k_trees_flip(i64 n, const i32* vcmap, i32* order) {
    const i64 i = thread_item();
    if (i >= n) return;
    if (vcmap[i] == ORD_MAX) order[i] = 2 * SPLIT - order[i];
}
__global__ void k_trees2(i64 n, const i32* hn, const i32* order, i32* vcmap) {
    const i64 i = thread_item();
    if (i >= n) return;
    if (vcmap[i] != ORD_MAX) return;
    i32 now = (i32)i, then = hn[i];
    while (__hip_atomic_load(&vcmap[now], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ORD_MAX &&
           order[now] > order[then]) {
        now = then;
        then = hn[then];
    }
    vcmap[i] = __hip_atomic_load(&vcmap[now], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ensure_gamma_connectivity (:41-187)
__global__ void k_gkeys(i64 n, const i32* vcmap, const i32* order, u64* key, i32* val) {
    const i64 i = thread_item();
    if (i >= n) return;
    key[i] = ((u64)(u32)vcmap[i] << 31) | (u64)(u32)order[i];
    val[i] = (i32)i;
}
__global__ void k_gstarts(i64 n, const i32* ids, const i32* vcmap, i64* cbeg, const i64* vw, i64* wsorted) {
    const i64 x = thread_item();
    if (x >= n) return;
    i32 i = ids[x];
    i32 c = vcmap[i];
    if (x == 0 || vcmap[ids[x - 1]] != c) cbeg[c] = x;
    wsorted[x] = vw[i];
}

// inner[i] = weight of i's edges to earlier members (by order) of its
// refined cluster; edge chunks (most rows are short), per-row partial sums
// added with atomics into the zeroed array (exact integer sums)
template <class WT>
__global__ void __launch_bounds__(EB) k_inner(const i64* rp, const i32* col, WT W, const i32* vcmap,
                                              const i32* order, i64 n, i64 nnz, i64* inner) {
    ChunkRows cr;
    chunk_rows(rp, n, nnz, cr);
    const i64 j0 = cr.eb + (i64)threadIdx.x * EPT;
    if (j0 >= cr.ee) return;
    i64 r = upper_row(rp, cr.rlo, cr.rhi, j0);
    i64 nxt = rp[r + 1];
    i32 ci = vcmap[r], oi = order[r];
    i64 acc = 0;
    for (int k = 0; k < EPT; k++) {
        const i64 j = j0 + k;
        if (j >= cr.ee) break;
        if (j >= nxt) {
            if (acc) atomicAdd((unsigned long long*)&inner[r], (unsigned long long)acc);
            acc = 0;
            r = advance_row(rp, r, cr.rhi, j, nxt);
            ci = vcmap[r];
            oi = order[r];
        }
        const i32 v = col[j];
        if (vcmap[v] == ci && order[v] < oi) acc += W(j);
    }
    if (acc) atomicAdd((unsigned long long*)&inner[r], (unsigned long long)acc);
}

__global__ void k_outer_vals(i64 n, const i32* ids, const i64* pv, const i64* inner, i64* ov) {
    const i64 x = thread_item();
    if (x >= n) return;
    i32 i = ids[x];
    ov[x] = pv[i] - 2 * inner[i];
}

__global__ void k_breaks(i64 n, const i32* ids, const i32* vcmap, const i32* louv, const i64* cbeg, const i64* tsz,
                         const i64* osc, const i64* inner, const i64* vw, const i64* vol, float gamma,
                         long long* breakers) {
    const i64 x = thread_item();
    if (x >= n) return;
    i32 i = ids[x];
    i32 c = vcmap[i];
    i64 cb = cbeg[c];
    i64 prev = tsz[x] - tsz[cb];
    if ((float)inner[i] < gamma * (float)vw[i] * (float)prev) atomicMin(&breakers[c], (long long)x);
    i64 cur = tsz[x] + vw[i] - tsz[cb];
    i64 other = vol[louv[i]] - cur;
    i64 out = osc[x + 1] - osc[cb];
    if ((float)out < gamma * (float)cur * (float)other) atomicMin(&breakers[c], (long long)(x + 1));
}

__global__ void k_break_apply(i64 n, const i32* ids, const long long* breakers, i32* vcmap) {
    const i64 x = thread_item();
    if (x >= n) return;
    i32 i = ids[x];
    i32 c = vcmap[i];
    if (x >= breakers[c]) vcmap[i] = i;
}

// contigitize_clusters (:189-210)
__global__ void k_root_flags(i64 n, const i32* vcmap, i32* f) {
    const i64 u = thread_item();
    if (u >= n) return;
    f[u] = vcmap[u] == u ? 1 : 0;
}
__global__ void k_root_relabel(i64 n, const i64* rank, const i32* in, i32* out) {
    const i64 u = thread_item();
    if (u >= n) return;
    out[u] = (i32)rank[in[u]];
}

static int bitsfor(i64 v) {
    int b = 0;
    while (b < 62 && ((i64)1 << b) <= v) b++;
    return b;
}

i64 leiden_refine(Ctx& c, const DGraph& g, const Orders& ord, const i32* louv, const i64* vol, double lambda,
                  const i64* pv, bool uniform, i32* vcmap_out) {
    const i64 n = g.n;
    const float gamma = (float)lambda;
    const dim3 G = grid_items(n, 256);
    // algorithmic bytes: two adjacency passes (neighbour choice: column,
    // weight, the neighbour's constraint, flag and vertex weight; inner
    // weights: column, weight, the neighbour's tree and order) and about 40
    // int32/int64 vertex arrays read or written by the flag, order, forest,
    // sort, scan, break and relabel steps
    const int wb = g.wk == W_UNIT ? 0 : g.wk == W_I32 ? 4 : 8;
    const int tk = c.kt_begin(KS_LEIDEN);
    Buf<i32> well(c, n + 1), hn(c, n + 1), order(c, n + 1), vcmap(c, n + 1);
    k_well<<<G, 256, 0, c.stream>>>(g.vw, louv, vol, pv, n, gamma, well);
    PGGC_LAUNCH_CHECK();
    if (uniform) {
        const i32 seed = (i32)c.next_draw();
        const i64 big = ord.off_large;  // order1: degree < 128 first
        if (big > 0)
            k_hn_uniform<<<grid_items(big, 256), 256, 0, c.stream>>>(g.rp, g.col, g.vw, louv, well, ord.o1, big,
                                                                     gamma, seed, hn);
        if (n > big)
            k_hn_uniform_wave<<<grid_items(n - big, 4), 256, 0, c.stream>>>(g.rp, g.col, g.vw, louv, well,
                                                                            ord.o1.p + big, n - big, gamma, seed,
                                                                            hn);
        PGGC_LAUNCH_CHECK();
    } else {
        const i32 seed = (i32)c.next_draw();
        const i64 big = ord.off_large;  // order1: degree < 128 first
        PGGC_WDISPATCH(g, {
            if (big > 0)
                k_hn_weighted_small<<<grid_items(big, 256), 256, 0, c.stream>>>(
                    g.rp, g.col, W, g.vw, louv, well, ord.o1, big, gamma, seed, hn);
            if (n > big)
                k_hn_weighted_large<<<grid_items(n - big, 4), 256, 0, c.stream>>>(
                    g.rp, g.col, W, g.vw, louv, well, ord.o1.p + big, n - big, gamma, seed, hn);
        });
        PGGC_LAUNCH_CHECK();
    }
    {
        const i32 seed = (i32)c.next_draw();
        k_order<<<G, 256, 0, c.stream>>>(n, seed, order);
        PGGC_LAUNCH_CHECK();
    }
    k_unselect<<<G, 256, 0, c.stream>>>(n, order, hn);
    fill_i32(c, vcmap, n, ORD_MAX);
    k_trees1<<<G, 256, 0, c.stream>>>(n, hn, order, vcmap);
    k_trees_flip<<<G, 256, 0, c.stream>>>(n, vcmap, order);
    k_trees2<<<G, 256, 0, c.stream>>>(n, hn, order, vcmap);
    PGGC_LAUNCH_CHECK();
    // sort vertices by (tree, order, id)
    Buf<u64> key(c, n);
    Buf<i32> ids(c, n);
    k_gkeys<<<G, 256, 0, c.stream>>>(n, vcmap, order, key, ids);
    PGGC_LAUNCH_CHECK();
    radix_sort_pairs(c, key, ids, n, 31 + bitsfor(n));
    Buf<i64> cbeg(c, n + 1), wsorted(c, n + 1), tsz(c, n + 1), inner(c, n + 1), ov(c, n + 1), osc(c, n + 1);
    k_gstarts<<<G, 256, 0, c.stream>>>(n, ids, vcmap, cbeg, g.vw, wsorted);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, wsorted, tsz, n);
    fill_i64(c, inner, n, 0);
    if (g.nnz > 0) {
        PGGC_WDISPATCH(g, (k_inner<<<(int)ceil_div(g.nnz, EC), EB, 0, c.stream>>>(g.rp, g.col, W, vcmap, order, n,
                                                                                g.nnz, inner)));
        PGGC_LAUNCH_CHECK();
    }
    k_outer_vals<<<G, 256, 0, c.stream>>>(n, ids, pv, inner, ov);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, ov, osc, n);
    Buf<i64> breakers(c, n + 1);
    fill_i64(c, breakers, n, n);
    k_breaks<<<G, 256, 0, c.stream>>>(n, ids, vcmap, louv, cbeg, tsz, osc, inner, g.vw, vol, gamma,
                                     (long long*)breakers.p);
    PGGC_LAUNCH_CHECK();
    k_break_apply<<<G, 256, 0, c.stream>>>(n, ids, (const long long*)breakers.p, vcmap);
    PGGC_LAUNCH_CHECK();
    // contiguous ids by root order
    k_root_flags<<<G, 256, 0, c.stream>>>(n, vcmap, well);
    PGGC_LAUNCH_CHECK();
    scan_i32_to_i64(c, well, tsz, n);
    k_root_relabel<<<G, 256, 0, c.stream>>>(n, tsz, vcmap, vcmap_out);
    PGGC_LAUNCH_CHECK();
    c.kt_end(tk, 2.0 * 8.0 * (n + 1) + (double)g.nnz * (32 + 2 * wb) + 250.0 * n);
    return c.read1(tsz.p + n);
}

}  // namespace pggc
