// Degree ordering, evaluator, own-cluster weights and contraction kernels.
#include <algorithm>
#include <vector>

#include "batching.h"
#include "device.h"
#include "graph.h"
#include "hash.h"

namespace pggc {

// ------------------------------------------------------ stable K-way partition
constexpr int PT = 256, PI = 8, PTILE = PT * PI;

template <int K>
__global__ void __launch_bounds__(PT) k_part_count(const i32* cls, i64 n, i64* counts, i64 nb) {
    i64 base = (i64)blockIdx.x * PTILE + (i64)threadIdx.x * PI;
    i64 c[K];
#pragma unroll
    for (int k = 0; k < K; k++) c[k] = 0;
#pragma unroll
    for (int q = 0; q < PI; q++) {
        i64 x = base + q;
        if (x < n) {
            int k = cls[x];
#pragma unroll
            for (int kk = 0; kk < K; kk++) c[kk] += (k == kk);
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        i64 s = block_sum_i64<PT>(c[k]);
        if (threadIdx.x == 0) counts[(i64)k * nb + blockIdx.x] = s;
    }
}

template <int K>
__global__ void __launch_bounds__(PT) k_part_write(const i32* cls, i64 n, const i64* offs, i64 nb, i32* out) {
    i64 base = (i64)blockIdx.x * PTILE + (i64)threadIdx.x * PI;
    int ck[PI];
    i64 c[K];
#pragma unroll
    for (int k = 0; k < K; k++) c[k] = 0;
#pragma unroll
    for (int q = 0; q < PI; q++) {
        i64 x = base + q;
        ck[q] = x < n ? cls[x] : -1;
#pragma unroll
        for (int kk = 0; kk < K; kk++) c[kk] += (ck[q] == kk);
    }
    i64 pos[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        i64 tot;
        pos[k] = block_excl_scan_i64(c[k], &tot) + offs[(i64)k * nb + blockIdx.x];
    }
#pragma unroll
    for (int q = 0; q < PI; q++) {
#pragma unroll
        for (int kk = 0; kk < K; kk++)
            if (ck[q] == kk) out[pos[kk]++] = (i32)(base + q);
    }
}

__global__ void k_gather_starts(const i64* offs, i64 nb, int K, i64* starts) {
    int k = threadIdx.x;
    if (k <= K) starts[k] = offs[(i64)k * nb];
}

template <int K>
static void partition_k(Ctx& c, const i32* cls, i64 n, i32* out, i64* starts) {
    i64 nb = ceil_div(std::max<i64>(n, 1), PTILE);
    Buf<i64> counts(c, K * nb + 1);
    k_part_count<K><<<(int)nb, PT, 0, c.stream>>>(cls, n, counts, nb);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, counts, counts, K * nb);
    k_part_write<K><<<(int)nb, PT, 0, c.stream>>>(cls, n, counts, nb, out);
    PGGC_LAUNCH_CHECK();
    Buf<i64> st(c, K + 1);
    k_gather_starts<<<1, 64, 0, c.stream>>>(counts, nb, K, st);
    PGGC_LAUNCH_CHECK();
    const i64* h = c.read(st, K + 1);
    for (int k = 0; k <= K; k++) starts[k] = h[k];
}

void partition_classes(Ctx& c, const i32* cls, i64 n, int K, i32* out, i64* starts) {
    switch (K) {
        case 2: partition_k<2>(c, cls, n, out, starts); break;
        case 3: partition_k<3>(c, cls, n, out, starts); break;
        case 4: partition_k<4>(c, cls, n, out, starts); break;
        case 8: partition_k<8>(c, cls, n, out, starts); break;
        case 11: partition_k<11>(c, cls, n, out, starts); break;
        default: throw Error(-1, "partition_classes: unsupported class count");
    }
}

// ------------------------------------------------------------------ orderings
__global__ void k_order_classes(const i64* rp, i64 n, i32* c1, i32* c2) {
    const i64 i = thread_item();
    if (i >= n) return;
    i64 d = rp[i + 1] - rp[i];
    c1[i] = d < LARGE_CUTOFF ? 0 : (d < MASSIVE_CUTOFF ? 1 : 2);
    c2[i] = d < MID_CUTOFF ? 0 : (d < MASSIVE_CUTOFF ? 1 : 2);
}

void make_orders(Ctx& c, const DGraph& g, Orders& o) {
    i64 n = g.n;
    // row bounds and two class bytes per vertex read, then per partition
    // the classes read twice (count, write) and the ids written
    const int tk = c.kt_begin(KS_ORDER);
    if (o.o1.n < (size_t)n) {
        o.o1 = Buf<i32>(c, std::max<i64>(n, 1));
        o.o2 = Buf<i32>(c, std::max<i64>(n, 1));
    }
    Buf<i32> c1(c, std::max<i64>(n, 1)), c2(c, std::max<i64>(n, 1));
    k_order_classes<<<grid_items(n, 256), 256, 0, c.stream>>>(g.rp, n, c1, c2);
    PGGC_LAUNCH_CHECK();
    i64 s1[4], s2[4];
    partition_k<3>(c, c1, n, o.o1, s1);
    partition_k<3>(c, c2, n, o.o2, s2);
    o.off_large = s1[1];
    o.off_large2 = s1[2];
    o.off_mid = s2[1];
    o.off_mid2 = s2[2];
    c.kt_end(tk, 8.0 * (n + 1) + 8.0 * n + 2.0 * 12.0 * n);
}

// ------------------------------------------------- edge-chunk row reductions
// (ChunkRows / chunk_rows in device.h)

template <class WT>
__global__ void __launch_bounds__(EB) k_eval_uncut(const i64* rp, const i32* col, WT W, const i32* lab, i64 n,
                                                    i64 nnz, i64* uncut) {
    ChunkRows cr;
    chunk_rows(rp, n, nnz, cr);
    i64 j0 = cr.eb + (i64)threadIdx.x * EPT;
    i64 acc = 0;
    if (j0 < cr.ee) {
        i64 r = upper_row(rp, cr.rlo, cr.rhi, j0);
        i64 nxt = rp[r + 1];
        i32 lr = lab[r];
        for (int k = 0; k < EPT; k++) {
            i64 j = j0 + k;
            if (j >= cr.ee) break;
            if (j >= nxt) {
                r = advance_row(rp, r, cr.rhi, j, nxt);
                lr = lab[r];
            }
            if (lab[col[j]] == lr) acc += W(j);
        }
    }
    acc = block_sum_i64<EB>(acc);
    if (threadIdx.x == 0 && acc) atomicAdd((unsigned long long*)uncut, (unsigned long long)acc);
}

__global__ void k_eval_vol(const i32* lab, const i64* vw, i64 n, i64* vol) {
    const i64 i = thread_item();
    if (i >= n) return;
    atomicAdd((unsigned long long*)&vol[lab[i]], (unsigned long long)vw[i]);
}

void evaluate(Ctx& c, const DGraph& g, const i32* labels, i64 L, i64* vol, i64* dev_uncut) {
    fill_i64(c, vol, L, 0);
    fill_i64(c, dev_uncut, 1, 0);
    if (g.nnz > 0) {
        // algorithmic bytes: row offsets, columns, labels of both ends, weights
        const double wb = g.wk == W_UNIT ? 0.0 : (g.wk == W_I32 ? 4.0 : 8.0);
        const double bytes = 8.0 * (g.n + 1) + 4.0 * g.n + (8.0 + wb) * g.nnz;
        const int tk = c.kt_begin(KS_EVAL);
        PGGC_WDISPATCH(g, (k_eval_uncut<<<(int)ceil_div(g.nnz, EC), EB, 0, c.stream>>>(g.rp, g.col, W, labels, g.n,
                                                                                       g.nnz, dev_uncut)));
        c.kt_end(tk, bytes);
        PGGC_LAUNCH_CHECK();
    }
    if (g.n > 0) {
        k_eval_vol<<<grid_items(g.n, 256), 256, 0, c.stream>>>(labels, g.vw, g.n, vol);
        PGGC_LAUNCH_CHECK();
    }
}

template <class WT>
__global__ void __launch_bounds__(EB) k_own_weight(const i64* rp, const i32* col, WT W, const i32* lab, i64 n,
                                                    i64 nnz, i64* out) {
    ChunkRows cr;
    chunk_rows(rp, n, nnz, cr);
    i64 j0 = cr.eb + (i64)threadIdx.x * EPT;
    if (j0 >= cr.ee) return;
    i64 r = upper_row(rp, cr.rlo, cr.rhi, j0);
    i64 nxt = rp[r + 1];
    i32 lr = lab[r];
    i64 acc = 0;
    for (int k = 0; k < EPT; k++) {
        i64 j = j0 + k;
        if (j >= cr.ee) break;
        if (j >= nxt) {
            if (acc) atomicAdd((unsigned long long*)&out[r], (unsigned long long)acc);
            acc = 0;
            r = advance_row(rp, r, cr.rhi, j, nxt);
            lr = lab[r];
        }
        if (lab[col[j]] == lr) acc += W(j);
    }
    if (acc) atomicAdd((unsigned long long*)&out[r], (unsigned long long)acc);
}

void own_weight(Ctx& c, const DGraph& g, const i32* labels, i64* out) {
    fill_i64(c, out, g.n, 0);
    if (g.nnz == 0) return;
    PGGC_WDISPATCH(g, (k_own_weight<<<(int)ceil_div(g.nnz, EC), EB, 0, c.stream>>>(g.rp, g.col, W, labels, g.n,
                                                                                   g.nnz, out)));
    PGGC_LAUNCH_CHECK();
}

__global__ void k_coarsen_vw(const i32* map, const i64* vw, i64 n, i64* out) {
    const i64 i = thread_item();
    if (i >= n) return;
    atomicAdd((unsigned long long*)&out[map[i]], (unsigned long long)vw[i]);
}

void coarsen_vertex_weights(Ctx& c, const DGraph& g, const i32* map, i64 nc, i64* out) {
    fill_i64(c, out, nc, 0);
    if (g.n == 0) return;
    k_coarsen_vw<<<grid_items(g.n, 256), 256, 0, c.stream>>>(map, g.vw, g.n, out);
    PGGC_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- contraction
// members and summed degrees per coarse vertex; a block's 256 vertices are
// merged per coarse id in LDS first (large clusters take many members of
// every block), then each distinct id gets one pair of HBM atomics
constexpr int CC_SLOTS = 512;
__global__ void __launch_bounds__(256) k_ccount(const i64* rp, const i32* vcmap, i64 n, i64* cnt, i64* bound) {
    __shared__ i32 lk[CC_SLOTS];
    __shared__ i64 lc[CC_SLOTS], lb[CC_SLOTS];
    for (int q = threadIdx.x; q < CC_SLOTS; q += 256) {
        lk[q] = EMPTY_KEY;
        lc[q] = 0;
        lb[q] = 0;
    }
    __syncthreads();
    const i64 i = thread_item();
    if (i < n) {
        const i32 u = vcmap[i];
        const i64 d = rp[i + 1] - rp[i];
        u32 h = slot_hash(u, CC_SLOTS - 1);
        while (true) {  // at most 256 keys in 512 slots: an empty slot is always found
            const i32 old = atomicCAS(&lk[h], EMPTY_KEY, u);
            if (old == EMPTY_KEY || old == u) break;
            h = (h + 1) & (CC_SLOTS - 1);
        }
        atomicAdd((unsigned long long*)&lc[h], 1ull);
        atomicAdd((unsigned long long*)&lb[h], (unsigned long long)d);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < CC_SLOTS; q += 256) {
        const i32 u = lk[q];
        if (u == EMPTY_KEY) continue;
        atomicAdd((unsigned long long*)&cnt[u], (unsigned long long)lc[q]);
        atomicAdd((unsigned long long*)&bound[u], (unsigned long long)lb[q]);
    }
}

// member lists: a block's vertices take ranks per coarse id in LDS, each
// distinct id reserves its run of the list with one HBM atomic, and every
// vertex writes at run start + rank (the order inside a list is free: the
// merges that read it are order-independent)
__global__ void __launch_bounds__(256) k_cscatter(const i32* vcmap, i64 n, i64* pos, i32* members) {
    __shared__ i32 lk[CC_SLOTS], lc[CC_SLOTS];
    __shared__ i64 lb[CC_SLOTS];
    for (int q = threadIdx.x; q < CC_SLOTS; q += 256) {
        lk[q] = EMPTY_KEY;
        lc[q] = 0;
    }
    __syncthreads();
    const i64 i = thread_item();
    u32 h = 0;
    i32 r = 0;
    if (i < n) {
        const i32 u = vcmap[i];
        h = slot_hash(u, CC_SLOTS - 1);
        while (true) {  // at most 256 keys in 512 slots
            const i32 old = atomicCAS(&lk[h], EMPTY_KEY, u);
            if (old == EMPTY_KEY || old == u) break;
            h = (h + 1) & (CC_SLOTS - 1);
        }
        r = atomicAdd(&lc[h], 1);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < CC_SLOTS; q += 256)
        if (lk[q] != EMPTY_KEY) lb[q] = (i64)atomicAdd((unsigned long long*)&pos[lk[q]], (unsigned long long)lc[q]);
    __syncthreads();
    if (i < n) members[lb[h] + r] = (i32)i;
}

constexpr i64 C_SMALL = 256;   // bound <= C_SMALL: 64-thread block, 512-slot LDS table
constexpr i64 C_HALF = 1024;   // bound <= C_HALF: 256-thread block, 2048-slot LDS table
constexpr i64 C_MID = 2048;    // bound <= C_MID: 256-thread block, 4096-slot LDS table

// classes: 0..2 LDS merge (small, half, mid), 3 big rows, 4 empty rows
__global__ void k_cclass(const i64* bound, i64 nc, i32* cls) {
    const i64 u = thread_item();
    if (u >= nc) return;
    i64 b = bound[u];
    cls[u] = b == 0 ? 4 : (b <= C_SMALL ? 0 : (b <= C_HALF ? 1 : (b <= C_MID ? 2 : 3)));
}

__host__ __device__ __forceinline__ i64 table_slots(i64 bound, i64 nc) {
    i64 cap = bound < nc ? bound : nc;
    i64 s = 2;
    while (s < 2 * cap) s <<= 1;
    return s;
}

__global__ void k_goff_sizes(const i32* list, i64 m, const i64* bound, i64 nc, i64* sz) {
    const i64 x = thread_item();
    if (x >= m) return;
    sz[x] = table_slots(bound[list[x]], nc);
}

// One block per coarse vertex: hash-merge the members' edges by coarse
// neighbour, then write the distinct (neighbour, weight) pairs to the
// temporary row at toff[u] (sorted when the table lives in LDS).
template <int BS, int SLOTS, bool GLOBAL, class WT>
__global__ void __launch_bounds__(BS) k_cagg(const i64* rp, const i32* col, WT W, const i32* vcmap, i64 nc,
                                              const i32* list, i64 nlist, const i64* mstart, const i32* members,
                                              const i64* bound, i32* gkeys, i64* gvals, const i64* goff,
                                              const i64* toff, i32* tcol, i64* tw, i64* ccnt) {
    __shared__ i32 skeys[GLOBAL ? 1 : SLOTS];
    __shared__ i64 svals[GLOBAL ? 1 : SLOTS];
    __shared__ i64 pre[BS];
    __shared__ i64 rs[BS];
    __shared__ int s_cnt;
    const int t = threadIdx.x;
    const i64 li = block_id();
    if (li >= nlist) return;
    {
        const i32 u = list[li];
        const i64 S = table_slots(bound[u], nc);
        const u32 mask = (u32)(S - 1);
        i32* keys = GLOBAL ? gkeys + goff[li] : skeys;
        i64* vals = GLOBAL ? gvals + goff[li] : svals;
        for (i64 s = t; s < S; s += BS) {
            keys[s] = EMPTY_KEY;
            vals[s] = 0;
        }
        if (t == 0) s_cnt = 0;
        __syncthreads();
        const i64 m0 = mstart[u], m1 = mstart[u + 1];
        for (i64 mb = m0; mb < m1; mb += BS) {
            i64 mi = mb + t;
            i64 d = 0, st = 0;
            if (mi < m1) {
                i32 v = members[mi];
                st = rp[v];
                d = rp[v + 1] - st;
            }
            i64 tot;
            i64 ex = block_excl_scan_i64(d, &tot);
            pre[t] = ex;
            rs[t] = st;
            __syncthreads();
            // groups of CR edges per thread: locate and gather all of them
            // (independent loads in flight together), then insert
            constexpr int CR = 4;
            for (i64 eb = 0; eb < tot; eb += (i64)BS * CR) {  // block-uniform loop control
                i32 cv[CR];
                i64 cw[CR];
#pragma unroll
                for (int r = 0; r < CR; r++) {
                    const i64 e = eb + (i64)r * BS + t;
                    cv[r] = u;  // no entry
                    cw[r] = 0;
                    if (e < tot) {
                        int lo = 0, hi = BS;
                        while (hi - lo > 1) {
                            int md = (lo + hi) >> 1;
                            if (pre[md] <= e) lo = md;
                            else hi = md;
                        }
                        const i64 j = rs[lo] + (e - pre[lo]);
                        cv[r] = vcmap[col[j]];
                        cw[r] = W(j);
                    }
                }
#pragma unroll
                for (int r = 0; r < CR; r++) {
                    if (cv[r] == u) continue;
                    if (GLOBAL) table_add_global(keys, vals, mask, cv[r], cw[r]);
                    else table_add_lds(keys, vals, mask, cv[r], cw[r]);
                }
            }
            __syncthreads();
        }
        if (!GLOBAL) {
            // Sort only the distinct entries: move them to the front of the
            // table (through registers, so no second LDS buffer), then rank
            // them directly when one block's threads suffice (keys are
            // distinct) or bitonic-sort the next power of two above them.
            constexpr int R = SLOTS / BS;
            i32 ek[R];
            i64 ev[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const i64 q = t + (i64)r * BS;
                ek[r] = q < S ? keys[q] : EMPTY_KEY;
                ev[r] = q < S ? vals[q] : 0;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; r++) {
                if (ek[r] != EMPTY_KEY) {
                    const int pos = atomicAdd(&s_cnt, 1);
                    keys[pos] = ek[r];
                    vals[pos] = ev[r];
                }
            }
            __syncthreads();
            const int c = s_cnt;
            const i64 base = toff[u];
            if (c <= BS) {
                if (t < c) {
                    const i32 mk = keys[t];
                    int rank = 0;
                    for (int j = 0; j < c; j++) rank += keys[j] < mk ? 1 : 0;
                    tcol[base + rank] = mk;
                    tw[base + rank] = vals[t];
                }
            } else {
                int P = 1;
                while (P < c) P <<= 1;  // P <= S: the table has at least twice the distinct keys
                for (int q = c + t; q < P; q += BS) {
                    keys[q] = 0x7fffffff;
                    vals[q] = 0;
                }
                __syncthreads();
                bitonic_sort_lds<BS>(keys, vals, P);
                for (int q = t; q < c; q += BS) {
                    tcol[base + q] = keys[q];
                    tw[base + q] = vals[q];
                }
            }
            if (t == 0) ccnt[u] = c;
        } else {
            __threadfence();
            __syncthreads();
            i64 base = toff[u];
            for (i64 s = t; s < S; s += BS) {
                i32 k = __hip_atomic_load(&keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (k != EMPTY_KEY) {
                    int p = atomicAdd(&s_cnt, 1);
                    tcol[base + p] = k;
                    tw[base + p] = __hip_atomic_load(&vals[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
            if (t == 0) ccnt[u] = s_cnt;
        }
        __syncthreads();
    }
}

// Big coarse rows (bound > C_MID): the flattened member-edge range of each
// such row is split into work items of BCH_E edges, each item a block that
// merges its edges in LDS and then inserts the distinct pairs into the row's
// HBM hash table; a second pass splits each table into slot items of BCH_S
// slots and compacts the occupied slots.
constexpr int BCH_E = 1024;      // member edges per insert item (4 per thread)
constexpr int BCH_S = 4096;      // table slots per extract item (16 per thread)
constexpr int BLDS = 2 * BCH_E;  // LDS pre-aggregation slots

__global__ void k_member_degs(const i32* members, const i64* rp, i64 n, i64* mdp) {
    const i64 k = thread_item();
    if (k >= n) return;
    i32 v = members[k];
    mdp[k] = rp[v + 1] - rp[v];
}
__global__ void k_items(const i32* list, i64 m, const i64* bound, const i64* goff, i64 nc, int slots, i64* items) {
    const i64 x = thread_item();
    if (x >= m) return;
    i64 len = slots ? goff[x + 1] - goff[x] : bound[list[x]];
    const i64 per = slots ? BCH_S : BCH_E;
    items[x] = (len + per - 1) / per;
}

template <class WT>
__global__ void __launch_bounds__(256) k_cbig_insert(const i64* rp, const i32* col, WT W, const i32* vcmap,
                                                      const i32* members, const i64* mstart, const i64* mdp,
                                                      const i32* big, i64 nbig, const i64* items, const i64* goff,
                                                      i32* gkeys, i64* gvals, u64* gtab) {
    // big, items and goff point at a batch of rows: items[0] and goff[0] are
    // the batch's first work item and first table slot
    const i64 b = block_id() + items[0];
    if (b >= items[nbig]) return;
    const i64 ui = upper_row(items, 0, nbig, b);
    const i32 u = big[ui];
    const i64 m0 = mstart[u], m1 = mstart[u + 1];
    const i64 e_lo = mdp[m0] + (b - items[ui]) * BCH_E;
    const i64 e_end = mdp[m1];
    const i64 e_hi = e_lo + BCH_E < e_end ? e_lo + BCH_E : e_end;
    __shared__ i32 lk[BLDS];
    __shared__ i64 lv[BLDS];
    lds_agg_clear<256, BLDS>(lk, lv);
    constexpr int EPI = BCH_E / 256;
    const i64 e0 = e_lo + (i64)threadIdx.x * EPI;
    if (e0 < e_hi) {
        i64 k = upper_row(mdp, m0, m1, e0);
        i64 nxt = mdp[k + 1];
        for (int q = 0; q < EPI; q++) {
            const i64 e = e0 + q;
            if (e >= e_hi) break;
            if (e >= nxt) k = advance_row(mdp, k, m1, e, nxt);
            const i32 v = members[k];
            const i64 j = rp[v] + (e - mdp[k]);
            const i32 cv = vcmap[col[j]];
            if (cv != u) table_add_lds(lk, lv, BLDS - 1, cv, W(j));
        }
    }
    const u32 gmask = (u32)(goff[ui + 1] - goff[ui] - 1);
    const i64 t0 = goff[ui] - goff[0];
    if (gtab) lds_agg_flush_packed<256, BLDS>(lk, lv, gtab + t0, gmask);  // kernel-uniform
    else lds_agg_flush<256, BLDS>(lk, lv, gkeys + t0, gvals + t0, gmask);
}

// gtab non-null: packed tables (key << 32 | weight), else gkeys / gvals
__global__ void __launch_bounds__(256) k_cbig_extract(const i32* big, i64 nbig, const i64* items, const i64* goff,
                                                       const i32* gkeys, const i64* gvals, const u64* gtab,
                                                       const i64* toff, i32* tcol, i64* tw, i64* ccnt) {
    __shared__ int s_cnt;
    __shared__ i64 s_base;
    // a batch of rows, as in k_cbig_insert; slot sl lives at sl - g0
    const i64 b = block_id() + items[0];
    if (b >= items[nbig]) return;
    const i64 g0 = goff[0];
    const i64 ui = upper_row(items, 0, nbig, b);
    const i32 u = big[ui];
    const i64 s_lo = goff[ui] + (b - items[ui]) * BCH_S;
    const i64 s_end = goff[ui + 1];
    const i64 s_hi = s_lo + BCH_S < s_end ? s_lo + BCH_S : s_end;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    i32 kk[16];
    int pos[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const i64 sl = s_lo + threadIdx.x + (i64)q * 256;
        if (gtab) {
            const u64 e = sl < s_hi ? gtab[sl - g0] : PEMPTY;
            kk[q] = e == PEMPTY ? EMPTY_KEY : (i32)(e >> 32);
        } else {
            kk[q] = sl < s_hi ? gkeys[sl - g0] : EMPTY_KEY;
        }
        pos[q] = kk[q] != EMPTY_KEY ? atomicAdd(&s_cnt, 1) : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base = (i64)atomicAdd((unsigned long long*)&ccnt[u], (unsigned long long)s_cnt);
    __syncthreads();
    const i64 base = toff[u] + s_base;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        if (pos[q] >= 0) {
            const i64 sl = s_lo + threadIdx.x + (i64)q * 256;
            tcol[base + pos[q]] = kk[q];
            tw[base + pos[q]] = gtab ? (i64)(gtab[sl - g0] & 0xffffffffull) : gvals[sl - g0];
        }
    }
}

// Copy temporary rows into the compact coarse CSR.  Consecutive lanes copy
// consecutive output entries (EPT steps of EB entries per block), so each
// wave-wide load and store covers whole cache lines.  The row of an entry is
// a binary search over the chunk's row starts, staged in LDS as offsets from
// the chunk start clamped to [0, EC]; a chunk spanning more than EC rows (runs
// of empty rows) searches the row pointers in global memory instead.  The
// search interval carries over between steps since entries only move right.
__global__ void __launch_bounds__(EB) k_crow_copy(const i64* crp, i64 nc, i64 nnzc, const i64* toff, const i32* tcol,
                                                   const i64* tw, i32* ocol, i64* ow) {
    __shared__ i32 s_st[EC + 1];
    ChunkRows cr;
    chunk_rows(crp, nc, nnzc, cr);
    const i64 nr = cr.rhi - cr.rlo;
    const bool staged = nr <= EC;
    if (staged) {
        for (int i = threadIdx.x; i <= (int)nr; i += EB) {
            const i64 v = crp[cr.rlo + i] - cr.eb;
            s_st[i] = (i32)(v < 0 ? 0 : (v > EC ? EC : v));
        }
    }
    __syncthreads();
    int lo = 0;
    i64 r = cr.rlo;
    for (int k = 0; k < EPT; k++) {
        const i64 j = cr.eb + (i64)k * EB + threadIdx.x;
        if (j >= cr.ee) break;
        if (staged) {
            // s_st[lo] <= o < s_st[hi]: s_st[0] = 0 and s_st[nr] >= ee - eb
            const int o = (int)(j - cr.eb);
            int hi = (int)nr;
            while (hi - lo > 1) {
                const int m = (lo + hi) >> 1;
                if (s_st[m] <= o) lo = m;
                else hi = m;
            }
            r = cr.rlo + lo;
        } else {
            r = upper_row(crp, r, cr.rhi, j);
        }
        const i64 src = toff[r] + (j - crp[r]);
        ocol[j] = tcol[src];
        ow[j] = tw[src];
    }
}

// Sorting of rows merged in global tables: flatten, radix sort by
// (row rank, column), scatter back.
__global__ void k_big_count(const i32* list, i64 m, const i64* crp, i64* cnt) {
    const i64 x = thread_item();
    if (x >= m) return;
    i32 u = list[x];
    cnt[x] = crp[u + 1] - crp[u];
}
// key = (big row index in the batch << cb) | column (columns < 2^cb), so the
// radix sort needs only cb + bits(rows in the batch) bits
__global__ void k_big_flatten(const i32* list, i64 m, const i64* crp, const i64* boff, const i32* ocol,
                              const i64* ow, int cb, u64* key, i64* val, i64* target) {
    const i64 x = block_id();
    if (x >= m) return;
    {
        i32 u = list[x];
        i64 s = crp[u], e = crp[u + 1], b = boff[x] - boff[0];  // within the batch
        for (i64 j = s + threadIdx.x; j < e; j += blockDim.x) {
            i64 f = b + (j - s);
            key[f] = ((u64)x << cb) | (u32)ocol[j];
            val[f] = ow[j];  // the weight travels with its key through the sort
            target[f] = j;
        }
    }
}
// the f-th smallest key of the batch goes to the f-th position of its rows
// in row order (rows keep their order and their lengths through the sort)
__global__ void k_big_scatter(const u64* key, const i64* target, i64 E, const i64* wsorted, int cb, i32* ocol,
                              i64* ow) {
    const u64 mask = ((u64)1 << cb) - 1;
    for (i64 f = (i64)blockIdx.x * blockDim.x + threadIdx.x; f < E; f += (i64)gridDim.x * blockDim.x) {
        i64 j = target[f];
        ocol[j] = (i32)(key[f] & mask);
        ow[j] = wsorted[f];
    }
}

// Algorithmic bytes of the coarse-row merge (timing runs only, counted after
// it ran): per coarse row its list entry, member range and output count; per
// member its id and row bounds; per member edge the column, weight and the
// neighbour's coarse id; per distinct coarse neighbour the column and int64
// weight written.
__global__ void __launch_bounds__(256) k_contract_bytes(const i32* __restrict__ list, i64 m,
                                                        const i64* __restrict__ mstart, const i64* __restrict__ bound,
                                                        const i64* __restrict__ ccnt, int wb, i64* acc) {
    const i64 x = thread_item();
    i64 b = 0;
    if (x < m) {
        const i32 u = list[x];
        b = 28 + 20 * (mstart[u + 1] - mstart[u]) + bound[u] * (8 + wb) + 12 * ccnt[u];
    }
    b = block_sum_i64<256>(b);
    if (threadIdx.x == 0 && b) atomicAdd((unsigned long long*)acc, (unsigned long long)b);
}
static void contract_bytes(Ctx& c, int tk, const i32* list, i64 m, const i64* mstart, const i64* bound,
                           const i64* ccnt, int wb) {
    if (tk < 0 || m <= 0) return;
    Buf<i64> acc(c, 1);
    fill_i64(c, acc, 1, 0);
    k_contract_bytes<<<grid_items(m, 256), 256, 0, c.stream>>>(list, m, mstart, bound, ccnt, wb, acc);
    PGGC_LAUNCH_CHECK();
    acct_add(c, c.kt_counter(tk), acc);
}


static void download_i64(Ctx& c, std::vector<i64>& dst, const i64* src, i64 count) {
    dst.resize((size_t)count);
    c.sync();
    PGGC_HIP(hipMemcpy(dst.data(), src, (size_t)count * sizeof(i64), hipMemcpyDeviceToHost));
}

static int bits_for(i64 v) {
    int b = 0;
    while (b < 62 && ((i64)1 << b) <= v) b++;
    return b;
}

__global__ void __launch_bounds__(256) k_max_i64(const i64* __restrict__ in, i64 n, u64* out) {
    u64 m = 0;
    // block-uniform loop control (the reduction below is cross-lane)
    for (i64 base = (i64)blockIdx.x * blockDim.x; base < n; base += (i64)gridDim.x * blockDim.x) {
        const i64 x = base + threadIdx.x;
        if (x < n) m = (u64)in[x] > m ? (u64)in[x] : m;  // weights are positive
    }
    m = block_max_u64<256>(m);
    if (threadIdx.x == 0 && m) atomicMax((unsigned long long*)out, (unsigned long long)m);
}
__global__ void k_narrow(const i64* __restrict__ in, i64 n, i32* __restrict__ out) {
    for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x)
        out[x] = (i32)in[x];
}

void narrow_weights(Ctx& c, const DGraph& fine, OwnedGraph& g) {
    if (g.nnz == 0 || !g.w.p) return;
    bool fits = fine.wk == W_UNIT && fine.nnz < ((i64)1 << 31);
    if (!fits) {
        Buf<u64> mx(c, 1);
        fill_i64(c, (i64*)mx.p, 1, 0);
        k_max_i64<<<grid_for(g.nnz, 256 * 8, 8192), 256, 0, c.stream>>>(g.w, g.nnz, mx);
        PGGC_LAUNCH_CHECK();
        fits = (u64)c.read1((const i64*)mx.p) <= 0x7fffffffull;
    }
    if (!fits) return;
    g.w32 = Buf<i32>(c, g.nnz);
    k_narrow<<<grid_for(g.nnz, 256 * 8, 8192), 256, 0, c.stream>>>(g.w, g.nnz, g.w32);
    PGGC_LAUNCH_CHECK();
    g.w.reset();
}

void contract(Ctx& c, const DGraph& g, const i32* vcmap, i64 nc, OwnedGraph& out) {
    const i64 n = g.n;
    Buf<i64> cnt(c, nc + 1), bound(c, nc + 1), mstart(c, nc + 1), toff(c, nc + 1), pos(c, nc + 1);
    Buf<i32> members(c, std::max<i64>(n, 1));
    fill_i64(c, cnt, nc + 1, 0);
    fill_i64(c, bound, nc + 1, 0);
    k_ccount<<<grid_items(n, 256), 256, 0, c.stream>>>(g.rp, vcmap, n, cnt, bound);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, cnt, mstart, nc);
    scan_i64(c, bound, toff, nc);
    copy_dev(c, pos, mstart, nc * sizeof(i64));
    k_cscatter<<<grid_items(n, 256), 256, 0, c.stream>>>(vcmap, n, pos, members);
    PGGC_LAUNCH_CHECK();
    const i64 tb = c.read1(toff.p + nc);
    c.mem_note("contract: before row staging");
    Buf<i32> tcol(c, std::max<i64>(tb, 1));
    Buf<i64> tw(c, std::max<i64>(tb, 1));
    Buf<i64> ccnt(c, nc + 1);
    Buf<i32> cls(c, std::max<i64>(nc, 1)), list(c, std::max<i64>(nc, 1));
    k_cclass<<<grid_items(nc, 256), 256, 0, c.stream>>>(bound, nc, cls);
    PGGC_LAUNCH_CHECK();
    i64 st[6];
    partition_k<5>(c, cls, nc, list, st);
    fill_i64(c, ccnt, nc + 1, 0);
    const i64 n0 = st[1] - st[0], n1 = st[2] - st[1], n1b = st[3] - st[2], n2 = st[4] - st[3];
    int tk = c.kt_begin(KS_CONTRACT_AGG);
    PGGC_WDISPATCH(g, {
        if (n0 > 0)
            k_cagg<64, 512, false><<<grid_blocks(n0), 64, 0, c.stream>>>(
                g.rp, g.col, W, vcmap, nc, list.p + st[0], n0, mstart, members, bound, nullptr, nullptr, nullptr,
                toff, tcol, tw, ccnt);
        if (n1 > 0)
            k_cagg<256, 2048, false><<<grid_blocks(n1), 256, 0, c.stream>>>(
                g.rp, g.col, W, vcmap, nc, list.p + st[1], n1, mstart, members, bound, nullptr, nullptr, nullptr,
                toff, tcol, tw, ccnt);
        if (n1b > 0)
            k_cagg<256, 4096, false><<<grid_blocks(n1b), 256, 0, c.stream>>>(
                g.rp, g.col, W, vcmap, nc, list.p + st[2], n1b, mstart, members, bound, nullptr, nullptr, nullptr,
                toff, tcol, tw, ccnt);
    });
    c.kt_end(tk, 0.0);
    PGGC_LAUNCH_CHECK();
    const int wb = g.wk == W_UNIT ? 0 : g.wk == W_I32 ? 4 : 8;
    contract_bytes(c, tk, list.p + st[0], n0 + n1 + n1b, mstart, bound, ccnt, wb);
    if (n2 > 0) {
        const i32* big = list.p + st[3];
        Buf<i64> goff(c, n2 + 1), items(c, n2 + 1), items2(c, n2 + 1), mdp(c, n + 1);
        k_goff_sizes<<<grid_items(n2, 256), 256, 0, c.stream>>>(big, n2, bound, nc, goff);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, goff, goff, n2);
        // packed tables (one CAS per new key) when every coarse weight fits
        // 32 bits: with unit weights a coarse weight counts fine edges
        const bool packed = g.wk == W_UNIT && g.nnz < ((i64)1 << 32) - 1;
        k_member_degs<<<grid_items(n, 256), 256, 0, c.stream>>>(members, g.rp, n, mdp);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, mdp, mdp, n);
        k_items<<<grid_items(n2, 256), 256, 0, c.stream>>>(big, n2, bound, goff, nc, 0, items);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, items, items, n2);
        k_items<<<grid_items(n2, 256), 256, 0, c.stream>>>(big, n2, bound, goff, nc, 1, items2);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, items2, items2, n2);
        // rows in batches whose tables hold at most c.contract_batch slots
        // together (a larger row is a batch of its own), so the tables'
        // memory stays bounded on graphs of billions of edges
        const i64 slot_cap = std::max<i64>(c.contract_batch, 1);
        std::vector<i64> hg, hi, hi2;
        download_i64(c, hg, goff, n2 + 1);
        download_i64(c, hi, items, n2 + 1);
        download_i64(c, hi2, items2, n2 + 1);
        i64 gmax = 0;
        const std::vector<i64> cuts = batch_cuts(hg, slot_cap, &gmax);
        Buf<i32> gkeys(c, packed ? 1 : gmax);
        Buf<i64> gvals(c, packed ? 1 : gmax);
        Buf<u64> gtab(c, packed ? gmax : 1);
        u64* gt = packed ? gtab.p : nullptr;
        c.mem_note("contract: big-row tables");
        tk = c.kt_begin(KS_CONTRACT_BIG);
        for (size_t q = 0; q + 1 < cuts.size(); q++) {
            const i64 r0 = cuts[q], r1 = cuts[q + 1], gsb = hg[r1] - hg[r0];
            if (packed) {
                fill_i64(c, (i64*)gtab.p, gsb, (i64)PEMPTY);
            } else {
                fill_i32(c, gkeys, gsb, EMPTY_KEY);
                fill_i64(c, gvals, gsb, 0);
            }
            const i64 I = hi[r1] - hi[r0];
            if (I > 0) {
                PGGC_WDISPATCH(g, (k_cbig_insert<<<grid_blocks(I), 256, 0, c.stream>>>(
                                      g.rp, g.col, W, vcmap, members, mstart, mdp, big + r0, r1 - r0, items.p + r0,
                                      goff.p + r0, gkeys, gvals, gt)));
                PGGC_LAUNCH_CHECK();
            }
            const i64 I2 = hi2[r1] - hi2[r0];
            if (I2 > 0) {
                k_cbig_extract<<<grid_blocks(I2), 256, 0, c.stream>>>(big + r0, r1 - r0, items2.p + r0, goff.p + r0,
                                                                      gkeys, gvals, gt, toff, tcol, tw, ccnt);
                PGGC_LAUNCH_CHECK();
            }
        }
        c.kt_end(tk, 0.0);
        contract_bytes(c, tk, big, n2, mstart, bound, ccnt, wb);
    }
    out.n = nc;
    out.rp = Buf<i64>(c, nc + 1);
    scan_i64(c, ccnt, out.rp, nc);
    out.nnz = c.read1(out.rp.p + nc);
    out.col = Buf<i32>(c, std::max<i64>(out.nnz, 1));
    out.w = Buf<i64>(c, std::max<i64>(out.nnz, 1));
    if (out.nnz > 0) {
        // algorithmic bytes: coarse offsets, row starts in the temporary rows,
        // (column, weight) read and written once
        const double bytes = 16.0 * (nc + 1) + 24.0 * out.nnz;
        const int tkc = c.kt_begin(KS_CONTRACT_COPY);
        k_crow_copy<<<(int)ceil_div(out.nnz, EC), EB, 0, c.stream>>>(out.rp, nc, out.nnz, toff, tcol, tw, out.col,
                                                                   out.w);
        c.kt_end(tkc, bytes);
        PGGC_LAUNCH_CHECK();
    }
    tcol.reset();  // the staged rows are copied out
    tw.reset();
    if (n2 > 0) {
        // the rows merged in HBM tables come out unsorted: sort each by
        // column, in batches of at most c.contract_batch / 4 entries (a
        // larger row is a batch of its own)
        const i64 sort_cap = std::max<i64>(c.contract_batch / 4, 1);
        const i32* big = list.p + st[3];
        Buf<i64> boff(c, n2 + 1);
        k_big_count<<<grid_items(n2, 256), 256, 0, c.stream>>>(big, n2, out.rp, boff);
        PGGC_LAUNCH_CHECK();
        scan_i64(c, boff, boff, n2);
        std::vector<i64> hb;
        download_i64(c, hb, boff, n2 + 1);
        i64 emax = 0;
        const std::vector<i64> cuts = batch_cuts(hb, sort_cap, &emax);
        if (emax > 0) {
            Buf<u64> key(c, emax);
            Buf<i64> val(c, emax), target(c, emax);
            const int cb = bits_for(nc - 1);
            for (size_t q = 0; q + 1 < cuts.size(); q++) {
                const i64 r0 = cuts[q], r1 = cuts[q + 1], E = hb[r1] - hb[r0];
                if (E == 0) continue;
                k_big_flatten<<<grid_blocks(r1 - r0), 256, 0, c.stream>>>(big + r0, r1 - r0, out.rp, boff.p + r0,
                                                                          out.col, out.w, cb, key, val, target);
                PGGC_LAUNCH_CHECK();
                radix_sort_pairs(c, key, val, E, cb + bits_for(r1 - r0));
                k_big_scatter<<<grid_for(E, 256 * 4, 8192), 256, 0, c.stream>>>(key, target, E, val, cb, out.col,
                                                                                out.w);
                PGGC_LAUNCH_CHECK();
            }
        }
    }
    (void)n1;
}

}  // namespace pggc
