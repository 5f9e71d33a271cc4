// Active-set predictor: feature, forward, backward and Adam kernels.
//
// Every input feature is an exact integer (or an exact ratio of integers)
// derived from a per-vertex pack of quantised values, so the input rows are
// reproducible bit for bit; only the fp32 accumulation order of the MFMA and
// of the neighbour means of layer 2 is implementation-defined.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "device.h"
#include "gnn.h"
#include "localmove.h"

namespace pggc {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int GNN_GRAD_BLOCKS = 512;

// fp32 -> bf16 with round to nearest even (values are finite)
__host__ __device__ __forceinline__ u16 bf16_rne(float f) {
    u32 u = __builtin_bit_cast(u32, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_f32(u16 b) { return __builtin_bit_cast(float, (u32)b << 16); }

// 8*floor(log2 v) + the three bits below the leading one (v >= 1)
__device__ __forceinline__ u32 logq(u64 v) {
    const int e = 63 - __clzll((long long)v);
    const u32 f = e >= 3 ? (u32)(v >> (e - 3)) & 7u : (u32)(v << (3 - e)) & 7u;
    return 8u * (u32)e + f;
}

// last gain of the vertex per (degree+1), in 1/64 steps, clamped to
// [-127, 127]; -128 for a cached NO_MOVE (gain -FLT_MAX); 0 before the
// first sweep of a call has scored the vertex
__device__ __forceinline__ int margin_q(float gain, i64 deg, bool first) {
    if (first) return 0;
    if (!(gain > -FLT_MAX)) return -128;
    float m = rintf(gain / (float)(deg + 1) * 64.0f);
    m = fminf(fmaxf(m, -127.0f), 127.0f);
    return (int)m;
}

__device__ __forceinline__ int pk_margin(u64 q) { return (int)(signed char)(unsigned char)((q >> 40) & 255u); }

template <class T, int BS>
__device__ __forceinline__ T block_sum_lds(T v) {
    __shared__ T sh[BS];
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] = sh[t] + sh[t + o];
        __syncthreads();
    }
    const T r = sh[0];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------ pack
// pk = part | ldeg_q << 32 | margin_q << 40 | moved << 48 | active << 49;
// blk[block] = active vertices of the block
__global__ void __launch_bounds__(256) k_gnn_pack(i64 n, const i64* __restrict__ rp, const i32* __restrict__ part,
                                                  const i32* __restrict__ dest, const float* __restrict__ gain,
                                                  const i32* __restrict__ stamp, i32 cur, int first,
                                                  u64* __restrict__ pk, i64* __restrict__ blk) {
    const i64 i = thread_item();
    i64 a = 0;
    if (i < n) {
        const i64 deg = rp[i + 1] - rp[i];
        const u32 lq = min(255u, logq((u64)deg + 1));
        const int mq = margin_q(gain[i], deg, first != 0);
        const u64 moved = stamp[i] == cur - 1 ? 1 : 0;
        const u64 act = dest[i] != NO_MOVE ? 1 : 0;
        pk[i] = (u64)(u32)part[i] | ((u64)lq << 32) | ((u64)(unsigned char)(signed char)mq << 40) | (moved << 48) |
                (act << 49);
        a = (i64)act;
    }
    const i64 s = block_sum_i64<256>(a);
    if (threadIdx.x == 0) blk[block_id()] = s;
}

// blk[block] = vertices with flag[i] == v (v != 0) or, for v == 0, flag[i] != 0
__global__ void __launch_bounds__(256) k_gnn_count(i64 n, const i32* __restrict__ flag, i32 v, i64* __restrict__ blk) {
    const i64 i = thread_item();
    const i64 a = (i < n && (v ? flag[i] == v : flag[i] != 0)) ? 1 : 0;
    const i64 s = block_sum_i64<256>(a);
    if (threadIdx.x == 0) blk[block_id()] = s;
}

// cnt[to] += cnt[from]
__global__ void k_gnn_acc(i64* cnt, int to, int from) {
    if (threadIdx.x == 0 && blockIdx.x == 0) cnt[to] += cnt[from];
}

// ------------------------------------------------------------- features
// Neighbour sums: 0 ldeg_q, 1 margin_q, 2 moved, 3 active, 4 same cluster,
// 5 margin above own, 6 same & active, 7 same & moved, 8 weight, 9 own
// weight.  Rows of degree <= 2048 carry them packed in three words
// (p0 = S0 | (S1 + 128 per edge) << 32, p1 = S2 | S3 << 21 | S4 << 42,
// p2 = S5 | S6 << 21 | S7 << 42; no field can carry into the next), and the
// weights only when the graph has them (unit weights: S8 = degree, S9 = S4).
constexpr int NAGG = 10;

struct AggPacked {
    u64 p0 = 0, p1 = 0, p2 = 0;
    i64 wa = 0, wo = 0;
};

template <class WT>
__device__ __forceinline__ void agg_edge_packed(u64 q, i64 w, i32 p, int mq_i, AggPacked& a) {
    const u64 same = (i32)(u32)q == p ? 1 : 0;
    const int mq = pk_margin(q);
    const u64 mv = (q >> 48) & 1u, ac = (q >> 49) & 1u;
    a.p0 += ((q >> 32) & 255u) | ((u64)(mq + 128) << 32);
    a.p1 += mv | (ac << 21) | (same << 42);
    a.p2 += (u64)(mq > mq_i ? 1 : 0) | ((same & ac) << 21) | ((same & mv) << 42);
    if (WT::bytes > 0) {
        a.wa += w;
        a.wo += same ? w : 0;
    }
}

template <class WT>
__device__ __forceinline__ void agg_unpack(const AggPacked& a, i64 d, i64* S) {
    constexpr u64 M21 = (1ull << 21) - 1;
    S[0] = (i64)(a.p0 & 0xFFFFFFFFull);
    S[1] = (i64)(a.p0 >> 32) - 128 * d;
    S[2] = (i64)(a.p1 & M21);
    S[3] = (i64)((a.p1 >> 21) & M21);
    S[4] = (i64)(a.p1 >> 42);
    S[5] = (i64)(a.p2 & M21);
    S[6] = (i64)((a.p2 >> 21) & M21);
    S[7] = (i64)(a.p2 >> 42);
    S[8] = WT::bytes > 0 ? a.wa : d;
    S[9] = WT::bytes > 0 ? a.wo : S[4];
}

__device__ __forceinline__ void agg_edge(u64 q, i64 w, i32 p, int mq_i, i64* a) {
    const i64 same = (i32)(u32)q == p ? 1 : 0;
    const int mq = pk_margin(q);
    const i64 mv = (i64)((q >> 48) & 1u), ac = (i64)((q >> 49) & 1u);
    a[0] += (i64)((q >> 32) & 255u);
    a[1] += mq;
    a[2] += mv;
    a[3] += ac;
    a[4] += same;
    a[5] += mq > mq_i ? 1 : 0;
    a[6] += same & ac;
    a[7] += same & mv;
    a[8] += w;
    a[9] += same ? w : 0;
}

__device__ void gnn_write_row(u64 me, i64 vwi, i64 volp, i64 deg, const i64* a, u16* out) {
    float x[GNN_IN];
    x[0] = (float)((me >> 32) & 255u) / 256.0f;
    x[1] = (float)logq(vwi > 0 ? (u64)vwi + 1 : 1) / 512.0f;
    x[2] = (float)pk_margin(me) / 128.0f;
    x[3] = (float)((me >> 48) & 1u);
    x[4] = (float)((me >> 49) & 1u);
    x[5] = a[8] > 0 ? (float)a[9] / (float)a[8] : 0.0f;
    x[6] = (float)logq(volp > 0 ? (u64)volp + 1 : 1) / 512.0f;
    x[7] = 1.0f;
    const float D = (float)(deg > 0 ? deg : 1);
    x[8] = (float)a[0] / (D * 256.0f);
    x[9] = (float)a[1] / (D * 128.0f);
#pragma unroll
    for (int k = 2; k < 8; k++) x[8 + k] = (float)a[k] / D;
    u32 w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = (u32)bf16_rne(x[2 * k]) | ((u32)bf16_rne(x[2 * k + 1]) << 16);
    uint4* o = (uint4*)out;
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// vertices of degree <= G, G lanes each (one edge per lane)
template <int G, class WT>
__global__ void __launch_bounds__(256) k_gnn_agg_group(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                       const i64* __restrict__ vw, const i32* __restrict__ list, i64 m,
                                                       const u64* __restrict__ pk, const i64* __restrict__ vol,
                                                       u16* __restrict__ in) {
    constexpr int GPB = 256 / G;
    const int gl = threadIdx.x & (G - 1);
    const i64 gi = block_id() * GPB + threadIdx.x / G;
    if (gi >= m) return;
    const i32 i = list[gi];
    const i64 s0 = rp[i];
    const i64 d = rp[i + 1] - s0;
    const u64 me = pk[i];
    const i32 p = (i32)(u32)me;
    AggPacked a;
    if (gl < d) agg_edge_packed<WT>(pk[col[s0 + gl]], W(s0 + gl), p, pk_margin(me), a);
    a.p0 = (u64)group_sum_i64<G>((i64)a.p0);
    a.p1 = (u64)group_sum_i64<G>((i64)a.p1);
    a.p2 = (u64)group_sum_i64<G>((i64)a.p2);
    if (WT::bytes > 0) {
        a.wa = group_sum_i64<G>(a.wa);
        a.wo = group_sum_i64<G>(a.wo);
    }
    if (gl == 0) {
        i64 S[NAGG];
        agg_unpack<WT>(a, d, S);
        gnn_write_row(me, vw[i], vol[p], d, S, in + (i64)i * GNN_IN);
    }
}

// one wave per vertex of degree 65..2048 (wave-uniform loop)
template <class WT>
__global__ void __launch_bounds__(256) k_gnn_agg_wave(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                      const i64* __restrict__ vw, const i32* __restrict__ list, i64 m,
                                                      const u64* __restrict__ pk, const i64* __restrict__ vol,
                                                      u16* __restrict__ in) {
    const int lane = threadIdx.x & 63;
    const i64 gi = block_id() * 4 + (threadIdx.x >> 6);
    if (gi >= m) return;
    const i32 i = list[gi];
    const i64 s0 = rp[i], e = rp[i + 1];
    const u64 me = pk[i];
    const i32 p = (i32)(u32)me;
    const int mqi = pk_margin(me);
    AggPacked a;
    for (i64 base = s0; base < e; base += 64) {
        const i64 j = base + lane;
        if (j < e) agg_edge_packed<WT>(pk[col[j]], W(j), p, mqi, a);
    }
    a.p0 = (u64)wave_sum_i64((i64)a.p0);
    a.p1 = (u64)wave_sum_i64((i64)a.p1);
    a.p2 = (u64)wave_sum_i64((i64)a.p2);
    if (WT::bytes > 0) {
        a.wa = wave_sum_i64(a.wa);
        a.wo = wave_sum_i64(a.wo);
    }
    if (lane == 0) {
        i64 S[NAGG];
        agg_unpack<WT>(a, e - s0, S);
        gnn_write_row(me, vw[i], vol[p], e - s0, S, in + (i64)i * GNN_IN);
    }
}

// one block per vertex of degree > 2048 (block-uniform loop, LDS tree)
template <class WT>
__global__ void __launch_bounds__(256) k_gnn_agg_block(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                       const i64* __restrict__ vw, const i32* __restrict__ list, i64 m,
                                                       const u64* __restrict__ pk, const i64* __restrict__ vol,
                                                       u16* __restrict__ in) {
    __shared__ i64 sh[NAGG][256];
    const i64 gi = block_id();
    if (gi >= m) return;
    const int t = threadIdx.x;
    const i32 i = list[gi];
    const i64 s0 = rp[i], e = rp[i + 1];
    const u64 me = pk[i];
    const i32 p = (i32)(u32)me;
    const int mqi = pk_margin(me);
    i64 a[NAGG];
#pragma unroll
    for (int k = 0; k < NAGG; k++) a[k] = 0;
    for (i64 base = s0; base < e; base += 256) {
        const i64 j = base + t;
        if (j < e) agg_edge(pk[col[j]], W(j), p, mqi, a);
    }
#pragma unroll
    for (int k = 0; k < NAGG; k++) sh[k][t] = a[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
#pragma unroll
            for (int k = 0; k < NAGG; k++) sh[k][t] += sh[k][t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < NAGG; k++) a[k] = sh[k][0];
        gnn_write_row(me, vw[i], vol[p], e - s0, a, in + (i64)i * GNN_IN);
    }
}

// --------------------------------------------------------------- layer 1
// One wave per 16 vertices: Z = IN[16x16] * W1[16x32] as two
// v_mfma_f32_16x16x16_bf16 (A: lane holds row lane%16, k = 4*(lane/16)+0..3;
// B: column lane%16, same k; D: rows 4*(lane/16)+0..3, column lane%16).
// Then h = bf16(ReLU(z + b1)), s = h.w2n and t = h.w2s + b2.
__global__ void __launch_bounds__(256) k_gnn_dense(const u16* __restrict__ in, i64 n, const float* __restrict__ prm,
                                                   const u16* __restrict__ w1b, u16* __restrict__ hout,
                                                   float* __restrict__ s, float* __restrict__ t) {
    const int lane = threadIdx.x & 63;
    const i64 tile = block_id() * 4 + (threadIdx.x >> 6);
    if (tile * 16 >= n) return;  // wave-uniform
    const int r = lane & 15, kq = lane >> 4;
    const i64 arow = tile * 16 + r;
    v4s a = {0, 0, 0, 0};
    if (arow < n) {
        const uint2 raw = *(const uint2*)(in + arow * GNN_IN + 4 * kq);
        a[0] = (short)(raw.x & 0xFFFFu);
        a[1] = (short)(raw.x >> 16);
        a[2] = (short)(raw.y & 0xFFFFu);
        a[3] = (short)(raw.y >> 16);
    }
    v4s b0, b1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        b0[j] = (short)w1b[(4 * kq + j) * GNN_HID + r];
        b1[j] = (short)w1b[(4 * kq + j) * GNN_HID + 16 + r];
    }
    v4f z0 = {0.0f, 0.0f, 0.0f, 0.0f}, z1 = {0.0f, 0.0f, 0.0f, 0.0f};
    z0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, z0, 0, 0, 0);
    z1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, z1, 0, 0, 0);
    const float c0 = prm[GP_B1 + r], c1 = prm[GP_B1 + 16 + r];
    const float ws0 = prm[GP_W2S + r], ws1 = prm[GP_W2S + 16 + r];
    const float wn0 = prm[GP_W2N + r], wn1 = prm[GP_W2N + 16 + r];
    float ps[4], pt[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const u16 hb0 = bf16_rne(fmaxf(z0[j] + c0, 0.0f));
        const u16 hb1 = bf16_rne(fmaxf(z1[j] + c1, 0.0f));
        const i64 row = tile * 16 + 4 * kq + j;
        if (hout && row < n) {
            hout[row * GNN_HID + r] = hb0;
            hout[row * GNN_HID + 16 + r] = hb1;
        }
        const float h0 = bf16_f32(hb0), h1 = bf16_f32(hb1);
        ps[j] = h0 * wn0 + h1 * wn1;
        pt[j] = h0 * ws0 + h1 * ws1;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        xlane_fence(ps[j]);
        xlane_fence(pt[j]);
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            ps[j] = ps[j] + __shfl_xor(ps[j], o, 16);
            pt[j] = pt[j] + __shfl_xor(pt[j], o, 16);
        }
    }
    if (r == 0) {
        const float b2 = prm[GP_B2];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const i64 row = tile * 16 + 4 * kq + j;
            if (row < n) {
                s[row] = ps[j];
                t[row] = pt[j] + b2;
            }
        }
    }
}

// --------------------------------------------------------------- layer 2
// logit_i = t_i + mean_{j in N(i)} s_j for active vertices; with `prune`,
// those below tau get a cached NO_MOVE and pruned[i] = cur.
__device__ __forceinline__ void logit_out(i32 i, float L, float tau, int prune, i32 cur, float* logit, i32* dest,
                                          i32* pruned) {
    if (logit) logit[i] = L;
    if (prune && L < tau) {
        dest[i] = NO_MOVE;
        pruned[i] = cur;
    }
}

__device__ __forceinline__ float wave_sum_f32(float v) {
    xlane_fence(v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    return v;
}

template <int G>
__global__ void __launch_bounds__(256) k_gnn_logit_group(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                         const i32* __restrict__ list, i64 m, i32* dest,
                                                         const float* __restrict__ s, const float* __restrict__ t,
                                                         float tau, int prune, i32 cur, float* logit, i32* pruned) {
    constexpr int GPB = 256 / G;
    const int gl = threadIdx.x & (G - 1);
    const i64 gi = block_id() * GPB + threadIdx.x / G;
    if (gi >= m) return;
    const i32 i = list[gi];
    if (dest[i] == NO_MOVE) return;
    const i64 s0 = rp[i], d = rp[i + 1] - s0;
    float v = gl < d ? s[col[s0 + gl]] : 0.0f;
    xlane_fence(v);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = v + __shfl_xor(v, o, G);
    if (gl == 0) logit_out(i, t[i] + v / (float)(d > 0 ? d : 1), tau, prune, cur, logit, dest, pruned);
}

__global__ void __launch_bounds__(256) k_gnn_logit_wave(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                        const i32* __restrict__ list, i64 m, i32* dest,
                                                        const float* __restrict__ s, const float* __restrict__ t,
                                                        float tau, int prune, i32 cur, float* logit, i32* pruned) {
    const int lane = threadIdx.x & 63;
    const i64 gi = block_id() * 4 + (threadIdx.x >> 6);
    if (gi >= m) return;
    const i32 i = list[gi];
    if (dest[i] == NO_MOVE) return;  // wave-uniform
    const i64 s0 = rp[i], e = rp[i + 1];
    float v = 0.0f;
    for (i64 base = s0; base < e; base += 64) {
        const i64 j = base + lane;
        if (j < e) v = v + s[col[j]];
    }
    v = wave_sum_f32(v);
    if (lane == 0) logit_out(i, t[i] + v / (float)(e > s0 ? e - s0 : 1), tau, prune, cur, logit, dest, pruned);
}

template <class T>
__device__ __forceinline__ T block_tree_256(T v, T* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) sh[t] = sh[t] + sh[t + o];
        __syncthreads();
    }
    return sh[0];
}

__global__ void __launch_bounds__(256) k_gnn_logit_block(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                         const i32* __restrict__ list, i64 m, i32* dest,
                                                         const float* __restrict__ s, const float* __restrict__ t,
                                                         float tau, int prune, i32 cur, float* logit, i32* pruned) {
    __shared__ float sh[256];
    const i64 gi = block_id();
    if (gi >= m) return;
    const i32 i = list[gi];
    if (dest[i] == NO_MOVE) return;  // block-uniform
    const i64 s0 = rp[i], e = rp[i + 1];
    float v = 0.0f;
    for (i64 base = s0; base < e; base += 256) {
        const i64 j = base + threadIdx.x;
        if (j < e) v = v + s[col[j]];
    }
    v = block_tree_256(v, sh);
    if (threadIdx.x == 0) logit_out(i, t[i] + v / (float)(e > s0 ? e - s0 : 1), tau, prune, cur, logit, dest, pruned);
}

// ------------------------------------------------------------- backward
// g_i = dLoss/dlogit_i (mean weighted BCE over the active vertices),
// r_i = g_i / deg_i; confusion counts and the summed loss for the record.
__global__ void __launch_bounds__(256) k_gnn_loss(i64 n, const i64* __restrict__ rp, const u64* __restrict__ pk,
                                                  const float* __restrict__ logit, const i32* __restrict__ labels,
                                                  const i32* __restrict__ stamp, i32 cur, float pw, float tau,
                                                  i64* cnt, float* __restrict__ g, float* __restrict__ r,
                                                  double* loss) {
    const i64 i = thread_item();
    i64 c4[4] = {0, 0, 0, 0};
    double l = 0.0;
    if (i < n) {
        float gi = 0.0f;
        if ((pk[i] >> 49) & 1u) {
            const float y = (labels ? labels[i] != 0 : stamp[i] == cur) ? 1.0f : 0.0f;
            const float L = logit[i];
            const i64 na = cnt[GC_SWEEP_ACTIVE];
            const float inv = 1.0f / (float)(na > 0 ? na : 1);
            const float sg = 1.0f / (1.0f + expf(-L));
            gi = (sg * (pw * y + 1.0f - y) - pw * y) * inv;
            const float lp = log1pf(expf(-fabsf(L)));
            const float sp_neg = fmaxf(-L, 0.0f) + lp;  // -log sigmoid(L)
            const float sp_pos = fmaxf(L, 0.0f) + lp;   // -log (1 - sigmoid(L))
            l = (double)((pw * y * sp_neg + (1.0f - y) * sp_pos) * inv);
            const bool pred = L >= tau;
            c4[pred ? (y > 0.0f ? 0 : 1) : (y > 0.0f ? 2 : 3)] = 1;
        }
        const i64 d = rp[i + 1] - rp[i];
        g[i] = gi;
        r[i] = gi / (float)(d > 0 ? d : 1);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const i64 s = block_sum_i64<256>(c4[k]);
        if (threadIdx.x == 0 && s) atomicAdd((unsigned long long*)(cnt + GC_TP + k), (unsigned long long)s);
    }
    const double ls = block_sum_lds<double, 256>(l);
    if (threadIdx.x == 0 && ls != 0.0) atomicAdd(loss, ls);
}

// ds_j = sum_{i in N(j)} r_i  (the graph is symmetric)
template <int G>
__global__ void __launch_bounds__(256) k_gnn_ds_group(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                      const i32* __restrict__ list, i64 m, const float* __restrict__ r,
                                                      float* __restrict__ ds) {
    constexpr int GPB = 256 / G;
    const int gl = threadIdx.x & (G - 1);
    const i64 gi = block_id() * GPB + threadIdx.x / G;
    if (gi >= m) return;
    const i32 j = list[gi];
    const i64 s0 = rp[j], d = rp[j + 1] - s0;
    float v = gl < d ? r[col[s0 + gl]] : 0.0f;
    xlane_fence(v);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = v + __shfl_xor(v, o, G);
    if (gl == 0) ds[j] = v;
}

__global__ void __launch_bounds__(256) k_gnn_ds_wave(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                     const i32* __restrict__ list, i64 m, const float* __restrict__ r,
                                                     float* __restrict__ ds) {
    const int lane = threadIdx.x & 63;
    const i64 gi = block_id() * 4 + (threadIdx.x >> 6);
    if (gi >= m) return;
    const i32 j = list[gi];
    const i64 s0 = rp[j], e = rp[j + 1];
    float v = 0.0f;
    for (i64 base = s0; base < e; base += 64) {
        const i64 x = base + lane;
        if (x < e) v = v + r[col[x]];
    }
    v = wave_sum_f32(v);
    if (lane == 0) ds[j] = v;
}

__global__ void __launch_bounds__(256) k_gnn_ds_block(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                      const i32* __restrict__ list, i64 m, const float* __restrict__ r,
                                                      float* __restrict__ ds) {
    __shared__ float sh[256];
    const i64 gi = block_id();
    if (gi >= m) return;
    const i32 j = list[gi];
    const i64 s0 = rp[j], e = rp[j + 1];
    float v = 0.0f;
    for (i64 base = s0; base < e; base += 256) {
        const i64 x = base + threadIdx.x;
        if (x < e) v = v + r[col[x]];
    }
    v = block_tree_256(v, sh);
    if (threadIdx.x == 0) ds[j] = v;
}

// Weight gradients.  Each wave walks 16-vertex tiles (wave-uniform loop);
// dW1 = IN^T dZ accumulates on MFMA (A: features x vertices, B: vertices x
// outputs, dZ rounded to bf16), the bias and layer-2 gradients in fp32 lanes.
// Per-block partials go to part[block][GNN_P].
__global__ void __launch_bounds__(256) k_gnn_grad(i64 n, const u16* __restrict__ in, const u16* __restrict__ h,
                                                  const float* __restrict__ g, const float* __restrict__ ds,
                                                  const float* __restrict__ prm, float* __restrict__ part) {
    __shared__ float red[4][GNN_P];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, kq = lane >> 4;
    const i64 tw = (i64)gridDim.x * 4;
    const i64 ntiles = (n + 15) / 16;
    const float ws0 = prm[GP_W2S + r], ws1 = prm[GP_W2S + 16 + r];
    const float wn0 = prm[GP_W2N + r], wn1 = prm[GP_W2N + 16 + r];
    v4f acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
    float pb0 = 0.0f, pb1 = 0.0f, ps0 = 0.0f, ps1 = 0.0f, pn0 = 0.0f, pn1 = 0.0f, pb2 = 0.0f;
    for (i64 tile = (i64)blockIdx.x * 4 + wv; tile < ntiles; tile += tw) {
        v4s a, b0, b1;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const i64 v = tile * 16 + 4 * kq + j;
            const bool ok = v < n;
            a[j] = ok ? (short)in[v * GNN_IN + r] : (short)0;
            const float gv = ok ? g[v] : 0.0f, dv = ok ? ds[v] : 0.0f;
            const float h0 = ok ? bf16_f32(h[v * GNN_HID + r]) : 0.0f;
            const float h1 = ok ? bf16_f32(h[v * GNN_HID + 16 + r]) : 0.0f;
            const float dz0 = h0 > 0.0f ? gv * ws0 + dv * wn0 : 0.0f;
            const float dz1 = h1 > 0.0f ? gv * ws1 + dv * wn1 : 0.0f;
            b0[j] = (short)bf16_rne(dz0);
            b1[j] = (short)bf16_rne(dz1);
            pb0 += dz0;
            pb1 += dz1;
            ps0 += gv * h0;
            ps1 += gv * h1;
            pn0 += dv * h0;
            pn1 += dv * h1;
            if (r == 0) pb2 += gv;
        }
        acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        red[wv][GP_W1 + (4 * kq + j) * GNN_HID + r] = acc0[j];
        red[wv][GP_W1 + (4 * kq + j) * GNN_HID + 16 + r] = acc1[j];
    }
    float v[7] = {pb0, pb1, ps0, ps1, pn0, pn1, pb2};
#pragma unroll
    for (int k = 0; k < 7; k++) {
        xlane_fence(v[k]);
        v[k] = v[k] + __shfl_xor(v[k], 16);
        v[k] = v[k] + __shfl_xor(v[k], 32);
    }
    if (kq == 0) {
        red[wv][GP_B1 + r] = v[0];
        red[wv][GP_B1 + 16 + r] = v[1];
        red[wv][GP_W2S + r] = v[2];
        red[wv][GP_W2S + 16 + r] = v[3];
        red[wv][GP_W2N + r] = v[4];
        red[wv][GP_W2N + 16 + r] = v[5];
        if (r == 0) red[wv][GP_B2] = v[6];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < GNN_P; p += 256)
        part[(i64)blockIdx.x * GNN_P + p] = red[0][p] + red[1][p] + red[2][p] + red[3][p];
}

// sums the block partials in block order; Adam step when `update`
__global__ void __launch_bounds__(1024) k_gnn_adam(const float* __restrict__ part, int nb, float* grad, float* prm,
                                                   float* m1, float* m2, u16* w1b, float lr, float c1, float c2,
                                                   int update) {
    for (int p = threadIdx.x; p < GNN_P; p += blockDim.x) {
        float gs = 0.0f;
        for (int b = 0; b < nb; b++) gs = gs + part[(i64)b * GNN_P + p];
        grad[p] = gs;
        if (!update) continue;
        const float m = 0.9f * m1[p] + 0.1f * gs;
        const float v = 0.999f * m2[p] + 0.001f * gs * gs;
        m1[p] = m;
        m2[p] = v;
        const float w = prm[p] - lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
        prm[p] = w;
        if (p < GNN_IN * GNN_HID) w1b[p] = bf16_rne(w);
    }
}

__global__ void k_gnn_stamp(const i32* moves, i64 k, i32* stamp, i32 cur) {
    const i64 x = thread_item();
    if (x >= k) return;
    stamp[moves[x]] = cur;
}

__global__ void __launch_bounds__(256) k_gnn_restore(i64 n, i32* pruned, i32* dest, i64* blk) {
    const i64 i = thread_item();
    i64 a = 0;
    if (i < n && pruned[i]) {
        pruned[i] = 0;
        dest[i] = NULL_PART;
        a = 1;
    }
    const i64 s = block_sum_i64<256>(a);
    if (threadIdx.x == 0) blk[block_id()] = s;
}

// ------------------------------------------------------------------ host
void Gnn::init(Ctx& c) {
    if (prm.p) return;
    prm = Buf<float>(c, GNN_P);
    m1 = Buf<float>(c, GNN_P);
    m2 = Buf<float>(c, GNN_P);
    grad = Buf<float>(c, GNN_P);
    w1b = Buf<u16>(c, GNN_IN * GNN_HID);
    dcnt = Buf<i64>(c, GC_COUNT);
    dloss = Buf<double>(c, 1);
    part = Buf<float>(c, (size_t)GNN_GRAD_BLOCKS * GNN_P);
    reset_counters(c);
}

void Gnn::reset_counters(Ctx& c) {
    fill_i64(c, dcnt, GC_COUNT, 0);
    fill_i64(c, (i64*)dloss.p, 1, 0);
    sweeps = fallbacks = train_steps = 0;
}

void Gnn::ensure(Ctx& c, i64 n, bool train) {
    init(c);
    if (cap < n) {
        cap = std::max<i64>(n, 1);
        pk = Buf<u64>(c, cap);
        in = Buf<u16>(c, cap * GNN_IN);
        s = Buf<float>(c, cap);
        t = Buf<float>(c, cap);
        logit = Buf<float>(c, cap);
        stamp = Buf<i32>(c, cap);
        pruned = Buf<i32>(c, cap);
        blk = Buf<i64>(c, ceil_div(cap, 256) + 1);
        fill_i32(c, stamp, cap, -100);
        fill_i32(c, pruned, cap, 0);
        hcap = false;
        h.reset();
        r.reset();
        ds.reset();
    }
    if (train && !hcap) {
        h = Buf<u16>(c, cap * GNN_HID);
        r = Buf<float>(c, cap);
        ds = Buf<float>(c, cap);
        hcap = true;
    }
}

void Gnn::set_params(Ctx& c, const float* host) {
    init(c);
    std::vector<u16> wb(GNN_IN * GNN_HID);
    for (int k = 0; k < GNN_IN * GNN_HID; k++) wb[k] = bf16_rne(host[GP_W1 + k]);
    copy_h2d(c, prm.p, host, GNN_P * sizeof(float));
    copy_h2d(c, w1b.p, wb.data(), wb.size() * sizeof(u16));
    fill_i32(c, (i32*)m1.p, GNN_P, 0);
    fill_i32(c, (i32*)m2.p, GNN_P, 0);
    c.sync();
    step = 0;
    have_params = true;
}

// blk[0 .. blocks of n) -> cnt[slot] (device-side, no host read)
static void gnn_sum_blocks(Ctx& c, Gnn& G, i64 n, int slot) { sum_i64(c, G.blk, ceil_div(n, 256), G.dcnt.p + slot); }

void gnn_predict(Ctx& c, Gnn& G, LevelPlan& lp, const DGraph& g, const i32* part, i32* dest, const float* gain,
                 const i64* vol, i32 cur, bool first, bool prune, bool keep_hidden) {
    const i64 n = g.n;
    G.ensure(c, n, keep_hidden);
    G.sweeps++;
    if (n == 0) {
        fill_i64(c, G.dcnt.p + GC_SWEEP_ACTIVE, 1, 0);
        return;
    }
    k_gnn_pack<<<grid_items(n, 256), 256, 0, c.stream>>>(n, g.rp, part, dest, gain, G.stamp, cur, first ? 1 : 0, G.pk,
                                                         G.blk);
    PGGC_LAUNCH_CHECK();
    gnn_sum_blocks(c, G, n, GC_SWEEP_ACTIVE);
    k_gnn_acc<<<1, 64, 0, c.stream>>>(G.dcnt, GC_ACTIVE, GC_SWEEP_ACTIVE);
    PGGC_LAUNCH_CHECK();
    const i64* st = lp.sc_start;
    const i32* L = lp.sc_list.p;
    i64 m;
    PGGC_WDISPATCH(g, {
        if ((m = st[1] - st[0]) > 0)
            k_gnn_agg_group<4><<<grid_items(m, 64), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[0], m, G.pk,
                                                                        vol, G.in);
        if ((m = st[2] - st[1]) > 0)
            k_gnn_agg_group<8><<<grid_items(m, 32), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[1], m, G.pk,
                                                                        vol, G.in);
        if ((m = st[3] - st[2]) > 0)
            k_gnn_agg_group<16><<<grid_items(m, 16), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[2], m, G.pk,
                                                                         vol, G.in);
        if ((m = st[4] - st[3]) > 0)
            k_gnn_agg_group<32><<<grid_items(m, 8), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[3], m, G.pk,
                                                                        vol, G.in);
        if ((m = st[5] - st[4]) > 0)
            k_gnn_agg_group<64><<<grid_items(m, 4), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[4], m, G.pk,
                                                                        vol, G.in);
        if ((m = st[CL_HUB] - st[CL_TABLE]) > 0)
            k_gnn_agg_wave<<<grid_items(m, 4), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[CL_TABLE], m, G.pk, vol,
                                                                   G.in);
        if ((m = st[CL_HUB + 1] - st[CL_HUB]) > 0)
            k_gnn_agg_block<<<grid_blocks(m), 256, 0, c.stream>>>(g.rp, g.col, W, g.vw, L + st[CL_HUB], m, G.pk, vol,
                                                                  G.in);
    });
    PGGC_LAUNCH_CHECK();
    const i64 ntiles = (n + 15) / 16;
    k_gnn_dense<<<grid_items(ntiles, 4), 256, 0, c.stream>>>(G.in, n, G.prm, G.w1b, keep_hidden ? G.h.p : nullptr,
                                                            G.s, G.t);
    PGGC_LAUNCH_CHECK();
    const float tau = G.cfg.tau;
    const int pr = prune ? 1 : 0;
    if ((m = st[1] - st[0]) > 0)
        k_gnn_logit_group<4><<<grid_items(m, 64), 256, 0, c.stream>>>(g.rp, g.col, L + st[0], m, dest, G.s, G.t, tau,
                                                                      pr, cur, G.logit, G.pruned);
    if ((m = st[2] - st[1]) > 0)
        k_gnn_logit_group<8><<<grid_items(m, 32), 256, 0, c.stream>>>(g.rp, g.col, L + st[1], m, dest, G.s, G.t, tau,
                                                                      pr, cur, G.logit, G.pruned);
    if ((m = st[3] - st[2]) > 0)
        k_gnn_logit_group<16><<<grid_items(m, 16), 256, 0, c.stream>>>(g.rp, g.col, L + st[2], m, dest, G.s, G.t,
                                                                       tau, pr, cur, G.logit, G.pruned);
    if ((m = st[4] - st[3]) > 0)
        k_gnn_logit_group<32><<<grid_items(m, 8), 256, 0, c.stream>>>(g.rp, g.col, L + st[3], m, dest, G.s, G.t, tau,
                                                                      pr, cur, G.logit, G.pruned);
    if ((m = st[5] - st[4]) > 0)
        k_gnn_logit_group<64><<<grid_items(m, 4), 256, 0, c.stream>>>(g.rp, g.col, L + st[4], m, dest, G.s, G.t, tau,
                                                                      pr, cur, G.logit, G.pruned);
    if ((m = st[CL_HUB] - st[CL_TABLE]) > 0)
        k_gnn_logit_wave<<<grid_items(m, 4), 256, 0, c.stream>>>(g.rp, g.col, L + st[CL_TABLE], m, dest, G.s, G.t, tau, pr,
                                                                 cur, G.logit, G.pruned);
    if ((m = st[CL_HUB + 1] - st[CL_HUB]) > 0)
        k_gnn_logit_block<<<grid_blocks(m), 256, 0, c.stream>>>(g.rp, g.col, L + st[CL_HUB], m, dest, G.s, G.t, tau, pr,
                                                                cur, G.logit, G.pruned);
    PGGC_LAUNCH_CHECK();
    if (prune) {
        k_gnn_count<<<grid_items(n, 256), 256, 0, c.stream>>>(n, G.pruned, cur, G.blk);
        PGGC_LAUNCH_CHECK();
        gnn_sum_blocks(c, G, n, GC_TMP);
        k_gnn_acc<<<1, 64, 0, c.stream>>>(G.dcnt, GC_PRUNED, GC_TMP);
        PGGC_LAUNCH_CHECK();
    }
}

void gnn_begin(Ctx& c, Gnn& G, i64 n, bool train) {
    G.ensure(c, n, train);
    fill_i32(c, G.stamp, n, -100);
    fill_i32(c, G.pruned, n, 0);
}

void gnn_stamp(Ctx& c, Gnn& G, const i32* moves, i64 k, i32 cur) {
    if (k <= 0) return;
    k_gnn_stamp<<<grid_items(k, 256), 256, 0, c.stream>>>(moves, k, G.stamp, cur);
    PGGC_LAUNCH_CHECK();
}

bool gnn_restore(Ctx& c, Gnn& G, i64 n, i32* dest) {
    if (!G.pruned.p || n == 0) return false;
    k_gnn_restore<<<grid_items(n, 256), 256, 0, c.stream>>>(n, G.pruned, dest, G.blk);
    PGGC_LAUNCH_CHECK();
    gnn_sum_blocks(c, G, n, GC_RESTORED);
    if (c.read1(G.dcnt.p + GC_RESTORED) == 0) return false;
    G.fallbacks++;
    return true;
}

void gnn_backward(Ctx& c, Gnn& G, LevelPlan& lp, const DGraph& g, const i32* labels, i32 cur, bool update) {
    const i64 n = g.n;
    if (n == 0) return;
    k_gnn_loss<<<grid_items(n, 256), 256, 0, c.stream>>>(n, g.rp, G.pk, G.logit, labels, G.stamp, cur,
                                                         G.cfg.pos_weight, G.cfg.tau, G.dcnt, G.s, G.r, G.dloss);
    PGGC_LAUNCH_CHECK();
    // G.s now holds g (dLoss/dlogit); s is not needed after the forward pass
    const i64* st = lp.sc_start;
    const i32* L = lp.sc_list.p;
    i64 m;
    if ((m = st[1] - st[0]) > 0)
        k_gnn_ds_group<4><<<grid_items(m, 64), 256, 0, c.stream>>>(g.rp, g.col, L + st[0], m, G.r, G.ds);
    if ((m = st[2] - st[1]) > 0)
        k_gnn_ds_group<8><<<grid_items(m, 32), 256, 0, c.stream>>>(g.rp, g.col, L + st[1], m, G.r, G.ds);
    if ((m = st[3] - st[2]) > 0)
        k_gnn_ds_group<16><<<grid_items(m, 16), 256, 0, c.stream>>>(g.rp, g.col, L + st[2], m, G.r, G.ds);
    if ((m = st[4] - st[3]) > 0)
        k_gnn_ds_group<32><<<grid_items(m, 8), 256, 0, c.stream>>>(g.rp, g.col, L + st[3], m, G.r, G.ds);
    if ((m = st[5] - st[4]) > 0)
        k_gnn_ds_group<64><<<grid_items(m, 4), 256, 0, c.stream>>>(g.rp, g.col, L + st[4], m, G.r, G.ds);
    if ((m = st[CL_HUB] - st[CL_TABLE]) > 0)
        k_gnn_ds_wave<<<grid_items(m, 4), 256, 0, c.stream>>>(g.rp, g.col, L + st[CL_TABLE], m, G.r, G.ds);
    if ((m = st[CL_HUB + 1] - st[CL_HUB]) > 0)
        k_gnn_ds_block<<<grid_blocks(m), 256, 0, c.stream>>>(g.rp, g.col, L + st[CL_HUB], m, G.r, G.ds);
    PGGC_LAUNCH_CHECK();
    const i64 ntiles = (n + 15) / 16;
    const int nb = (int)std::min<i64>(GNN_GRAD_BLOCKS, std::max<i64>(1, (ntiles + 3) / 4));
    k_gnn_grad<<<nb, 256, 0, c.stream>>>(n, G.in, G.h, G.s, G.ds, G.prm, G.part);
    PGGC_LAUNCH_CHECK();
    float c1 = 1.0f, c2 = 1.0f;
    if (update) {
        G.step++;
        G.train_steps++;
        c1 = (float)(1.0 - std::pow(0.9, (double)G.step));
        c2 = (float)(1.0 - std::pow(0.999, (double)G.step));
    }
    k_gnn_adam<<<1, 1024, 0, c.stream>>>(G.part, nb, G.grad, G.prm, G.m1, G.m2, G.w1b, G.cfg.lr, c1, c2,
                                          update ? 1 : 0);
    PGGC_LAUNCH_CHECK();
}

}  // namespace pggc
