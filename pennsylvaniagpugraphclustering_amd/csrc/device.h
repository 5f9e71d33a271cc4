// Device-side building blocks: wave and block reductions and scans (wave64).
#pragma once
#include "common.h"

namespace pggc {

// Cross-lane guard.  On gfx950 with this toolchain a ds_bpermute (the
// instruction behind __shfl*) issued right after the EXEC restore that ends a
// divergent loop or branch was observed to read stale lanes (run-to-run
// different reductions); 16 wait states before the cross-lane sequence made
// every reduction exact and deterministic.  Every cross-lane helper and every
// kernel-level shuffle/ballot sequence starts with this guard.
__device__ __forceinline__ void xlane_guard() { asm volatile("s_nop 7\n\ts_nop 7" ::: "memory"); }

__device__ __forceinline__ i64 wave_sum_i64(i64 v) {
    xlane_guard();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ u64 wave_max_u64(u64 v) {
    xlane_guard();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        u64 y = __shfl_xor(v, o);
        v = y > v ? y : v;
    }
    return v;
}

// group reductions over `G` consecutive lanes (G a power of two <= 64)
template <int G>
__device__ __forceinline__ i64 group_sum_i64(i64 v) {
    xlane_guard();
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}
template <int G>
__device__ __forceinline__ u64 group_max_u64(u64 v) {
    xlane_guard();
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        u64 y = __shfl_xor(v, o, G);
        v = y > v ? y : v;
    }
    return v;
}
template <int G>
__device__ __forceinline__ i64 group_max_i64(i64 v) {
    xlane_guard();
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        i64 y = __shfl_xor(v, o, G);
        v = y > v ? y : v;
    }
    return v;
}

// Sum over the whole block; every thread receives the total.
__device__ __forceinline__ i64 block_sum_i64(i64 v) {
    __shared__ i64 ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    v = wave_sum_i64(v);
    if (lane == 0) ws[w] = v;
    __syncthreads();
    i64 t = 0;
    for (int k = 0; k < nw; k++) t += ws[k];
    __syncthreads();
    return t;
}

__device__ __forceinline__ u64 block_max_u64(u64 v) {
    __shared__ u64 ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    v = wave_max_u64(v);
    if (lane == 0) ws[w] = v;
    __syncthreads();
    u64 t = 0;
    for (int k = 0; k < nw; k++) t = ws[k] > t ? ws[k] : t;
    __syncthreads();
    return t;
}

// Exclusive prefix sum over the block in thread order; *total = block sum.
__device__ __forceinline__ i64 block_excl_scan_i64(i64 v, i64* total) {
    __shared__ i64 ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    i64 x = v;
    xlane_guard();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        i64 y = __shfl_up(x, o);
        x += (lane >= o) ? y : (i64)0;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    i64 pre = 0, tot = 0;
    for (int k = 0; k < nw; k++) {
        if (k < w) pre += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

// Row lookup helper: largest r in [lo, hi) with rp[r] <= j.
__device__ __forceinline__ i64 upper_row(const i64* rp, i64 lo, i64 hi, i64 j) {
    while (hi - lo > 1) {
        i64 m = (lo + hi) >> 1;
        if (rp[m] <= j) lo = m;
        else hi = m;
    }
    return lo;
}

// Advance row r (with nxt = rp[r+1]) until it contains edge j: a few linear
// steps, then a binary search, so runs of empty rows cost O(log) instead of
// their length.
__device__ __forceinline__ i64 advance_row(const i64* rp, i64 r, i64 rhi, i64 j, i64& nxt) {
    int steps = 0;
    while (j >= nxt) {
        if (++steps > 4) {
            r = upper_row(rp, r, rhi, j);
            nxt = rp[r + 1];
            return r;
        }
        r++;
        nxt = rp[r + 1];
    }
    return r;
}

}  // namespace pggc
