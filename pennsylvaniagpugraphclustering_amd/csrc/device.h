// Device-side building blocks: wave and block reductions and scans (wave64).
#pragma once
#include "common.h"

namespace pggc {

// Cross-lane fence.  On gfx950 a ds_bpermute (the instruction behind
// __shfl*) issued a few instructions after the EXEC restore that ends a
// divergent loop or branch gave run-to-run different reductions, while the
// same reduction after a wave-uniform loop, or through LDS, was exact.  Every
// value that enters a shuffle first passes through this fence: 16 wait
// states in an asm statement that takes the value as an in/out operand, so
// the compiler cannot hoist the shuffle above the wait states (a plain asm
// with a memory clobber does not order register-only shuffles).
template <class T>
__device__ __forceinline__ void xlane_fence(T& v) {
    asm volatile("s_nop 7\n\ts_nop 7" : "+v"(v));
}

__device__ __forceinline__ i64 wave_sum_i64(i64 v) {
    xlane_fence(v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ u64 wave_max_u64(u64 v) {
    xlane_fence(v);
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
    xlane_fence(v);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}
template <int G>
__device__ __forceinline__ u64 group_max_u64(u64 v) {
    xlane_fence(v);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        u64 y = __shfl_xor(v, o, G);
        v = y > v ? y : v;
    }
    return v;
}
template <int G>
__device__ __forceinline__ i64 group_max_i64(i64 v) {
    xlane_fence(v);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        i64 y = __shfl_xor(v, o, G);
        v = y > v ? y : v;
    }
    return v;
}

// Block reductions through an LDS halving tree (no lane shuffles, so they
// are safe after loops whose lanes exit at different iterations).  BS is the
// block size (a power of two); every thread must call them and receives the
// result.
template <int BS>
__device__ __forceinline__ i64 block_sum_i64(i64 v) {
    __shared__ i64 sh[BS];
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] = sh[t] + sh[t + o];
        __syncthreads();
    }
    const i64 r = sh[0];
    __syncthreads();
    return r;
}

template <int BS>
__device__ __forceinline__ u64 block_max_u64(u64 v) {
    __shared__ u64 sh[BS];
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int o = BS / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] = sh[t + o] > sh[t] ? sh[t + o] : sh[t];
        __syncthreads();
    }
    const u64 r = sh[0];
    __syncthreads();
    return r;
}

// Exclusive prefix sum over the block in thread order; *total = block sum.
__device__ __forceinline__ i64 block_excl_scan_i64(i64 v, i64* total) {
    __shared__ i64 ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    i64 x = v;
    xlane_fence(x);
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

// Edge-chunk row reductions: each block owns a contiguous range of EC edges;
// each thread owns EPT consecutive edges and walks their rows with a
// monotone row pointer (per-row partial sums then go out with atomics).
constexpr int EB = 256, EPT = 16, EC = EB * EPT;

struct ChunkRows {
    i64 eb, ee, rlo, rhi;
};
__device__ __forceinline__ void chunk_rows(const i64* rp, i64 n, i64 nnz, ChunkRows& cr) {
    __shared__ i64 s_lo, s_hi;
    cr.eb = (i64)blockIdx.x * EC;
    cr.ee = cr.eb + EC < nnz ? cr.eb + EC : nnz;
    if (threadIdx.x == 0) {
        s_lo = upper_row(rp, 0, n, cr.eb);
        s_hi = upper_row(rp, 0, n, cr.ee - 1) + 1;
    }
    __syncthreads();
    cr.rlo = s_lo;
    cr.rhi = s_hi;
}

}  // namespace pggc
