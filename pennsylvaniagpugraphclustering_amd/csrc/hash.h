// Open-addressing (linear probing) cluster-id -> int64 weight tables in LDS or
// HBM, and an in-LDS bitonic sort of (key, value) pairs.
#pragma once
#include "common.h"

namespace pggc {

constexpr i32 EMPTY_KEY = -1;  // cluster ids are >= 0

__device__ __forceinline__ void table_add_lds(i32* keys, i64* vals, u32 mask, i32 key, i64 w) {
    u32 h = slot_hash(key, mask);
    while (true) {
        i32 old = atomicCAS(&keys[h], EMPTY_KEY, key);
        if (old == EMPTY_KEY || old == key) {
            atomicAdd((unsigned long long*)&vals[h], (unsigned long long)w);
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ void table_add_global(i32* keys, i64* vals, u32 mask, i32 key, i64 w) {
    u32 h = slot_hash(key, mask);
    while (true) {
        i32 cur = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == key) break;
        if (cur == EMPTY_KEY) {
            i32 old = atomicCAS(&keys[h], EMPTY_KEY, key);
            if (old == EMPTY_KEY || old == key) break;
        }
        h = (h + 1) & mask;
    }
    atomicAdd((unsigned long long*)&vals[h], (unsigned long long)w);
}

// Packed HBM table: one u64 per slot, key in the high 32 bits and the summed
// weight in the low 32 bits (callers guarantee every sum fits 32 bits), so an
// insert touches one cache line and a new key costs a single CAS.
constexpr u64 PEMPTY = 0xFFFFFFFF00000000ull;  // key -1, weight 0

__device__ __forceinline__ void ptable_add_from(u64* tab, u32 mask, i32 key, u64 w, u32 h) {
    while (true) {
        const u64 cur = __hip_atomic_load(&tab[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((u32)(cur >> 32) == (u32)key) break;
        if (cur == PEMPTY) {
            const u64 old = atomicCAS((unsigned long long*)&tab[h], (unsigned long long)PEMPTY,
                                      (unsigned long long)(((u64)(u32)key << 32) | w));
            if (old == PEMPTY) return;
            if ((u32)(old >> 32) == (u32)key) break;
        }
        h = (h + 1) & mask;
    }
    atomicAdd((unsigned long long*)&tab[h], (unsigned long long)w);
}
__device__ __forceinline__ void ptable_add(u64* tab, u32 mask, i32 key, u64 w) {
    ptable_add_from(tab, mask, key, w, slot_hash(key, mask));
}

// Probe sequences of the hub tables: a table of mask + 1 slots is split into
// slices of smask + 1 slots (smask <= mask, both powers of two minus one);
// a key hashes over the whole table and probes linearly within its slice,
// wrapping at the slice's end.  smask == mask is plain linear probing over
// the table.  Slices let a build hold one slice in LDS at a time.
__device__ __forceinline__ u32 probe_next(u32 h, u32 smask) { return (h & ~smask) | ((h + 1) & smask); }

// Signed update of a table that is kept current between sweeps.  At most
// smask+1 probes; returns false when the key is neither found nor
// insertable (slice full, or a decrement of a key the table does not hold),
// which leaves the table stale.  A decrement never borrows across the packed
// key: the slot holds at least the weight being moved out of it.  A key new
// to the table appends its slot to the table's occupied-slot list
// (list[*cnt], *cnt += 1).
__device__ __forceinline__ bool ptable_update(u64* tab, u32 mask, u32 smask, i32 key, i64 w, u32* list, i32* cnt) {
    u32 h = slot_hash(key, mask);
    for (u32 probe = 0; probe <= smask; probe++) {
        const u64 cur = __hip_atomic_load(&tab[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((u32)(cur >> 32) == (u32)key) {
            atomicAdd((unsigned long long*)&tab[h], (unsigned long long)w);
            return true;
        }
        if (cur == PEMPTY) {
            if (w < 0) return false;
            const u64 old = atomicCAS((unsigned long long*)&tab[h], (unsigned long long)PEMPTY,
                                      (unsigned long long)(((u64)(u32)key << 32) | (u64)w));
            if (old == PEMPTY) {
                list[atomicAdd(cnt, 1)] = h;
                return true;
            }
            if ((u32)(old >> 32) == (u32)key) {
                atomicAdd((unsigned long long*)&tab[h], (unsigned long long)w);
                return true;
            }
        }
        h = probe_next(h, smask);
    }
    return false;
}
__device__ __forceinline__ bool table_update(i32* keys, i64* vals, u32 mask, u32 smask, i32 key, i64 w, u32* list,
                                             i32* cnt) {
    u32 h = slot_hash(key, mask);
    for (u32 probe = 0; probe <= smask; probe++) {
        const i32 cur = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool hit = cur == key;
        if (!hit && cur == EMPTY_KEY) {
            if (w < 0) return false;
            const i32 old = atomicCAS(&keys[h], EMPTY_KEY, key);
            if (old == EMPTY_KEY) list[atomicAdd(cnt, 1)] = h;
            hit = old == EMPTY_KEY || old == key;
        }
        if (hit) {
            atomicAdd((unsigned long long*)&vals[h], (unsigned long long)w);
            return true;
        }
        h = probe_next(h, smask);
    }
    return false;
}
// weight of `key` in a table (0 if absent), at most smask+1 probes
__device__ __forceinline__ i64 ptable_get(const u64* tab, u32 mask, u32 smask, i32 key) {
    u32 h = slot_hash(key, mask);
    for (u32 probe = 0; probe <= smask; probe++) {
        const u64 cur = tab[h];
        if ((u32)(cur >> 32) == (u32)key) return (i64)(cur & 0xffffffffull);
        if (cur == PEMPTY) return 0;
        h = probe_next(h, smask);
    }
    return 0;
}
__device__ __forceinline__ i64 table_get(const i32* keys, const i64* vals, u32 mask, u32 smask, i32 key) {
    u32 h = slot_hash(key, mask);
    for (u32 probe = 0; probe <= smask; probe++) {
        const i32 cur = keys[h];
        if (cur == key) return vals[h];
        if (cur == EMPTY_KEY) return 0;
        h = probe_next(h, smask);
    }
    return 0;
}

// Work-item pre-aggregation: a block merges the (key, weight) pairs of one
// work item (at most LSLOTS / 2 pairs) in an LDS table, then adds each
// distinct key once to the HBM table.  Hot keys (a hub whose neighbours sit in
// a few clusters) then cost one global atomic per item instead of one per
// edge.
template <int BS, int LSLOTS>
__device__ __forceinline__ void lds_agg_clear(i32* lk, i64* lv) {
    for (int q = threadIdx.x; q < LSLOTS; q += BS) {
        lk[q] = EMPTY_KEY;
        lv[q] = 0;
    }
    __syncthreads();
}
template <int BS, int LSLOTS>
__device__ __forceinline__ void lds_agg_flush(const i32* lk, const i64* lv, i32* gkeys, i64* gvals, u32 gmask) {
    __syncthreads();
    for (int q = threadIdx.x; q < LSLOTS; q += BS) {
        const i32 key = lk[q];
        const i64 w = lv[q];
        if (key != EMPTY_KEY) table_add_global(gkeys, gvals, gmask, key, w);
    }
}
// Batched flush into a packed table: each thread first tries to claim the
// home slots of all its keys with CAS (independent atomics in flight
// together; a key new to the table then costs one atomic), then adds to the
// slots that turned out to hold the key already (no-return atomics); only
// keys that meet another key fall back to probing.
template <int BS, int LSLOTS>
__device__ __forceinline__ void lds_agg_flush_packed(const i32* lk, const i64* lv, u64* tab, u32 gmask) {
    __syncthreads();
    constexpr int R = LSLOTS / BS;
    i32 key[R];
    u64 w[R], cur[R];
    u32 h[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        key[r] = lk[threadIdx.x + r * BS];
        w[r] = (u64)lv[threadIdx.x + r * BS];
        h[r] = slot_hash(key[r], gmask);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        cur[r] = key[r] != EMPTY_KEY ? (u64)atomicCAS((unsigned long long*)&tab[h[r]], (unsigned long long)PEMPTY,
                                                      (unsigned long long)(((u64)(u32)key[r] << 32) | w[r]))
                                     : PEMPTY;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (key[r] == EMPTY_KEY || cur[r] == PEMPTY) continue;  // done, or the CAS claimed the slot
        if ((u32)(cur[r] >> 32) == (u32)key[r]) atomicAdd((unsigned long long*)&tab[h[r]], (unsigned long long)w[r]);
        else ptable_add_from(tab, gmask, key[r], w[r], (h[r] + 1) & gmask);
    }
}

// Ascending bitonic sort of S (power of two) pairs held in LDS by a block of
// BS threads.  Must be called by every thread of the block.
template <int BS>
__device__ __forceinline__ void bitonic_sort_lds(i32* keys, i64* vals, int S) {
    for (int k = 2; k <= S; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < S; i += BS) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool up = (i & k) == 0;
                    i32 a = keys[i], b = keys[ixj];
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[ixj] = a;
                        i64 t = vals[i];
                        vals[i] = vals[ixj];
                        vals[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace pggc
