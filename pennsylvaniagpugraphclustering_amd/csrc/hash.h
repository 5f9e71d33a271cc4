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
