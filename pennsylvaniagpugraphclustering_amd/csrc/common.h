// Shared definitions for the PGGC HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace pggc {

using i32 = int32_t;
using i64 = int64_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

// Sentinel labels of the local mover (src/local_mover.cpp:74-80).
constexpr i32 NULL_PART = -1;
constexpr i32 NO_MOVE = -3;
constexpr i32 NEW_PART = -4;
constexpr float AFTERBURNER_EPS = 0.1f;
// Degree cutoffs of the ordering pass (src/header/ordering.h:17-19).
constexpr i64 MID_CUTOFF = 64;
constexpr i64 LARGE_CUTOFF = 128;
constexpr i64 MASSIVE_CUTOFF = 15000;
// leidenR constants (src/leidenR.cpp:38-39).
constexpr i32 ORD_MAX = 0x7fffffff;
constexpr i32 SPLIT = 1000000000;

constexpr int WAVE = 64;

// Error raised inside the library; converted to a status code at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PGGC_HIP(call)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::pggc::Error(-3, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

#define PGGC_LAUNCH_CHECK() PGGC_HIP(hipGetLastError())

// Edge-weight accessors: unit (uniform graphs), int32 (input graphs) and
// int64 (coarse graphs, whose summed weights can exceed int32).
struct WUnit {
    static constexpr int bytes = 0;  // weight bytes read per edge
    __device__ __forceinline__ i64 operator()(i64) const { return 1; }
};
struct W32 {
    static constexpr int bytes = 4;
    const i32* p;
    __device__ __forceinline__ i64 operator()(i64 j) const { return p[j]; }
};
struct W64 {
    static constexpr int bytes = 8;
    const i64* p;
    __device__ __forceinline__ i64 operator()(i64 j) const { return p[j]; }
};

enum WKind : int { W_UNIT = 0, W_I32 = 1, W_I64 = 2 };

// One level of the multilevel hierarchy, resident in HBM.
struct DGraph {
    i64 n = 0;
    i64 nnz = 0;
    const i64* rp = nullptr;   // row offsets [n+1]
    const i32* col = nullptr;  // column ids [nnz]
    const void* w = nullptr;   // weights [nnz] (null when W_UNIT)
    int wk = W_UNIT;
    const i64* vw = nullptr;   // vertex weights [n]
};

#define PGGC_WDISPATCH(G, ...)                               \
    do {                                                      \
        if ((G).wk == ::pggc::W_UNIT) {                       \
            ::pggc::WUnit W{};                                \
            __VA_ARGS__;                                      \
        } else if ((G).wk == ::pggc::W_I32) {                 \
            ::pggc::W32 W{(const ::pggc::i32*)(G).w};         \
            __VA_ARGS__;                                      \
        } else {                                              \
            ::pggc::W64 W{(const ::pggc::i64*)(G).w};         \
            __VA_ARGS__;                                      \
        }                                                     \
    } while (0)

// xorshift32 of the local mover (src/local_mover.cpp:82-87).
__host__ __device__ __forceinline__ u32 xorshift32(u32 x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// MurmurHash3_x86_32 of a 4-byte key with seed 0 (the hash behind the
// reference's Kokkos::pod_hash<int32>, src/leidenR.cpp:37).
__host__ __device__ __forceinline__ u32 murmur3_i32(i32 key) {
    u32 k1 = (u32)key;
    u32 h1 = 0;
    k1 *= 0xcc9e2d51u;
    k1 = (k1 << 15) | (k1 >> 17);
    k1 *= 0x1b873593u;
    h1 ^= k1;
    h1 = (h1 << 13) | (h1 >> 19);
    h1 = h1 * 5u + 0xe6546b64u;
    h1 ^= 4u;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    h1 ^= h1 >> 16;
    return h1;
}

// k-th seed draw of a run (stand-in for std::random_device in leidenR).
__host__ __device__ __forceinline__ u32 seed_draw(u64 seed, u64 k) {
    u64 z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (u32)z;
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11):
// ten rounds of two 32x32->64 multiplies with the published multipliers and
// Weyl key increments.  The synthetic graph generators draw from it so that
// a CPU build of the same draws reproduces the same graph.
struct Philox4 {
    u32 v[4];
};
__host__ __device__ __forceinline__ Philox4 philox4x32_10(u32 c0, u32 c1, u32 c2, u32 c3, u32 k0, u32 k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const u64 p0 = (u64)0xD2511F53u * c0;
        const u64 p1 = (u64)0xCD9E8D57u * c2;
        const u32 n0 = (u32)(p1 >> 32) ^ c1 ^ k0;
        const u32 n2 = (u32)(p0 >> 32) ^ c3 ^ k1;
        c1 = (u32)p1;
        c3 = (u32)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{{c0, c1, c2, c3}};
}
// streams of the generators (counter word 3)
constexpr u32 STREAM_RMAT = 0, STREAM_WEIGHT = 1, STREAM_LABEL = 2;
// LFR-style generator: degrees, community sizes, internal and external stub
// keys, vertex permutation
constexpr u32 STREAM_LFR_DEG = 3, STREAM_LFR_COMM = 4, STREAM_LFR_ISTUB = 5, STREAM_LFR_ESTUB = 6,
              STREAM_LFR_PERM = 7;

// Monotone map float -> u32 (for packed argmax keys).
__device__ __forceinline__ u32 ord_f32(float f) {
    u32 b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(u32 o) {
    u32 b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(b);
}

// Hash slot for a cluster id in a power-of-two table.
__device__ __forceinline__ u32 slot_hash(i32 key, u32 mask) {
    return ((u32)key * 0x9E3779B1u) >> 7 & mask;
}

inline i64 ceil_div(i64 a, i64 b) { return (a + b - 1) / b; }

// Exact launch grids: one work item per thread, lane group, wave or block.
// Blocks are laid out on a 2-D grid so counts beyond one dimension's limit
// still launch; kernels recover the linear block id with block_id().
//
// Kernels whose per-item body has divergent control flow or divergent stores
// take one item per thread (no per-lane grid-stride loop).  On gfx950 a
// per-lane grid-stride loop around such a body gave run-to-run differences in
// the afterburner at 160k+ items, while the same body with one item per
// thread was exact; grid-stride loops remain only around straight-line
// bodies (fills, copies, reductions).
constexpr i64 GRID_X = 32768;
inline dim3 grid_blocks(i64 blocks) {
    if (blocks < 1) blocks = 1;
    if (blocks <= GRID_X) return dim3((unsigned)blocks);
    return dim3((unsigned)GRID_X, (unsigned)ceil_div(blocks, GRID_X));
}
inline dim3 grid_items(i64 items, int per_block) { return grid_blocks(ceil_div(items, per_block)); }
__device__ __forceinline__ i64 block_id() { return (i64)blockIdx.y * gridDim.x + blockIdx.x; }
__device__ __forceinline__ i64 thread_item() { return block_id() * blockDim.x + threadIdx.x; }

inline int grid_for(i64 work, int per_block, int cap = 1 << 20) {
    i64 g = ceil_div(work, per_block);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

}  // namespace pggc
