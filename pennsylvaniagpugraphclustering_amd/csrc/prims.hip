// Context allocator and data-parallel primitives: fill, scan, ordered
// compaction, stable LSD radix sort, exact integer reductions.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "ctx.h"
#include "device.h"

namespace pggc {

// ------------------------------------------------------------------ allocator
static size_t round_bytes(size_t b) {
    if (b < 256) return 256;
    if (b < (1u << 20)) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t mb = 1u << 20;
    return (b + mb - 1) / mb * mb;
}

void* Ctx::alloc(size_t bytes) {
    size_t sz = round_bytes(bytes);
    auto it = free_blocks.lower_bound(sz);
    if (it != free_blocks.end() && it->first <= sz + sz / 4 + 4096) {
        void* p = it->second;
        size_t s = it->first;
        free_blocks.erase(it);
        bytes_cached -= s;
        live[p] = s;
        bytes_live += s;
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        trim();
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(-2, "device allocation of " + std::to_string(sz) + " bytes failed (" +
                                std::to_string(bytes_live) + " bytes live in this context)");
        }
    }
    live[p] = sz;
    bytes_live += sz;
    return p;
}

void Ctx::mem_note(const char* where) const {
    if (verbose < 2) return;
    fprintf(stderr, "[mem] %s: %.2f GiB live, %.2f GiB cached\n", where, bytes_live / 1073741824.0,
            bytes_cached / 1073741824.0);
}

void Ctx::release(void* p) {
    auto it = live.find(p);
    if (it == live.end()) return;
    free_blocks.emplace(it->second, p);
    bytes_cached += it->second;
    bytes_live -= it->second;
    live.erase(it);
}

void Ctx::trim() {
    (void)hipStreamSynchronize(stream);
    for (auto& kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    bytes_cached = 0;
}

hipEvent_t Ctx::kt_event() {
    if (!kt.pool.empty()) {
        hipEvent_t e = kt.pool.back();
        kt.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    PGGC_HIP(hipEventCreate(&e));
    return e;
}

constexpr int KT_MAX = 4096;  // pending regions before a resolve

int Ctx::kt_start(int slot) {
    if (kt.open == 0 && kt.pend.size() >= KT_MAX) kt_resolve();
    if (!kt.dcnt) {
        kt.dcnt = (u64*)alloc(KT_MAX * sizeof(u64));
        fill_i64(*this, (i64*)kt.dcnt, KT_MAX, 0);
    }
    KTimer::Pending p{slot, kt_event(), kt_event(), 0.0, false};
    PGGC_HIP(hipEventRecord(p.a, stream));
    kt.pend.push_back(p);
    kt.open++;
    return (int)kt.pend.size() - 1;
}

void Ctx::kt_stop(int tok, double bytes) {
    KTimer::Pending& p = kt.pend[tok];
    p.bytes = bytes;
    PGGC_HIP(hipEventRecord(p.b, stream));
    p.stopped = true;
    kt.open--;
}

void Ctx::kt_resolve() {
    if (kt.pend.empty()) return;
    std::vector<u64> cnt(kt.pend.size(), 0);
    PGGC_HIP(hipMemcpyAsync(cnt.data(), kt.dcnt, cnt.size() * sizeof(u64), hipMemcpyDeviceToHost, stream));
    PGGC_HIP(hipStreamSynchronize(stream));
    fill_i64(*this, (i64*)kt.dcnt, (i64)cnt.size(), 0);
    for (size_t x = 0; x < kt.pend.size(); x++) {
        const auto& p = kt.pend[x];
        if (p.stopped) {  // regions left open by an error are dropped
            float t = 0.0f;
            PGGC_HIP(hipEventElapsedTime(&t, p.a, p.b));
            kt.launches[p.slot]++;
            kt.ms[p.slot] += t;
            kt.bytes[p.slot] += p.bytes + (double)cnt[x];
        }
        kt.pool.push_back(p.a);
        kt.pool.push_back(p.b);
    }
    kt.pend.clear();
    kt.open = 0;
}

void Ctx::kt_free() {
    for (auto& p : kt.pend) {
        kt.pool.push_back(p.a);
        kt.pool.push_back(p.b);
    }
    kt.pend.clear();
    for (hipEvent_t e : kt.pool) (void)hipEventDestroy(e);
    kt.pool.clear();
}

// Small readbacks: a one-wave kernel on the context stream stores the values
// into the pinned host slots (coherent host memory), makes the stores
// visible at system scope, then the host waits for the stream.
__global__ void k_readback(const i64* __restrict__ dev, i64* host, int count) {
    const int t = threadIdx.x;
    if (t < count) host[t] = dev[t];
    __threadfence_system();
}

const i64* Ctx::read(const i64* dev, int count) {
    if (count < 1 || count > 64) throw Error(-5, "read: 1..64 values");
    k_readback<<<1, 64, 0, stream>>>(dev, pinned_dev, count);
    PGGC_LAUNCH_CHECK();
    PGGC_HIP(hipStreamSynchronize(stream));
    return pinned;
}

// ---------------------------------------------------------------------- fill
template <class T>
__global__ void k_fill(T* p, i64 n, T v) {
    for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x) p[x] = v;
}
__global__ void k_iota(i32* p, i64 n) {
    for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x) p[x] = (i32)x;
}

void fill_i32(Ctx& c, i32* p, i64 n, i32 v) {
    if (n <= 0) return;
    k_fill<i32><<<grid_for(n, 256 * 4, 8192), 256, 0, c.stream>>>(p, n, v);
    PGGC_LAUNCH_CHECK();
}
void fill_i64(Ctx& c, i64* p, i64 n, i64 v) {
    if (n <= 0) return;
    k_fill<i64><<<grid_for(n, 256 * 4, 8192), 256, 0, c.stream>>>(p, n, v);
    PGGC_LAUNCH_CHECK();
}
// device-to-device copy as a kernel on the context stream (16-byte words
// when both ends are 16-byte aligned, else 4-byte words; every copy in the
// library is a multiple of 4 bytes)
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, i64 n) {
    const i64 i = thread_item();
    if (i < n) dst[i] = src[i];
}
__global__ void k_copy4(const u32* __restrict__ src, u32* __restrict__ dst, i64 n) {
    const i64 i = thread_item();
    if (i < n) dst[i] = src[i];
}
void copy_dev(Ctx& c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0 || dst == src) return;
    if (bytes % 4 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 4 != 0)
        throw Error(-5, "copy_dev: size and addresses must be multiples of 4 bytes");
    if (((uintptr_t)dst | (uintptr_t)src) % 16 == 0) {
        // 16-byte words for the bulk, 4-byte words for a tail of up to 12 bytes
        const i64 n16 = (i64)(bytes / 16);
        if (n16 > 0) {
            k_copy16<<<grid_items(n16, 256), 256, 0, c.stream>>>((const uint4*)src, (uint4*)dst, n16);
            PGGC_LAUNCH_CHECK();
        }
        const i64 n4 = (i64)(bytes % 16) / 4;
        if (n4 > 0) {
            k_copy4<<<1, 64, 0, c.stream>>>((const u32*)src + 4 * n16, (u32*)dst + 4 * n16, n4);
            PGGC_LAUNCH_CHECK();
        }
        return;
    }
    const i64 n = (i64)(bytes / 4);
    k_copy4<<<grid_items(n, 256), 256, 0, c.stream>>>((const u32*)src, (u32*)dst, n);
    PGGC_LAUNCH_CHECK();
}

void copy_h2d(Ctx& c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    c.sync();  // the staging area is free and nothing queued still reads dst
    if (c.stage_cap < bytes) {
        if (c.stage) (void)hipHostFree(c.stage);
        c.stage = nullptr;
        c.stage_cap = 0;
        const size_t cap = std::max<size_t>(bytes, 1 << 16);
        PGGC_HIP(hipHostMalloc(&c.stage, cap, hipHostMallocCoherent));
        PGGC_HIP(hipHostGetDevicePointer(&c.stage_dev, c.stage, 0));
        c.stage_cap = cap;
    }
    std::memcpy(c.stage, src, bytes);
    copy_dev(c, dst, c.stage_dev, bytes);
    c.sync();
}

void iota_i32(Ctx& c, i32* p, i64 n) {
    if (n <= 0) return;
    k_iota<<<grid_for(n, 256 * 4, 8192), 256, 0, c.stream>>>(p, n);
    PGGC_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------- scan
constexpr int ST = 256;          // threads per scan block
constexpr int SI = 8;            // items per thread
constexpr int STILE = ST * SI;   // elements per tile

template <class Load>
__global__ void __launch_bounds__(ST) k_tile_sums(Load ld, i64 n, i64* bsum) {
    i64 base = (i64)blockIdx.x * STILE;
    i64 s = 0;
    for (int k = 0; k < SI; k++) {
        i64 x = base + (i64)k * ST + threadIdx.x;
        if (x < n) s += ld(x);
    }
    i64 tot = block_sum_i64<ST>(s);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <class Load>
__global__ void __launch_bounds__(ST) k_tile_scan(Load ld, i64 n, const i64* boff, i64* out) {
    i64 base = (i64)blockIdx.x * STILE + (i64)threadIdx.x * SI;
    i64 v[SI];
    i64 s = 0;
#pragma unroll
    for (int k = 0; k < SI; k++) {
        i64 x = base + k;
        v[k] = x < n ? ld(x) : 0;
        s += v[k];
    }
    i64 tot;
    i64 pre = block_excl_scan_i64(s, &tot) + (boff ? boff[blockIdx.x] : 0);
#pragma unroll
    for (int k = 0; k < SI; k++) {
        i64 x = base + k;
        if (x < n) out[x] = pre;
        pre += v[k];
        if (x == n - 1) out[n] = pre;
    }
}

struct LoadI64 {
    const i64* p;
    __device__ __forceinline__ i64 operator()(i64 x) const { return p[x]; }
};
struct LoadI32 {
    const i32* p;
    __device__ __forceinline__ i64 operator()(i64 x) const { return p[x]; }
};
struct LoadFlag {
    const i32* p;
    __device__ __forceinline__ i64 operator()(i64 x) const { return p[x] != 0 ? 1 : 0; }
};

template <class Load>
static void scan_generic(Ctx& c, Load ld, i64* out, i64 n) {
    if (n <= 0) {
        fill_i64(c, out, 1, 0);
        return;
    }
    i64 nb = ceil_div(n, STILE);
    if (nb == 1) {
        k_tile_scan<Load><<<1, ST, 0, c.stream>>>(ld, n, nullptr, out);
        PGGC_LAUNCH_CHECK();
        return;
    }
    Buf<i64> bs(c, nb + 1);
    k_tile_sums<Load><<<(int)nb, ST, 0, c.stream>>>(ld, n, bs);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, bs, bs, nb);
    k_tile_scan<Load><<<(int)nb, ST, 0, c.stream>>>(ld, n, bs, out);
    PGGC_LAUNCH_CHECK();
}

void scan_i64(Ctx& c, const i64* in, i64* out, i64 n) {
    if (in == out && n > 0) {
        // the tile scan reads a whole tile before writing it, and tiles are
        // disjoint, so in-place operation is safe
    }
    scan_generic(c, LoadI64{in}, out, n);
}
void scan_i32_to_i64(Ctx& c, const i32* in, i64* out, i64 n) { scan_generic(c, LoadI32{in}, out, n); }

// ---------------------------------------------------------- ordered compaction
__global__ void __launch_bounds__(ST) k_compact_write(const i32* flag, const i32* src, i64 n, const i64* boff,
                                                       i32* out, i64* count) {
    i64 base = (i64)blockIdx.x * STILE + (i64)threadIdx.x * SI;
    int f[SI];
    i64 s = 0;
#pragma unroll
    for (int k = 0; k < SI; k++) {
        i64 x = base + k;
        f[k] = (x < n && flag[x] != 0) ? 1 : 0;
        s += f[k];
    }
    i64 tot;
    i64 pre = block_excl_scan_i64(s, &tot) + boff[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SI; k++) {
        i64 x = base + k;
        if (f[k]) out[pre++] = src ? src[x] : (i32)x;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count = boff[blockIdx.x] + tot;
}

i64 compact_flags(Ctx& c, const i32* flag, const i32* src, i64 n, i32* out, i64* count_dev) {
    if (n <= 0) {
        fill_i64(c, count_dev, 1, 0);
        return 0;
    }
    i64 nb = ceil_div(n, STILE);
    Buf<i64> bs(c, nb + 1);
    k_tile_sums<LoadFlag><<<(int)nb, ST, 0, c.stream>>>(LoadFlag{flag}, n, bs);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, bs, bs, nb);
    k_compact_write<<<(int)nb, ST, 0, c.stream>>>(flag, src, n, bs, out, count_dev);
    PGGC_LAUNCH_CHECK();
    return c.read1(count_dev);
}

// ---------------------------------------------------------------- radix sort
// Stable LSD sort of (key, value) pairs on the low `bits` key bits: rocPRIM's
// onesweep radix sort (a plain library sort, compiled from the ROCm headers),
// with its temporary storage and the second buffers from the context's pool.
template <class V>
void radix_sort_pairs_t(Ctx& c, u64* keys, V* vals, i64 n, int bits) {
    if (n <= 1 || bits <= 0) return;
    Buf<u64> k2(c, n);
    Buf<V> v2(c, n);
    rocprim::double_buffer<u64> kd(keys, k2.p);
    rocprim::double_buffer<V> vd(vals, v2.p);
    const unsigned int end_bit = (unsigned int)std::min(bits, 64);
    size_t tb = 0;
    PGGC_HIP(rocprim::radix_sort_pairs(nullptr, tb, kd, vd, (size_t)n, 0u, end_bit, c.stream));
    Buf<u64> tmp(c, std::max<size_t>((tb + 7) / 8, 1));
    PGGC_HIP(rocprim::radix_sort_pairs((void*)tmp.p, tb, kd, vd, (size_t)n, 0u, end_bit, c.stream));
    if (kd.current() != keys) copy_dev(c, keys, kd.current(), n * sizeof(u64));
    if (vd.current() != vals) copy_dev(c, vals, vd.current(), n * sizeof(V));
}
void radix_sort_pairs(Ctx& c, u64* keys, i32* vals, i64 n, int bits) { radix_sort_pairs_t(c, keys, vals, n, bits); }
void radix_sort_pairs(Ctx& c, u64* keys, i64* vals, i64 n, int bits) { radix_sort_pairs_t(c, keys, vals, n, bits); }

void radix_sort_keys(Ctx& c, u64* keys, i64 n, int bits) {
    if (n <= 1 || bits <= 0) return;
    Buf<u64> k2(c, n);
    rocprim::double_buffer<u64> kd(keys, k2.p);
    const unsigned int end_bit = (unsigned int)std::min(bits, 64);
    size_t tb = 0;
    PGGC_HIP(rocprim::radix_sort_keys(nullptr, tb, kd, (size_t)n, 0u, end_bit, c.stream));
    Buf<u64> tmp(c, std::max<size_t>((tb + 7) / 8, 1));
    PGGC_HIP(rocprim::radix_sort_keys((void*)tmp.p, tb, kd, (size_t)n, 0u, end_bit, c.stream));
    if (kd.current() != keys) copy_dev(c, keys, kd.current(), n * sizeof(u64));
}

// ------------------------------------------------------------ exact reductions
__global__ void __launch_bounds__(256) k_sum_i64(const i64* in, i64 n, i64* out) {
    i64 s = 0;
    // block-uniform loop control (the reduction below is cross-lane)
    for (i64 base = (i64)blockIdx.x * blockDim.x; base < n; base += (i64)gridDim.x * blockDim.x) {
        const i64 x = base + threadIdx.x;
        if (x < n) s += in[x];
    }
    s = block_sum_i64<256>(s);
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)out, (unsigned long long)s);
}

__global__ void k_acct_add(u64* dst, const i64* src) {
    if (threadIdx.x == 0) *dst += (u64)*src;
}
void acct_add(Ctx& c, u64* dst, const i64* src) {
    if (!dst) return;
    k_acct_add<<<1, 64, 0, c.stream>>>(dst, src);
    PGGC_LAUNCH_CHECK();
}

void sum_i64(Ctx& c, const i64* in, i64 n, i64* out) {
    fill_i64(c, out, 1, 0);
    if (n <= 0) return;
    k_sum_i64<<<grid_for(n, 256 * 8, 2048), 256, 0, c.stream>>>(in, n, out);
    PGGC_LAUNCH_CHECK();
}

struct U128 {
    u64 lo, hi;
};
__device__ __forceinline__ U128 add128(U128 a, U128 b) {
    U128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0);
    return r;
}

__global__ void __launch_bounds__(256) k_sumsq(const i64* in, i64 n, U128* part) {
    U128 s{0, 0};
    for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x) {
        i64 v = in[x];
        u64 a = (u64)(v < 0 ? -v : v);
        U128 q{a * a, __umul64hi(a, a)};
        s = add128(s, q);
    }
    __shared__ U128 sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = add128(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_sumsq_final(const U128* part, int np, i64* out) {
    U128 s{0, 0};
    for (int x = threadIdx.x; x < np; x += blockDim.x) s = add128(s, part[x]);
    __shared__ U128 sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = add128(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = (i64)sh[0].hi;
        out[1] = (i64)sh[0].lo;
    }
}

void sum_sq_u128(Ctx& c, const i64* in, i64 n, i64* out) {
    int g = grid_for(std::max<i64>(n, 1), 256 * 8, 1024);
    Buf<U128> part(c, g);
    k_sumsq<<<g, 256, 0, c.stream>>>(in, n, part);
    PGGC_LAUNCH_CHECK();
    k_sumsq_final<<<1, 256, 0, c.stream>>>(part, g, out);
    PGGC_LAUNCH_CHECK();
}

}  // namespace pggc
