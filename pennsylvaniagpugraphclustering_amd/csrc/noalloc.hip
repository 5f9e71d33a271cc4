// Allocation-free modularity evaluator: pggc_evaluate for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  Everything it needs beyond the caller's
// inputs and outputs lives in one file-scope device array, `scratch`:
//
//   words [0, NA_MAX_CHUNKS)           uncut partial sum of each edge chunk
//   words [NA_MAX_CHUNKS, +2*NA_SQB)   sum-of-squares partials (lo, hi)
//   words [.., +3)                     results: uncut, sum(vol^2) high, low
//
// Every partial and result word is written before it is read in the same
// call, so nothing in the array needs clearing; the caller's vol_out is
// zeroed by a fill kernel.  Each GPU holds its own copy of the array until
// the process ends, and every call on that GPU shares it: calls that overlap
// on one GPU (two streams, two host threads) race on it.
//
// The host never names the array directly (its name is a host-side stand-in,
// not a device address): hipGetSymbolAddress resolves it once per device at
// context creation, and kernels receive that pointer.
#include <mutex>
#include <string>

#include "ctx.h"
#include "device.h"
#include "noalloc.h"

namespace pggc {

constexpr int NA_EB = 256;                      // threads per edge chunk
constexpr int NA_EPT = 16;                      // consecutive edges per thread
constexpr i64 NA_EC = (i64)NA_EB * NA_EPT;      // edges per chunk
constexpr i64 NA_MAX_CHUNKS = (i64)1 << 20;     // nnz <= 2^32
constexpr int NA_SQB = 1024;                    // sum-of-squares blocks
constexpr i64 NA_RES = NA_MAX_CHUNKS + 2 * NA_SQB;
constexpr i64 NA_WORDS = NA_RES + 4;

static __device__ __align__(256) u64 scratch[NA_WORDS];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static u64* g_na_ptr[NA_MAX_DEVICES] = {};

i64 noalloc_max_nnz() { return NA_MAX_CHUNKS * NA_EC; }
i64 noalloc_scratch_bytes() { return NA_WORDS * (i64)sizeof(u64); }

void noalloc_init(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) throw Error(-1, "device index outside the scratch table");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return;
    void* p = nullptr;
    PGGC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(-3, "hipGetSymbolAddress returned a null scratch address");
    g_na_ptr[device] = (u64*)p;
}

static u64* scratch_ptr(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) throw Error(-1, "device index outside the scratch table");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) throw Error(-5, "scratch of this device was not resolved at context creation");
    return g_na_ptr[device];
}

// Uncut weight of one chunk of NA_EC edges: each thread walks NA_EPT
// consecutive edges with a monotone row pointer; the block's sum goes to
// part[chunk] (every block writes its word, zero included).
template <class WT>
__global__ void __launch_bounds__(NA_EB) k_na_uncut(const i64* __restrict__ rp, const i32* __restrict__ col, WT W,
                                                    const i32* __restrict__ lab, i64 n, i64 nnz, i64 chunks,
                                                    u64* __restrict__ part) {
    __shared__ i64 s_lo, s_hi;
    const i64 b = block_id();
    if (b >= chunks) return;  // block-uniform
    const i64 eb = b * NA_EC;
    const i64 ee = eb + NA_EC < nnz ? eb + NA_EC : nnz;
    if (threadIdx.x == 0) {
        s_lo = upper_row(rp, 0, n, eb);
        s_hi = upper_row(rp, 0, n, ee - 1) + 1;
    }
    __syncthreads();
    const i64 rlo = s_lo, rhi = s_hi;
    const i64 j0 = eb + (i64)threadIdx.x * NA_EPT;
    i64 acc = 0;
    if (j0 < ee) {
        i64 r = upper_row(rp, rlo, rhi, j0);
        i64 nxt = rp[r + 1];
        i32 lr = lab[r];
        for (int k = 0; k < NA_EPT; k++) {
            const i64 j = j0 + k;
            if (j >= ee) break;
            if (j >= nxt) {
                r = advance_row(rp, r, rhi, j, nxt);
                lr = lab[r];
            }
            if (lab[col[j]] == lr) acc += W(j);
        }
    }
    acc = block_sum_i64<NA_EB>(acc);  // LDS tree
    if (threadIdx.x == 0) part[b] = (u64)acc;
}

__global__ void __launch_bounds__(256) k_na_vol(const i32* __restrict__ lab, const i64* __restrict__ vw, i64 n,
                                                i64* vol) {
    const i64 i = thread_item();
    if (i >= n) return;
    atomicAdd((unsigned long long*)&vol[lab[i]], (unsigned long long)vw[i]);
}

struct NaU128 {
    u64 lo, hi;
};
__device__ __forceinline__ NaU128 na_add(NaU128 a, NaU128 b) {
    NaU128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0);
    return r;
}

// exact 128-bit sum of vol[l]^2 per block: sq[2b] = low word, sq[2b+1] = high
__global__ void __launch_bounds__(256) k_na_sumsq(const i64* __restrict__ vol, i64 L, u64* __restrict__ sq) {
    NaU128 s{0, 0};
    for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < L; x += (i64)gridDim.x * blockDim.x) {
        const i64 v = vol[x];
        const u64 a = (u64)(v < 0 ? -v : v);
        s = na_add(s, NaU128{a * a, __umul64hi(a, a)});
    }
    __shared__ NaU128 sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = na_add(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sq[2 * blockIdx.x] = sh[0].lo;
        sq[2 * blockIdx.x + 1] = sh[0].hi;
    }
}

// res[0] = sum of the chunk partials, res[1] / res[2] = high / low word of
// the sum-of-squares partials
__global__ void __launch_bounds__(1024) k_na_final(const u64* __restrict__ part, i64 chunks,
                                                   const u64* __restrict__ sq, int nsq, u64* __restrict__ res) {
    i64 acc = 0;
    for (i64 x = threadIdx.x; x < chunks; x += 1024) acc += (i64)part[x];
    NaU128 s{0, 0};
    for (int x = threadIdx.x; x < nsq; x += 1024) s = na_add(s, NaU128{sq[2 * x], sq[2 * x + 1]});
    acc = block_sum_i64<1024>(acc);
    __shared__ NaU128 sh[1024];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = na_add(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        res[0] = (u64)acc;
        res[1] = sh[0].hi;
        res[2] = sh[0].lo;
    }
}

void evaluate_noalloc(Ctx& c, const DGraph& g, const i32* labels, i64 L, i64* vol, i64* uncut_out, u64* sq_out) {
    if (g.nnz > noalloc_max_nnz())
        throw Error(-4, "pggc_evaluate_noalloc: nnz " + std::to_string(g.nnz) + " exceeds the scratch limit of " +
                            std::to_string(noalloc_max_nnz()) + " nonzeros");
    if (!g.vw && g.n > 0) throw Error(-1, "vertex weights required");
    if (g.n > 0 && (L < 1 || !labels || !vol)) throw Error(-1, "labels, L >= 1 and vol_out are required");
    if (L < 0) throw Error(-1, "L must be >= 0");
    u64* s = scratch_ptr(c.device);
    u64* part = s;
    u64* sq = s + NA_MAX_CHUNKS;
    u64* res = s + NA_RES;
    const i64 chunks = ceil_div(g.nnz, NA_EC);
    fill_i64(c, vol, L, 0);
    if (chunks > 0) {
        PGGC_WDISPATCH(g, (k_na_uncut<<<grid_blocks(chunks), NA_EB, 0, c.stream>>>(g.rp, g.col, W, labels, g.n,
                                                                                   g.nnz, chunks, part)));
        PGGC_LAUNCH_CHECK();
    }
    if (g.n > 0) {
        k_na_vol<<<grid_items(g.n, 256), 256, 0, c.stream>>>(labels, g.vw, g.n, vol);
        PGGC_LAUNCH_CHECK();
    }
    const int nsq = grid_for(std::max<i64>(L, 1), 256 * 8, NA_SQB);
    k_na_sumsq<<<nsq, 256, 0, c.stream>>>(vol, L, sq);
    PGGC_LAUNCH_CHECK();
    k_na_final<<<1, 1024, 0, c.stream>>>(part, chunks, sq, nsq, res);
    PGGC_LAUNCH_CHECK();
    const i64* h = c.read((const i64*)res, 3);
    *uncut_out = h[0];
    sq_out[0] = (u64)h[1];
    sq_out[1] = (u64)h[2];
}

}  // namespace pggc
