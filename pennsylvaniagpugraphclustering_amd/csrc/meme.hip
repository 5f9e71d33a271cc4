// Building blocks of the memetic loop (app/memetic_clustering.cpp):
// intersection of two clusterings, the cut-edge signature of a clustering and
// the Hamming distance of two signatures.
#include <algorithm>

#include "ctx.h"
#include "device.h"
#include "meme.h"

namespace pggc {

// gen_signature (app/memetic_clustering.cpp:71-106): bit j of the signature is
// set when edge j (CSR position) joins two clusters.  One 32-edge word per
// thread, rows found by binary search and then walked.
__global__ void __launch_bounds__(256) k_cut_signature(const i64* __restrict__ rp, const i32* __restrict__ col,
                                                       i64 n, i64 nnz, const i32* __restrict__ part, u32* sig,
                                                       i64 words) {
    const i64 w = thread_item();
    if (w >= words) return;
    const i64 j0 = w * 32;
    const i64 j1 = j0 + 32 < nnz ? j0 + 32 : nnz;
    i64 r = upper_row(rp, 0, n, j0);
    i64 nxt = rp[r + 1];
    i32 pr = part[r];
    u32 bits = 0;
    for (i64 j = j0; j < j1; j++) {
        if (j >= nxt) {
            r = advance_row(rp, r, n, j, nxt);
            pr = part[r];
        }
        bits |= (u32)(pr != part[col[j]]) << (u32)(j - j0);
    }
    sig[w] = bits;
}

// get_cut_diff (:108-116): popcount of the XOR
__global__ void __launch_bounds__(256) k_signature_diff(const u32* __restrict__ a, const u32* __restrict__ b,
                                                        i64 words, i64* out) {
    const i64 w = thread_item();
    const i64 x = w < words ? (i64)__popc(a[w] ^ b[w]) : 0;
    const i64 s = block_sum_i64<256>(x);
    if (threadIdx.x == 0 && s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
}

void cut_signature(Ctx& c, const DGraph& g, const i32* part, u32* sig) {
    const i64 words = (g.nnz + 31) / 32;
    if (words == 0) return;
    k_cut_signature<<<grid_items(words, 256), 256, 0, c.stream>>>(g.rp, g.col, g.n, g.nnz, part, sig, words);
    PGGC_LAUNCH_CHECK();
}

i64 signature_diff(Ctx& c, const u32* a, const u32* b, i64 words) {
    Buf<i64> out(c, 1);
    fill_i64(c, out, 1, 0);
    if (words > 0) {
        k_signature_diff<<<grid_items(words, 256), 256, 0, c.stream>>>(a, b, words, out);
        PGGC_LAUNCH_CHECK();
    }
    return c.read1(out);
}

// intersection_cluster (:27-62): vertices share a label iff they share their
// labels in both inputs.  The reference numbers the (a, b) pairs in the order
// of their hash slots; here they are numbered in increasing (a, b) order, the
// same partition under other ids.
__global__ void k_pair_keys(const i32* c1, const i32* c2, i64 n, i64 l2, u64* key, i32* val) {
    const i64 i = thread_item();
    if (i >= n) return;
    key[i] = (u64)c1[i] * (u64)l2 + (u64)c2[i];
    val[i] = (i32)i;
}
__global__ void k_key_starts(const u64* key, i64 n, i32* flag) {
    const i64 x = thread_item();
    if (x >= n) return;
    flag[x] = (x == 0 || key[x] != key[x - 1]) ? 1 : 0;
}
__global__ void k_pair_labels(const i64* rank, const i32* flag, const i32* val, i64 n, i32* out) {
    const i64 x = thread_item();
    if (x >= n) return;
    out[val[x]] = (i32)(rank[x] + flag[x] - 1);
}

i64 intersect_clusterings(Ctx& c, i64 n, const i32* c1, const i32* c2, i64 l2, i32* out) {
    if (n == 0) return 0;
    Buf<u64> key(c, n);
    Buf<i32> val(c, n), flag(c, n);
    Buf<i64> rank(c, n + 1);
    k_pair_keys<<<grid_items(n, 256), 256, 0, c.stream>>>(c1, c2, n, l2, key, val);
    PGGC_LAUNCH_CHECK();
    // keys < (max label of c1 + 1) * l2 <= 2^62
    int bits = 1;
    const u64 top = (u64)(n + 1) * (u64)std::max<i64>(l2, 1);
    while (bits < 63 && ((u64)1 << bits) <= top) bits++;
    radix_sort_pairs(c, key, val, n, bits);
    k_key_starts<<<grid_items(n, 256), 256, 0, c.stream>>>(key, n, flag);
    PGGC_LAUNCH_CHECK();
    scan_i32_to_i64(c, flag, rank, n);
    k_pair_labels<<<grid_items(n, 256), 256, 0, c.stream>>>(rank, flag, val, n, out);
    PGGC_LAUNCH_CHECK();
    return c.read1(rank.p + n);
}

}  // namespace pggc
