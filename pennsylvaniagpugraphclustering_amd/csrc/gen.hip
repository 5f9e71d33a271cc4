// Device-side graph construction: the symmetric CSR build of an edge list
// (construct_from_edgelist, app/io/io_mtx.cpp:95-201: self loops and zero
// weights dropped, undirected duplicates merged, both directions stored,
// rows sorted) and the counter-based generators behind the synthetic
// workloads (R-MAT edge draws, per-pair weights, random labellings).
//
// Build: every undirected pair becomes the key (lo << B) | hi with B =
// ceil(log2 n); a radix sort of the keys groups duplicates and orders the
// pairs by (lo, hi), and a second sort of the transposed keys (hi << B) | lo
// orders them by (hi, lo).  With firstF[u] / firstT[u] the positions of the
// first pair with lo >= u / hi >= u, row u of the CSR is the transposed run
// of u (columns < u, ascending) followed by the forward run of u (columns > u,
// ascending), so rp[u] = firstF[u] + firstT[u], and pair k lands at
// firstT[lo + 1] + k (forward) and firstF[hi] + k (transposed) -- no atomics.
#include <algorithm>
#include <cmath>

#include "ctx.h"
#include "device.h"
#include "gen.h"

namespace pggc {

static int key_bits(i64 n) {
    int b = 1;
    while (((i64)1 << b) < n) b++;
    return b;
}

// ---------------------------------------------------------------- draws
// R-MAT quadrant thresholds on a u32 draw (Chakrabarti, Zhan, Faloutsos,
// SDM'04): r < t0 -> (0,0), < t1 -> (0,1), < t2 -> (1,0), else (1,1).
void rmat_thresholds(double a, double b, double c, u64 t[3]) {
    t[0] = (u64)std::floor(a * 4294967296.0);
    t[1] = (u64)std::floor((a + b) * 4294967296.0);
    t[2] = (u64)std::floor((a + b + c) * 4294967296.0);
}

// Edge draw e: one Philox block per 4 levels (counter = e lo, e hi, block,
// stream 0; key = seed), level l (most significant bit first) uses word l % 4.
__global__ void __launch_bounds__(256) k_rmat_keys(i64 m, int scale, u64 t0, u64 t1, u64 t2, u64 seed, int B,
                                                   u64 invalid, u64* __restrict__ keys) {
    const i64 e = thread_item();
    if (e >= m) return;
    u32 s = 0, d = 0;
    for (int blk = 0; blk * 4 < scale; blk++) {
        const Philox4 r = philox4x32_10((u32)e, (u32)((u64)e >> 32), (u32)blk, STREAM_RMAT, (u32)seed,
                                        (u32)(seed >> 32));
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (blk * 4 + q < scale) {
                const u64 x = r.v[q];
                s = (s << 1) | (x >= t1 ? 1u : 0u);
                d = (d << 1) | (((x >= t0) && (x < t1)) || (x >= t2) ? 1u : 0u);
            }
        }
    }
    const u32 lo = s < d ? s : d, hi = s < d ? d : s;
    keys[e] = lo == hi ? invalid : ((u64)lo << B) | hi;
}

// keys of an edge list; ids outside [0, n) or negative weights raise *bad
__global__ void __launch_bounds__(256) k_edge_keys(i64 m, const i32* __restrict__ src, const i32* __restrict__ dst,
                                                   const i32* __restrict__ w, i64 n, int B, u64 invalid,
                                                   u64* __restrict__ keys, i32* __restrict__ wk, i64* bad) {
    const i64 j = thread_item();
    if (j >= m) return;
    const i32 u = src[j], v = dst[j];
    const i32 x = w ? w[j] : 1;
    if (u < 0 || v < 0 || (i64)u >= n || (i64)v >= n || x < 0) {
        atomicOr((unsigned long long*)bad, 1ull);
        keys[j] = invalid;
        if (wk) wk[j] = 0;
        return;
    }
    const u32 lo = (u32)(u < v ? u : v), hi = (u32)(u < v ? v : u);
    keys[j] = (lo == hi || x == 0) ? invalid : ((u64)lo << B) | hi;
    if (wk) wk[j] = x;
}

// weight of the undirected pair lo < hi: uniform in [wmin, wmax]
__device__ __forceinline__ i32 pair_weight(u32 lo, u32 hi, u64 seed, i32 wmin, u32 span) {
    const Philox4 r = philox4x32_10(lo, hi, 0u, STREAM_WEIGHT, (u32)seed, (u32)(seed >> 32));
    return wmin + (i32)(r.v[0] % span);
}

__global__ void __launch_bounds__(256) k_random_labels(i64 n, i64 L, u64 seed, i32* __restrict__ out) {
    const i64 i = thread_item();
    if (i >= n) return;
    const Philox4 r = philox4x32_10((u32)i, (u32)((u64)i >> 32), 0u, STREAM_LABEL, (u32)seed, (u32)(seed >> 32));
    out[i] = (i32)(r.v[0] % (u32)L);
}

// ------------------------------------------------------------ unique pairs
// Keys are sorted; a key survives when it is valid and differs from its
// predecessor (the first of a run, so with a stable sort the first
// occurrence's weight is kept).
constexpr int UT = 256, UI = 8, UTILE = UT * UI;

__device__ __forceinline__ bool uniq_flag(const u64* keys, i64 x, u64 invalid) {
    const u64 k = keys[x];
    return k != invalid && (x == 0 || keys[x - 1] != k);
}

__global__ void __launch_bounds__(UT) k_uniq_sums(const u64* __restrict__ keys, i64 m, u64 invalid, i64* bsum) {
    const i64 base = block_id() * UTILE;
    i64 s = 0;
    for (int k = 0; k < UI; k++) {
        const i64 x = base + (i64)k * UT + threadIdx.x;
        if (x < m && uniq_flag(keys, x, invalid)) s++;
    }
    s = block_sum_i64<UT>(s);
    if (threadIdx.x == 0) bsum[block_id()] = s;
}

__global__ void __launch_bounds__(UT) k_uniq_write(const u64* __restrict__ keys, const i32* __restrict__ wk, i64 m,
                                                   u64 invalid, const i64* __restrict__ boff,
                                                   u64* __restrict__ okeys, i32* __restrict__ ow) {
    const i64 base = block_id() * UTILE + (i64)threadIdx.x * UI;
    int f[UI];
    i64 s = 0;
#pragma unroll
    for (int k = 0; k < UI; k++) {
        const i64 x = base + k;
        f[k] = (x < m && uniq_flag(keys, x, invalid)) ? 1 : 0;
        s += f[k];
    }
    i64 tot;
    i64 pre = block_excl_scan_i64(s, &tot) + boff[block_id()];
#pragma unroll
    for (int k = 0; k < UI; k++) {
        if (f[k]) {
            okeys[pre] = keys[base + k];
            if (ow) ow[pre] = wk[base + k];
            pre++;
        }
    }
}

// unique valid keys (and their weights) of sorted keys[m]; returns the count
static i64 unique_pairs(Ctx& c, const u64* keys, const i32* wk, i64 m, u64 invalid, u64* okeys, i32* ow) {
    if (m <= 0) return 0;
    const i64 nb = ceil_div(m, UTILE);
    Buf<i64> bs(c, nb + 1);
    k_uniq_sums<<<grid_blocks(nb), UT, 0, c.stream>>>(keys, m, invalid, bs);
    PGGC_LAUNCH_CHECK();
    scan_i64(c, bs, bs, nb);
    k_uniq_write<<<grid_blocks(nb), UT, 0, c.stream>>>(keys, wk, m, invalid, bs, okeys, ow);
    PGGC_LAUNCH_CHECK();
    return c.read1(bs.p + nb);
}

// ------------------------------------------------------------------ CSR fill
__global__ void __launch_bounds__(256) k_transpose(i64 M, const u64* __restrict__ keys, int B, u64* __restrict__ kt) {
    const i64 k = thread_item();
    if (k >= M) return;
    const u64 x = keys[k];
    const u64 mask = ((u64)1 << B) - 1;
    kt[k] = ((x & mask) << B) | (x >> B);
}

// first[u] = number of keys below u << B, for u in [0, n]
__global__ void __launch_bounds__(256) k_first(i64 n, const u64* __restrict__ keys, i64 M, int B, i64* first) {
    const i64 u = thread_item();
    if (u > n) return;
    const u64 t = (u64)u << B;
    i64 lo = 0, hi = M;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if (keys[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    first[u] = lo;
}

__global__ void __launch_bounds__(256) k_rowptr(i64 n, const i64* __restrict__ fF, const i64* __restrict__ fT,
                                                i64* __restrict__ rp) {
    const i64 u = thread_item();
    if (u > n) return;
    rp[u] = fF[u] + fT[u];
}

// wmode 0: unit (no weights), 1: carried weights wk, 2: generated per pair
template <int WMODE, bool FWD>
__global__ void __launch_bounds__(256) k_fill(i64 M, const u64* __restrict__ keys, const i32* __restrict__ wk, int B,
                                              const i64* __restrict__ other_first, i32* __restrict__ col,
                                              i32* __restrict__ w, u64 seed, i32 wmin, u32 span) {
    const i64 k = thread_item();
    if (k >= M) return;
    const u64 x = keys[k];
    const u64 mask = ((u64)1 << B) - 1;
    const u32 row = (u32)(x >> B), nb = (u32)(x & mask);
    // forward pairs (row = lo): after the row's transposed run; transposed
    // pairs (row = hi): at the row's start
    const i64 pos = (FWD ? other_first[(i64)row + 1] : other_first[row]) + k;
    col[pos] = (i32)nb;
    if (WMODE == 1) w[pos] = wk[k];
    if (WMODE == 2) w[pos] = FWD ? pair_weight(row, nb, seed, wmin, span) : pair_weight(nb, row, seed, wmin, span);
}

// Sorted unique pairs (and weights) -> symmetric CSR in out.
static void pairs_to_csr(Ctx& c, i64 n, u64* fkeys, i32* fw, i64 M, int B, int wmode, u64 seed, i32 wmin,
                         i32 wmax, DevCSR& out) {
    out.n = n;
    out.nnz = 2 * M;
    PGGC_HIP(hipMalloc(&out.rp, (size_t)(n + 1) * sizeof(i64)));
    PGGC_HIP(hipMalloc(&out.col, (size_t)std::max<i64>(out.nnz, 1) * sizeof(i32)));
    if (wmode != 0) PGGC_HIP(hipMalloc(&out.w, (size_t)std::max<i64>(out.nnz, 1) * sizeof(i32)));
    Buf<i64> fF(c, n + 1), fT(c, n + 1);
    Buf<u64> tkeys(c, std::max<i64>(M, 1));
    Buf<i32> tw;
    k_transpose<<<grid_items(M, 256), 256, 0, c.stream>>>(M, fkeys, B, tkeys);
    PGGC_LAUNCH_CHECK();
    if (wmode == 1) {
        tw = Buf<i32>(c, std::max<i64>(M, 1));
        copy_dev(c, tw.p, fw, (size_t)M * sizeof(i32));
        radix_sort_pairs(c, tkeys, tw, M, 2 * B);
    } else {
        radix_sort_keys(c, tkeys, M, 2 * B);
    }
    k_first<<<grid_items(n + 1, 256), 256, 0, c.stream>>>(n, fkeys, M, B, fF);
    k_first<<<grid_items(n + 1, 256), 256, 0, c.stream>>>(n, tkeys, M, B, fT);
    k_rowptr<<<grid_items(n + 1, 256), 256, 0, c.stream>>>(n, fF, fT, out.rp);
    PGGC_LAUNCH_CHECK();
    const u32 span = wmode == 2 ? (u32)(wmax - wmin + 1) : 1u;
    const dim3 g = grid_items(M, 256);
    if (M > 0) {
        if (wmode == 0) {
            k_fill<0, true><<<g, 256, 0, c.stream>>>(M, fkeys, nullptr, B, fT, out.col, nullptr, seed, wmin, span);
            k_fill<0, false><<<g, 256, 0, c.stream>>>(M, tkeys, nullptr, B, fF, out.col, nullptr, seed, wmin, span);
        } else if (wmode == 1) {
            k_fill<1, true><<<g, 256, 0, c.stream>>>(M, fkeys, fw, B, fT, out.col, out.w, seed, wmin, span);
            k_fill<1, false><<<g, 256, 0, c.stream>>>(M, tkeys, tw, B, fF, out.col, out.w, seed, wmin, span);
        } else {
            k_fill<2, true><<<g, 256, 0, c.stream>>>(M, fkeys, nullptr, B, fT, out.col, out.w, seed, wmin, span);
            k_fill<2, false><<<g, 256, 0, c.stream>>>(M, tkeys, nullptr, B, fF, out.col, out.w, seed, wmin, span);
        }
        PGGC_LAUNCH_CHECK();
    }
    c.sync();
}

void DevCSR::release() {
    if (rp) (void)hipFree(rp);
    if (col) (void)hipFree(col);
    if (w) (void)hipFree(w);
    rp = nullptr;
    col = nullptr;
    w = nullptr;
}

void rmat_csr(Ctx& c, int scale, int edge_factor, double a, double b, double cc, u64 seed, i32 wmin, i32 wmax,
              DevCSR& out) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || edge_factor > 1024)
        throw Error(-1, "R-MAT: scale must be in [1, 30] and edge_factor in [1, 1024]");
    if (!(a >= 0 && b >= 0 && cc >= 0 && a + b + cc <= 1.0)) throw Error(-1, "R-MAT: bad quadrant probabilities");
    if (wmax > 0 && (wmin < 1 || wmin > wmax)) throw Error(-1, "R-MAT: weights need 1 <= wmin <= wmax");
    const i64 n = (i64)1 << scale;
    const i64 m = (i64)edge_factor << scale;
    const int B = scale;
    const u64 invalid = 2 * B >= 64 ? ~(u64)0 : (((u64)1 << (2 * B)) - 1);
    u64 t[3];
    rmat_thresholds(a, b, cc, t);
    i64 M = 0;
    Buf<u64> ukeys;
    {
        Buf<u64> keys(c, m);
        k_rmat_keys<<<grid_items(m, 256), 256, 0, c.stream>>>(m, scale, t[0], t[1], t[2], seed, B, invalid, keys);
        PGGC_LAUNCH_CHECK();
        radix_sort_keys(c, keys, m, 2 * B);
        ukeys = Buf<u64>(c, m);
        M = unique_pairs(c, keys, nullptr, m, invalid, ukeys, nullptr);
    }
    pairs_to_csr(c, n, ukeys, nullptr, M, B, wmax > 0 ? 2 : 0, seed, wmin, wmax, out);
    ukeys.reset();
    c.trim();
}

void edges_csr(Ctx& c, i64 n, i64 m, const i32* src, const i32* dst, const i32* w, DevCSR& out) {
    if (n < 0 || n >= 0x7fffffff || m < 0) throw Error(-1, "edge list: need 0 <= n < 2^31 - 1 and m >= 0");
    if (m > 0 && (!src || !dst)) throw Error(-1, "edge list: null src/dst");
    const int B = key_bits(std::max<i64>(n, 2));
    const u64 invalid = ((u64)1 << (2 * B)) - 1;
    i64 M = 0;
    Buf<u64> ukeys;
    Buf<i32> uw;
    {
        Buf<u64> keys(c, std::max<i64>(m, 1));
        Buf<i32> wk;
        if (w) wk = Buf<i32>(c, std::max<i64>(m, 1));
        Buf<i64> bad(c, 1);
        fill_i64(c, bad, 1, 0);
        if (m > 0) {
            k_edge_keys<<<grid_items(m, 256), 256, 0, c.stream>>>(m, src, dst, w, n, B, invalid, keys, wk.p, bad);
            PGGC_LAUNCH_CHECK();
        }
        if (c.read1(bad.p) != 0) throw Error(-1, "edge list: vertex id outside [0, n) or negative weight");
        if (w) radix_sort_pairs(c, keys, wk, m, 2 * B);
        else radix_sort_keys(c, keys, m, 2 * B);
        ukeys = Buf<u64>(c, std::max<i64>(m, 1));
        if (w) uw = Buf<i32>(c, std::max<i64>(m, 1));
        M = unique_pairs(c, keys, wk.p, m, invalid, ukeys, uw.p);
    }
    pairs_to_csr(c, n, ukeys, uw.p, M, B, w ? 1 : 0, 0, 0, 0, out);
    ukeys.reset();
    uw.reset();
    c.trim();
}

// ------------------------------------------------------------------ LFR
// Power-law tables are built on the host in double precision (the CPU
// restatement builds the same tables with the same operations); the device
// only compares integer draws against them, so CPU and GPU draws agree bit
// for bit.
static void pl_probs(double tau, double x0, i32 top, std::vector<double>& p, i32& k0) {
    k0 = (i32)std::floor(x0);
    const double a = std::pow(x0, 1.0 - tau);
    const double b = std::pow((double)top + 1.0, 1.0 - tau);
    const double z = a - b;
    p.assign((size_t)(top - k0 + 1), 0.0);
    double prev = a;  // x^(1-tau) at the lower end of the current bin
    for (i32 k = k0; k <= top; k++) {
        const double nxt = std::pow((double)k + 1.0, 1.0 - tau);
        p[(size_t)(k - k0)] = (prev - nxt) / z;
        prev = nxt;
    }
}

static double pl_mean(double tau, double x0, i32 top) {
    std::vector<double> p;
    i32 k0;
    pl_probs(tau, x0, top, p, k0);
    double m = 0.0;
    for (size_t i = 0; i < p.size(); i++) m += (double)(k0 + (i32)i) * p[i];
    return m;
}

PowerLawTable power_law_table(double tau, i32 lo, i32 top, double mean) {
    if (!(tau > 0.0) || std::fabs(tau - 1.0) < 1e-6 || lo < 1 || top < lo)
        throw Error(-1, "power law: need tau > 0, tau != 1 and 1 <= lo <= top");
    double x0 = (double)lo;
    if (mean > 0.0) {
        double a = (double)lo, b = (double)top;
        if (mean < pl_mean(tau, a, top) || mean > pl_mean(tau, b, top))
            throw Error(-1, "power law: the requested mean is outside the reachable range");
        for (int it = 0; it < 200; it++) {  // the mean grows with the lower end
            const double mid = 0.5 * (a + b);
            if (pl_mean(tau, mid, top) < mean) a = mid;
            else b = mid;
        }
        x0 = 0.5 * (a + b);
    }
    std::vector<double> p;
    PowerLawTable t;
    pl_probs(tau, x0, top, p, t.k0);
    t.thr.resize(p.size());
    double cum = 0.0;
    for (size_t i = 0; i < p.size(); i++) {
        cum += p[i];
        double f = std::floor(cum * 4294967296.0);
        if (f < 0.0) f = 0.0;
        if (f > 4294967296.0) f = 4294967296.0;
        t.thr[i] = (u64)f;
    }
    t.thr.back() = (u64)1 << 32;
    return t;
}

// first i with r < thr[i] (thr non-decreasing, thr[K-1] = 2^32)
__host__ __device__ __forceinline__ i32 pl_index(const u64* thr, i32 K, u32 r) {
    i32 lo = 0, hi = K - 1;
    while (lo < hi) {
        const i32 m = (lo + hi) >> 1;
        if ((u64)r < thr[m]) hi = m;
        else lo = m + 1;
    }
    return lo;
}

std::vector<i64> lfr_community_sizes(const LfrParams& p) {
    const PowerLawTable t = power_law_table(p.tau2, p.cmin, p.cmax, 0.0);
    std::vector<i64> sz;
    i64 sum = 0;
    for (u64 j = 0; sum < p.n; j++) {
        const Philox4 r = philox4x32_10((u32)j, (u32)(j >> 32), 0u, STREAM_LFR_COMM, (u32)p.seed,
                                        (u32)(p.seed >> 32));
        const i64 s = t.k0 + pl_index(t.thr.data(), (i32)t.thr.size(), r.v[0]);
        sz.push_back(s);
        sum += s;
    }
    if (!sz.empty()) {
        // the last community takes what is left; a remainder below cmin joins
        // the community before it
        const i64 rem = p.n - (sum - sz.back());
        sz.back() = rem;
        if (rem < p.cmin && sz.size() > 1) {
            sz[sz.size() - 2] += rem;
            sz.pop_back();
        }
    }
    return sz;
}

// Vertex v lies in community c with cs[c] <= v < cs[c+1].  Degree from the
// degree table; internal degree round((1 - mu) * degree) (mu in thousandths),
// at most the community size - 1; the rest of the degree is external.
__global__ void __launch_bounds__(256) k_lfr_degrees(i64 n, i64 nc, const i64* __restrict__ cs,
                                                     const u64* __restrict__ thr, i32 K, i32 k0, i64 mu_m, u64 seed,
                                                     i32* __restrict__ kin, i32* __restrict__ kext,
                                                     i32* __restrict__ comm) {
    const i64 v = thread_item();
    if (v >= n) return;
    const i64 c = upper_row(cs, 0, nc, v);
    const i64 size = cs[c + 1] - cs[c];
    const Philox4 r = philox4x32_10((u32)v, (u32)((u64)v >> 32), 0u, STREAM_LFR_DEG, (u32)seed, (u32)(seed >> 32));
    const i64 deg = k0 + pl_index(thr, K, r.v[0]);
    i64 ki = (deg * (1000 - mu_m) + 500) / 1000;
    if (ki > size - 1) ki = size - 1;
    kin[v] = (i32)ki;
    kext[v] = (i32)(deg - ki);
    comm[v] = (i32)c;
}

// stub p of the vertex v with off[v] <= p < off[v+1]: key = (community << 32
// when WITH_COMM) | Philox draw of p
template <bool WITH_COMM>
__global__ void __launch_bounds__(256) k_lfr_stubs(i64 S, i64 n, const i64* __restrict__ off,
                                                   const i32* __restrict__ comm, u64 seed, u32 stream,
                                                   u64* __restrict__ keys, i32* __restrict__ vals) {
    const i64 p = thread_item();
    if (p >= S) return;
    const i64 v = upper_row(off, 0, n, p);
    const Philox4 r = philox4x32_10((u32)p, (u32)((u64)p >> 32), 0u, stream, (u32)seed, (u32)(seed >> 32));
    keys[p] = WITH_COMM ? (((u64)(u32)comm[v] << 32) | r.v[0]) : (u64)r.v[0];
    vals[p] = (i32)v;
}

__device__ __forceinline__ u64 lfr_pair_key(const i32* pi, i32 a, i32 b, int B, u64 invalid) {
    const u32 u = (u32)pi[a], w = (u32)pi[b];
    if (u == w) return invalid;
    const u32 lo = u < w ? u : w, hi = u < w ? w : u;
    return ((u64)lo << B) | hi;
}

// internal stubs sorted by (community, draw): the stubs of community c fill
// [oin[cs[c]], oin[cs[c+1]]), paired (first, second), (third, fourth), ...
__global__ void __launch_bounds__(256) k_lfr_pair_int(i64 S, const u64* __restrict__ keys, const i32* __restrict__ vals,
                                                      const i64* __restrict__ oin, const i64* __restrict__ cs,
                                                      const i32* __restrict__ pi, int B, u64 invalid,
                                                      u64* __restrict__ ek) {
    const i64 p = thread_item();
    if (p >= S) return;
    const i64 c = (i64)(keys[p] >> 32);
    const i64 s0 = oin[cs[c]], len = oin[cs[c + 1]] - s0;
    const i64 q = p - s0;
    ek[p] = ((q & 1) == 0 && q + 1 < len) ? lfr_pair_key(pi, vals[p], vals[p + 1], B, invalid) : invalid;
}

// external stubs sorted by draw, paired consecutively over the whole graph
__global__ void __launch_bounds__(256) k_lfr_pair_ext(i64 S, const i32* __restrict__ vals,
                                                      const i32* __restrict__ pi, int B, u64 invalid,
                                                      u64* __restrict__ ek) {
    const i64 p = thread_item();
    if (p >= S) return;
    ek[p] = ((p & 1) == 0 && p + 1 < S) ? lfr_pair_key(pi, vals[p], vals[p + 1], B, invalid) : invalid;
}

__global__ void __launch_bounds__(256) k_lfr_perm_keys(i64 n, u64 seed, u64* __restrict__ keys,
                                                       i32* __restrict__ vals) {
    const i64 v = thread_item();
    if (v >= n) return;
    const Philox4 r = philox4x32_10((u32)v, (u32)((u64)v >> 32), 0u, STREAM_LFR_PERM, (u32)seed, (u32)(seed >> 32));
    keys[v] = r.v[0];
    vals[v] = (i32)v;
}
__global__ void __launch_bounds__(256) k_rank_scatter(i64 n, const i32* __restrict__ order, i32* __restrict__ pi) {
    const i64 r = thread_item();
    if (r >= n) return;
    pi[order[r]] = (i32)r;
}
__global__ void __launch_bounds__(256) k_lfr_gt(i64 n, const i32* __restrict__ pi, const i32* __restrict__ comm,
                                                i32* __restrict__ gt) {
    const i64 v = thread_item();
    if (v >= n) return;
    gt[pi[v]] = comm[v];
}

void lfr_csr(Ctx& c, const LfrParams& p, DevCSR& out, i32* gt) {
    if (p.n < 1 || p.n >= 0x7fffffff) throw Error(-1, "LFR: need 1 <= n < 2^31 - 1");
    if (p.kmax < 1 || p.cmin < 1 || p.cmax < p.cmin) throw Error(-1, "LFR: need kmax >= 1 and 1 <= cmin <= cmax");
    if (!(p.mu >= 0.0 && p.mu <= 1.0)) throw Error(-1, "LFR: mixing mu must be in [0, 1]");
    if (p.wmax > 0 && (p.wmin < 1 || p.wmin > p.wmax)) throw Error(-1, "LFR: weights need 1 <= wmin <= wmax");
    const PowerLawTable dt = power_law_table(p.tau1, 1, p.kmax, p.kmean);
    const std::vector<i64> sz = lfr_community_sizes(p);
    const i64 n = p.n, nc = (i64)sz.size();
    std::vector<i64> cs(nc + 1, 0);
    for (i64 k = 0; k < nc; k++) cs[k + 1] = cs[k] + sz[k];
    const i64 mu_m = std::llround(p.mu * 1000.0);
    const i32 K = (i32)dt.thr.size();
    Buf<i64> dcs(c, nc + 1);
    Buf<u64> dthr(c, K);
    copy_h2d(c, dcs.p, cs.data(), (nc + 1) * sizeof(i64));
    copy_h2d(c, dthr.p, dt.thr.data(), K * sizeof(u64));
    Buf<i32> kin(c, n), kext(c, n), comm(c, n);
    k_lfr_degrees<<<grid_items(n, 256), 256, 0, c.stream>>>(n, nc, dcs, dthr, K, dt.k0, mu_m, p.seed, kin, kext, comm);
    PGGC_LAUNCH_CHECK();
    Buf<i64> oin(c, n + 1), oext(c, n + 1);
    scan_i32_to_i64(c, kin, oin, n);
    scan_i32_to_i64(c, kext, oext, n);
    const i64 Sin = c.read1(oin.p + n), Sext = c.read1(oext.p + n);
    Buf<i32> pi(c, n);
    {
        Buf<u64> pk(c, n);
        Buf<i32> pv(c, n);
        k_lfr_perm_keys<<<grid_items(n, 256), 256, 0, c.stream>>>(n, p.seed, pk, pv);
        PGGC_LAUNCH_CHECK();
        radix_sort_pairs(c, pk, pv, n, 32);
        k_rank_scatter<<<grid_items(n, 256), 256, 0, c.stream>>>(n, pv, pi);
        PGGC_LAUNCH_CHECK();
    }
    if (gt) {
        k_lfr_gt<<<grid_items(n, 256), 256, 0, c.stream>>>(n, pi, comm, gt);
        PGGC_LAUNCH_CHECK();
    }
    const int B = key_bits(std::max<i64>(n, 2));
    const u64 invalid = ((u64)1 << (2 * B)) - 1;
    const i64 M = Sin + Sext;
    Buf<u64> ek(c, std::max<i64>(M, 1));
    if (Sin > 0) {
        Buf<u64> sk(c, Sin);
        Buf<i32> sv(c, Sin);
        k_lfr_stubs<true><<<grid_items(Sin, 256), 256, 0, c.stream>>>(Sin, n, oin, comm, p.seed, STREAM_LFR_ISTUB,
                                                                       sk, sv);
        PGGC_LAUNCH_CHECK();
        radix_sort_pairs(c, sk, sv, Sin, 32 + key_bits(std::max<i64>(nc, 2)));
        k_lfr_pair_int<<<grid_items(Sin, 256), 256, 0, c.stream>>>(Sin, sk, sv, oin, dcs, pi, B, invalid, ek);
        PGGC_LAUNCH_CHECK();
    }
    if (Sext > 0) {
        Buf<u64> sk(c, Sext);
        Buf<i32> sv(c, Sext);
        k_lfr_stubs<false><<<grid_items(Sext, 256), 256, 0, c.stream>>>(Sext, n, oext, comm, p.seed,
                                                                        STREAM_LFR_ESTUB, sk, sv);
        PGGC_LAUNCH_CHECK();
        radix_sort_pairs(c, sk, sv, Sext, 32);
        k_lfr_pair_ext<<<grid_items(Sext, 256), 256, 0, c.stream>>>(Sext, sv, pi, B, invalid, ek.p + Sin);
        PGGC_LAUNCH_CHECK();
    }
    kin.reset();
    kext.reset();
    comm.reset();
    i64 U = 0;
    Buf<u64> uk(c, std::max<i64>(M, 1));
    if (M > 0) {
        radix_sort_keys(c, ek, M, 2 * B);
        U = unique_pairs(c, ek, nullptr, M, invalid, uk, nullptr);
    }
    ek.reset();
    pairs_to_csr(c, n, uk, nullptr, U, B, p.wmax > 0 ? 2 : 0, p.seed, p.wmin, p.wmax, out);
    uk.reset();
    c.trim();
}

void random_labels(Ctx& c, i64 n, i64 L, u64 seed, i32* out) {
    if (n < 0 || L < 1 || L > 0x7fffffff) throw Error(-1, "random labels: need n >= 0 and 1 <= L < 2^31");
    if (n == 0) return;
    k_random_labels<<<grid_items(n, 256), 256, 0, c.stream>>>(n, L, seed, out);
    PGGC_LAUNCH_CHECK();
}

}  // namespace pggc
