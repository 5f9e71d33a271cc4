// Device context: stream, caching allocator, pinned readback slots, and the
// primitive operations (scan, compaction, radix sort, reductions) shared by
// the hot-path kernels.
#pragma once
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace pggc {

// Kernel timing slots: HIP events recorded on the context stream around the
// launches of the kernels named here, with the algorithmic bytes each launch
// moves (measurement hooks; off unless enabled).
enum KSlot {
    KS_SCORE_GROUP = 0,  // k_score_group<G>
    KS_SCORE_TABLE,      // k_score_table / k_score_raw (LDS tables, block arg-max)
    KS_SCORE_HUB,        // hub scoring: table build (k_hub_insert), k_hub_pv, k_hub_scan, k_hub_decide
    KS_AFTERBURNER,      // k_ab_pack + k_ab_serial / k_ab_wave / k_ab_block
    KS_MOVES,            // hub-table updates + k_delta_uncut + k_apply_moves
    KS_CONTRACT_AGG,     // k_cagg (LDS merge of coarse rows)
    KS_CONTRACT_BIG,     // k_cbig_insert + k_cbig_extract
    KS_CONTRACT_COPY,    // k_crow_copy
    KS_EVAL,             // k_eval_uncut
    KS_LEIDEN,           // leidenR refinement (all its kernels)
    KS_ORDER,            // ordering pass (classes + two stable partitions)
    KS_COUNT
};

struct KTimer {
    struct Pending {
        int slot;
        hipEvent_t a, b;
        double bytes;
        bool stopped;
    };
    bool on = false;
    int open = 0;  // regions started and not yet stopped
    std::vector<Pending> pend;
    u64* dcnt = nullptr;  // per-region device byte counters (kernels add to dcnt[token])
    std::vector<hipEvent_t> pool;
    i64 launches[KS_COUNT] = {0};
    double ms[KS_COUNT] = {0};
    double bytes[KS_COUNT] = {0};
};

struct Ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    std::string last_error;
    // caching allocator: blocks are kept for reuse until the context ends
    std::multimap<size_t, void*> free_blocks;
    std::unordered_map<void*, size_t> live;
    size_t bytes_live = 0, bytes_cached = 0;
    i64* pinned = nullptr;      // 64 pinned int64 readback slots
    i64* pinned_dev = nullptr;  // the same slots as seen from the device
    void* stage = nullptr;      // pinned coherent staging area for uploads
    void* stage_dev = nullptr;
    size_t stage_cap = 0;
    u64 seed = 0;           // leidenR seed stream
    u64 draws = 0;
    int verbose = 0;        // >0: per-iteration progress lines on stderr; >=2: memory by stage
    // per-outer-iteration callback (iteration, seconds, cut, objective, labels)
    void (*progress)(void*, int64_t, double, int64_t, double, int32_t) = nullptr;
    void* progress_user = nullptr;
    // per-sweep trace of the local mover (level n, nnz, phase, sweep, active,
    // candidates, moves, microseconds); synchronises per sweep when set
    void (*sweep_trace)(void*, const int64_t*, int) = nullptr;
    void* sweep_user = nullptr;
    double last_contract_s = 0.0, last_refine_s = 0.0;  // of the last clustering call
    // tuning knobs (pggc_ctx_set_knob): hub tables are updated from the
    // movers while a sweep moves at most n / hub_update_div vertices and
    // rebuilt otherwise (0: always rebuilt)
    i64 hub_update_div = 32;
    // Execution-shape knobs of the scoring kernels (results do not depend on
    // them; pggc_autotune picks them by timing):
    //   table_threads: threads per vertex of the LDS-table classes with
    //     64 < degree <= 512 (64, 128 or 256);
    //   table_spe: LDS hash-table slots per edge of a class's degree bound
    //     (2 or 4), classes 64 < degree <= 2048;
    //   hub_lds_max: vertices of degree 2049..hub_lds_max are scored from an
    //     8192-slot LDS table instead of an HBM hub table (2048 or 4096);
    //   hub_edges_item / hub_slots_item: edges per hub-table build item
    //     (512, 1024 or 2048) and table slots per hub scan item (2048, 4096
    //     or 8192).
    int table_threads = 64;
    int table_spe = 2;
    i64 hub_lds_max = 2048;
    i64 hub_edges_item = 1024;
    i64 hub_slots_item = 4096;
    // the first sweep of a call from singletons scores the raw adjacency
    // (every neighbour its own cluster); 0 scores it with the tables
    int raw_sweeps = 1;
    // hub tables: 0 builds each table slice in LDS and writes it out whole,
    // 1 adds edge work items into the HBM tables with atomics
    int hub_build = 0;
    // the top level's local mover runs on a copy renumbered by degree:
    // 0 never, 1 always, 2 when a call makes several top-level passes
    int renumber = 2;
    // big-row contraction batches: table slots per batch (sort entries per
    // batch: a quarter of it)
    i64 contract_batch = (i64)1 << 30;
    u64 knob_epoch = 0;  // bumped by every knob change; level plans are rebuilt when it moves
    // the active-set predictor runs on at most this many sweeps of a
    // local-move call (0: on every eligible sweep)
    i64 gnn_max_sweeps = 0;
    KTimer kt;

    void* alloc(size_t bytes);
    // verbose >= 2: a line on stderr with the bytes this context holds
    void mem_note(const char* where) const;
    void release(void* p);
    void trim();
    void sync() { PGGC_HIP(hipStreamSynchronize(stream)); }
    // copy `count` int64 device values to the pinned slots and wait
    const i64* read(const i64* dev, int count);
    i64 read1(const i64* dev) { return read(dev, 1)[0]; }
    u32 next_draw() { return seed_draw(seed, draws++); }
    // kernel timing: kt_start/kt_stop bracket launches on `stream`;
    // kt_resolve waits for the stream and folds the pending pairs into the
    // per-slot totals.
    hipEvent_t kt_event();
    int kt_start(int slot);
    void kt_stop(int token, double bytes);
    void kt_resolve();
    void kt_free();
    // timed region on `stream`: token (or -1 when timing is off) ...
    int kt_begin(int slot) { return kt.on ? kt_start(slot) : -1; }
    // ... closed with the algorithmic bytes it moved (added to what the
    // kernels counted in kt_counter(tok))
    void kt_end(int tok, double bytes) {
        if (tok >= 0) kt_stop(tok, bytes);
    }
    u64* kt_counter(int tok) { return tok >= 0 ? kt.dcnt + tok : nullptr; }
};

// RAII device buffer from the context's caching allocator.
template <class T>
struct Buf {
    Ctx* c = nullptr;
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(Ctx& ctx, size_t count) : c(&ctx), n(count) { p = (T*)ctx.alloc(count * sizeof(T) + 16); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept : c(o.c), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    Buf& operator=(Buf&& o) noexcept {
        if (this != &o) {
            reset();
            c = o.c; p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~Buf() { reset(); }
    void reset() {
        if (p && c) c->release(p);
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
    operator T*() const { return p; }
};

// ---- primitives (prims.hip) ----
void fill_i32(Ctx& c, i32* p, i64 n, i32 v);
void fill_i64(Ctx& c, i64* p, i64 n, i64 v);
void iota_i32(Ctx& c, i32* p, i64 n);
void copy_dev(Ctx& c, void* dst, const void* src, size_t bytes);  // on c.stream
// host -> device copy, done when it returns: the bytes go through a pinned
// coherent staging area and a copy kernel on the stream, so later kernels
// read them through the same cache path as data written by kernels
void copy_h2d(Ctx& c, void* dst, const void* src, size_t bytes);
// out[0..n] = exclusive scan of in[0..n); out[n] = total.  in may alias out.
void scan_i64(Ctx& c, const i64* in, i64* out, i64 n);
void scan_i32_to_i64(Ctx& c, const i32* in, i64* out, i64 n);
// Ordered compaction of the indices x in [0,n) with flag[x] != 0, writing
// src ? src[x] : x; returns the count (read back to the host).
i64 compact_flags(Ctx& c, const i32* flag, const i32* src, i64 n, i32* out, i64* count_dev);
// Stable LSD radix sort of (key, val) pairs on the low `bits` key bits.
void radix_sort_pairs(Ctx& c, u64* keys, i32* vals, i64 n, int bits);
void radix_sort_pairs(Ctx& c, u64* keys, i64* vals, i64 n, int bits);
void radix_sort_keys(Ctx& c, u64* keys, i64 n, int bits);
// *dst += *src on the stream (algorithmic-byte counters of timed regions)
void acct_add(Ctx& c, u64* dst, const i64* src);
// sum of int64 values (device result in *out)
void sum_i64(Ctx& c, const i64* in, i64 n, i64* out);
// exact sum of squares of int64 values: out[0] = high word, out[1] = low word
void sum_sq_u128(Ctx& c, const i64* in, i64 n, i64* out);

}  // namespace pggc
