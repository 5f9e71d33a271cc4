// Local mover (src/local_mover.cpp) on one level graph.
#pragma once
#include <vector>

#include "graph.h"

namespace pggc {

// Objective state: the cluster_data mirror (interface/cluster_data.hpp:18-88).
struct RefineData {
    i64* vol = nullptr;   // total_deg, device, capacity `cap` labels
    i64 cap = 0;
    i64 uncut = 0;
    i64 top_nnz = 0;
    double obj = -1.0;
    i32 label_count = 0;
    double lambda = 1.0;
    double scale = 1.0;   // get_objective() = obj * scale + offset
    double offset = 0.0;
    double objective() const { return obj * scale + offset; }
};

// obj = uncut - lambda * sum(vol^2) with the exact 128-bit sum of squares
// given as (hi, lo) words (interface/cluster_data.hpp:76-87).
double objective_value(i64 uncut, double lambda, u64 sq_hi, u64 sq_lo);
// recompute rd.obj from the device volumes
void update_objective(Ctx& c, RefineData& rd);

struct MoveStats {
    i64 sweeps = 0, moves = 0, cands = 0;
};

// Scoring classes by degree: 0..4 lane groups (d <= 4, 8, 16, 32, 64),
// CL_TABLE..CL_BIG-1 one wave per vertex with an LDS table sized for the class
// (d <= 128, 256, 512), CL_BIG..CL_HUB-1 one block per vertex (256 threads for
// d <= 1024, 512 for d <= 2048), CL_HUB hubs.
constexpr int NCLS = 11;
constexpr int CL_TABLE = 5, CL_BIG = 8, CL_HUB = 10;

// Per-level cached work partitioning (reused by every local_move on a level).
struct LevelPlan {
    const void* key = nullptr;  // graph identity (row offsets pointer)
    i64 n = -1;
    Orders ord;
    Buf<i32> sc_list;           // vertices grouped by scoring class
    i64 sc_start[NCLS + 1] = {0};
    // hubs (largest degree class): HBM hash tables of neighbour-cluster
    // weights (own cluster included), edge and slot work items per hub.
    // A table is built at the first table sweep of a local-move call and
    // then kept current by the movers of each sweep (weight moved from the
    // old to the new cluster key), so later sweeps only scan it; a table that
    // runs out of slots marks the set stale and the next sweep rebuilds it.
    // packed: one u64 per slot (key << 32 | weight), used when every hub's
    // row weight fits 32 bits; otherwise separate key / int64 weight arrays.
    // A table is split into slices of at most hub_slice(packed) slots: a key
    // probes within the slice its hash falls in, so a build holds one slice
    // in LDS per block (soff: slice prefix per hub) and writes it out whole.
    // A slice that cannot hold its keys (or the hub_build knob) sends its
    // hub to the fallback build, edge items adding into the table with
    // atomics and probing over the whole table: hflag[h] = 1, and hsm[h]
    // (the slice mask every table user probes with) becomes the table mask.
    // imap / smap / bmap: block -> build item, scan item or slice, so that
    // all work of one hub runs on one XCD (blocks are dispatched round-robin
    // over the 8 XCDs) and its table stays in that XCD's L2; -1 = idle block.
    // hidx: vertex -> hub index, -1 for other vertices.
    Buf<i64> goff, items, sitems, soff, hub_pv;
    Buf<u64> hub_best;
    Buf<i32> gkeys;
    Buf<i64> gvals;
    Buf<u64> gtab;
    Buf<i32> imap, smap, bmap, hidx, hflag, hovf;
    Buf<u32> hsm;
    // occupied-slot list per table (slot indices from goff[h] on, hcnt[h]
    // of them, in no particular order): scans read occupied slots only
    Buf<u32> hlist;
    Buf<i32> hcnt;
    bool packed = false;
    bool hub_live = false;  // tables match the current clustering
    i64 gslots = 0, nitems = 0, nsitems = 0, nslices = 0, igrid = 0, sgrid = 0, bgrid = 0;
    i64 hs1 = 0, hs2 = 0;  // slice-class bounds (xcd_maps)
    i64 hg = 0, ghg = 0, ihg = 0, shg = 0;  // first fallback-built hub, its first slot, build item and slice
    // the hub class is sorted by degree; its first nlds vertices (degree <=
    // Ctx::hub_lds_max) are scored in LDS, the rest (from hub0 on) in HBM
    // tables; he / hs: edges per build item and slots per scan item
    u64 epoch = ~(u64)0;
    i64 nlds = 0, hub0 = 0, nhub = 0;
    i64 he = 1024, hs = 4096;
};

// Device scratch sized for the top level, shared by every level.
struct Gnn;

// What the afterburner reads of a neighbour, in one 16-byte load: its gain,
// destination and cluster of this sweep and its vertex weight as the fp32
// value the gain arithmetic uses.
struct __align__(16) AbInfo {
    float gain;
    i32 dest;
    i32 part;
    float vwf;
};

struct MoveScratch {
    i64 cap = 0;
    Gnn* gnn = nullptr;  // active-set predictor (null or mode off: plain sweeps)
    Buf<AbInfo> abi;
    Buf<u64> mvp;  // per vertex: cluster | moved << 31, and the cluster after this sweep's moves
    Buf<i32> part, dest, mflag, cand, moves, flag, tmp, tmp2;
    Buf<float> gain;
    Buf<i64> prio, scal;
    void ensure(Ctx& c, i64 n);
};

void plan_level(Ctx& c, const DGraph& g, LevelPlan& p);

// local_move<constrained> (src/local_mover.cpp:1360-1417) and
// local_move_strict (:1419-1483).  best_part and cons are device arrays.
// ids (optional, device, n): g is a renumbering of the caller's graph and
// ids[i] is vertex i's caller id, which then stands in for i wherever the
// result depends on vertex ids (afterburner priorities, new cluster ids).
void local_move(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i32* best_part, RefineData& best,
                bool is_initial, bool constrained, i32* cons, bool SA, MoveStats* st, const i32* ids = nullptr);
void local_move_strict(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, i32* best_part, RefineData& best,
                       bool is_initial, bool constrained, i32* cons, MoveStats* st);

// ---------------------------------------------------------------- batches
// Many independent graphs held as one disjoint union graph: instance k owns
// the contiguous vertex range [voff[k], voff[k+1]) of every level, and the
// label range [loff[k], loff[k] + n0[k]) at every level (n0 = its top-level
// vertex count), so its labels are its single-run labels plus loff[k].
// Kernels run once over the union; every per-instance decision of the
// reference's control flow (sweep phases, early stops, best-state tracking,
// new cluster ids, relabelling) is taken per instance, so each instance
// ends exactly as it would alone.
struct BatchLevel {
    i64 K = 0;
    std::vector<i64> voff;  // K+1 vertex offsets at this level (host)
    Buf<i64> dvoff;
    Buf<i32> inst;  // instance of every vertex
    Buf<i32> loc;   // vertex id within its instance
    void build(Ctx& c, const std::vector<i64>& offsets);
};
struct BatchState {
    i64 K = 0;
    std::vector<i64> loff;  // K+1 label offsets (host)
    Buf<i64> dloff;
    Buf<i32> linst;          // instance owning every label
    Buf<float> dlam;         // float lambda per instance
    Buf<i64> vol;            // volumes over the whole label space
    std::vector<RefineData> rd;  // per-instance objective state (rd[k].vol unused)
};

// local_move (src/local_mover.cpp:1360-1417, SA on, unconstrained) of every
// instance k with active[k] at once; best_part holds labels (label space).
void local_move_batch(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, const BatchLevel& bl,
                      i32* best_part, BatchState& best, bool is_initial, const std::vector<char>& active,
                      std::vector<MoveStats>& stats);

// One scoring + afterburner sweep from a fixed state (parity entry point).
// Returns the number of accepted moves written to moves_out (order1 order).
i64 sweep_once(Ctx& c, MoveScratch& s, LevelPlan& lp, const DGraph& g, const i32* part, const i64* vol, i64 L,
               double lambda, float ratio, bool raw, const i32* cons, i32* dest, float* gain, i32* moves_out,
               i64* ncand);

}  // namespace pggc
