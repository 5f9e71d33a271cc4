// Internal vertex order of the top level for the local mover: vertices
// renumbered by descending degree (ties by id), so that the records the
// sweeps gather at random neighbour ids (cluster, gain, destination, weight)
// of the high-degree vertices, which most edges point at, sit together in
// few cache lines.  Results are those of the caller's order, bit for bit:
//  - every row keeps its entries in the caller's order, so each per-vertex
//    reduction (the afterburner's FP32 sum) adds the same terms in the same
//    order;
//  - labels stay in the caller's id space (vertex i' starts in cluster
//    ids[i'] when the clustering starts at singletons);
//  - the afterburner's priority hash is taken of the caller's id;
//  - new cluster ids go to the NEW_PART movers in the caller's move order
//    (order class, then caller's id);
//  - the degree classes, and so every kernel's work split, are unchanged.
#pragma once
#include "graph.h"
#include "localmove.h"

namespace pggc {

struct Renum {
    const void* key = nullptr;  // row offsets of the graph it was built for
    i64 n = 0;
    Buf<i64> rp, vw;
    Buf<i32> col, w32, ids, part;
    Buf<i64> w64;
    DGraph g;       // the renumbered graph
    LevelPlan plan;
    Orders ord;     // orderings of the caller's graph (for the refinement), built on first use
    bool ord_built = false;
    bool built_for(const DGraph& top) const { return key == (const void*)top.rp && n == top.n; }
};

// Builds r for top (no-op when r already holds it).
void renum_build(Ctx& c, const DGraph& top, Renum& r);

// out[i'] = in[ids[i']]: a per-vertex array of the caller's order in the
// renumbered order
void renum_gather(Ctx& c, const Renum& r, const i32* in, i32* out);

// local_move on the renumbered graph: part (caller's order) is read and
// written in place; cons is indexed by label, so it is shared unchanged.
void local_move_renum(Ctx& c, MoveScratch& s, Renum& r, i32* part, RefineData& best, bool is_initial,
                      bool constrained, i32* cons, bool SA, MoveStats* st);

}  // namespace pggc
