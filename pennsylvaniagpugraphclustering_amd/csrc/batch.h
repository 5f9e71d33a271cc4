// Batched pLouvain over a disjoint union of graphs (batch.hip).
#pragma once
#include <vector>

#include "localmove.h"

namespace pggc {

struct BatchInfo {
    i64 levels = 0, total_nnz = 0;
    MoveStats mv;
};

// louvain_part (src/clustering_methods.cpp:139-222, unconstrained) of every
// graph of gs at once; init[k] is instance k's initial objective state
// (get_objective), outs[k] (device, gs[k].n entries) receives its
// clustering, final_rd[k] its final state.
void louvain_batch(Ctx& c, MoveScratch& ms, const std::vector<DGraph>& gs, const std::vector<RefineData>& init,
                   const std::vector<i32*>& outs, std::vector<RefineData>& final_rd, std::vector<BatchInfo>& info);

}  // namespace pggc
