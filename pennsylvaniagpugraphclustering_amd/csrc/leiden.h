// Leiden refinement of one level (src/leidenR.cpp:263-500).
#pragma once
#include "graph.h"

namespace pggc {

// Refines the Louvain clustering `louv` (volumes `vol`, own-cluster weights
// `pv`) into gamma-connected sub-clusters.  Draws two seeds from the context
// seed stream.  Writes contiguous labels to vcmap_out and returns their count.
i64 leiden_refine(Ctx& c, const DGraph& g, const Orders& ord, const i32* louv, const i64* vol, double lambda,
                  const i64* pv, bool uniform, i32* vcmap_out);

}  // namespace pggc
