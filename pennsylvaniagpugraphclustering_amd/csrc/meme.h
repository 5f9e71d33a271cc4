// Building blocks of the memetic loop (app/memetic_clustering.cpp:27-116).
#pragma once
#include "ctx.h"

namespace pggc {

// bit j of sig[(nnz+31)/32] = edge j joins two clusters of part
void cut_signature(Ctx& c, const DGraph& g, const i32* part, u32* sig);
// number of differing bits of two signatures
i64 signature_diff(Ctx& c, const u32* a, const u32* b, i64 words);
// labels of the intersection of two clusterings (c2 has l2 labels); returns
// the label count
i64 intersect_clusterings(Ctx& c, i64 n, const i32* c1, const i32* c2, i64 l2, i32* out);

}  // namespace pggc
