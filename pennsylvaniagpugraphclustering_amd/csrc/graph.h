// Level-graph operations: degree ordering, evaluator, own-cluster weights,
// contraction, and the device-resident coarse graph type.
#pragma once
#include "ctx.h"

namespace pggc {

// A graph level that owns its device arrays (coarse levels).
struct OwnedGraph {
    Buf<i64> rp;
    Buf<i32> col;
    Buf<i64> w;   // int64 coarse weights
    Buf<i64> vw;
    i64 n = 0, nnz = 0;
    DGraph view() const {
        DGraph g;
        g.n = n;
        g.nnz = nnz;
        g.rp = rp.p;
        g.col = col.p;
        g.w = w.p;
        g.wk = W_I64;
        g.vw = vw.p;
        return g;
    }
};

// Stable degree bucketing (src/ordering.cpp:131-164).
struct Orders {
    Buf<i32> o1, o2;
    i64 off_large = 0, off_large2 = 0, off_mid = 0, off_mid2 = 0;
};
void make_orders(Ctx& c, const DGraph& g, Orders& o);

// Stable K-way partition of [0,n) by a per-vertex class (class array given);
// out receives the vertex ids, starts[0..K] the class boundaries (host).
void partition_classes(Ctx& c, const i32* cls, i64 n, int K, i32* out, i64* starts);

// Evaluator (app/verify_objective.cpp:64-89): vol[l] += vw[i], uncut.
// vol must hold L entries; result scalars are left in dev_out[0] (uncut).
void evaluate(Ctx& c, const DGraph& g, const i32* labels, i64 L, i64* vol, i64* dev_uncut);

// out[i] = sum of w(i,j) over neighbours j with labels[j] == labels[i].
void own_weight(Ctx& c, const DGraph& g, const i32* labels, i64* out);

// Contraction (src/contract.cpp:234-336) with sorted coarse rows.
void contract(Ctx& c, const DGraph& g, const i32* vcmap, i64 nc, OwnedGraph& out);

// next.vw[map[i]] += g.vw[i]  (src/clustering_methods.cpp:33-38)
void coarsen_vertex_weights(Ctx& c, const DGraph& g, const i32* map, i64 nc, i64* out);

}  // namespace pggc
