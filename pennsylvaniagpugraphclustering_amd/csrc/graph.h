// Level-graph operations: degree ordering, evaluator, own-cluster weights,
// contraction, and the device-resident coarse graph type.
#pragma once
#include "ctx.h"

namespace pggc {

// A graph level that owns its device arrays (coarse levels).
struct OwnedGraph {
    Buf<i64> rp;
    Buf<i32> col;
    Buf<i64> w;    // int64 coarse weights, or
    Buf<i32> w32;  // the same weights in 32 bits when every one fits (narrow_weights)
    Buf<i64> vw;
    i64 n = 0, nnz = 0;
    DGraph view() const {
        DGraph g;
        g.n = n;
        g.nnz = nnz;
        g.rp = rp.p;
        g.col = col.p;
        if (w32.p) {
            g.w = w32.p;
            g.wk = W_I32;
        } else {
            g.w = w.p;
            g.wk = W_I64;
        }
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

// Replaces g's int64 coarse weights by int32 ones when every weight fits:
// always when the level it was contracted from has unit weights and fewer
// than 2^31 nonzeros (a coarse weight then counts fine edges), else after a
// max reduction.  The values, and so every result, are unchanged; the
// sweeps read 4 instead of 8 weight bytes per edge and the level holds 8
// instead of 12 bytes per nonzero.
void narrow_weights(Ctx& c, const DGraph& fine, OwnedGraph& g);

// next.vw[map[i]] += g.vw[i]  (src/clustering_methods.cpp:33-38)
void coarsen_vertex_weights(Ctx& c, const DGraph& g, const i32* map, i64 nc, i64* out);

}  // namespace pggc
