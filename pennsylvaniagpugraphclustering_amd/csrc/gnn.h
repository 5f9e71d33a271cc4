// Active-set predictor for the local mover: a two-layer graph network that
// estimates, before a sweep, which active vertices will move in it.  It has
// no counterpart in the reference; it hooks into the reference's own active
// set, the cached NO_MOVE destination that makes a sweep skip a vertex until
// a neighbour moves (src/local_mover.cpp:477-479,717,798), by caching NO_MOVE
// for active vertices the network predicts to stay.  Every vertex it lets
// through is still scored by the exact gain, and a sweep that finds nothing
// to move is repeated with every pruned vertex restored, so a local-move
// call that stops for lack of moves stops under the plain method's
// condition (a call that stops at its sweep limit leaves pruned vertices to
// the next call, which starts with every vertex active).
//
// Model (hidden width 32):
//   per-vertex pack  pk_i = part_i | ldeg_q | margin_q | moved | active
//   input row        in_i = 8 own features + 8 neighbour means  (16 x bf16)
//   layer 1          h_i  = bf16(ReLU(in_i W1 + b1))             (MFMA bf16, fp32 accumulate)
//   layer 2          logit_i = h_i.w2s + b2 + mean_{j in N(i)} h_j.w2n
// Training: weighted binary cross-entropy on "moved in this sweep" over the
// active vertices of sweeps recorded from the plain method, Adam updates.
#pragma once
#include "common.h"
#include "ctx.h"

namespace pggc {

struct LevelPlan;

constexpr int GNN_IN = 16, GNN_HID = 32;
constexpr int GP_W1 = 0;                         // W1[16][32], row = input feature
constexpr int GP_B1 = GP_W1 + GNN_IN * GNN_HID;  // b1[32]
constexpr int GP_W2S = GP_B1 + GNN_HID;          // w2s[32] (own hidden state)
constexpr int GP_W2N = GP_W2S + GNN_HID;         // w2n[32] (neighbour mean)
constexpr int GP_B2 = GP_W2N + GNN_HID;          // b2
constexpr int GNN_P = GP_B2 + 1;                 // 609 parameters

enum GnnMode { GNN_OFF = 0, GNN_TRAIN = 1, GNN_PRUNE = 2 };

// device counter slots
enum GnnCount {
    GC_ACTIVE = 0,    // active vertices seen by the predictor (all sweeps)
    GC_PRUNED = 1,    // prune decisions (all sweeps)
    GC_RESTORED = 2,  // vertices put back by the last restore
    GC_TP = 3, GC_FP = 4, GC_FN = 5, GC_TN = 6,  // training sweeps: prediction (logit >= tau) vs moved
    GC_SWEEP_ACTIVE = 7,  // active vertices of the current sweep
    GC_TMP = 8,
    GC_COUNT = 9
};

struct GnnConfig {
    int mode = GNN_OFF;
    float tau = 0.0f;          // prune when logit < tau
    i64 min_nnz = 1 << 20;     // levels with fewer nonzeros run plain
    int full_every = 0;        // every k-th sweep of a call restores and scores all (0: never)
    float lr = 1e-3f;
    float pos_weight = 4.0f;   // weight of the "moved" class in the loss
};

struct Gnn {
    GnnConfig cfg;
    bool have_params = false;
    i64 step = 0;                     // Adam steps taken
    i64 sweeps = 0, fallbacks = 0, train_steps = 0;
    Buf<float> prm, m1, m2, grad;     // fp32 parameters, Adam moments, last gradient
    Buf<u16> w1b;                     // bf16 copy of W1
    Buf<i64> dcnt;                    // GnnCount slots
    Buf<double> dloss;                // summed training loss
    Buf<float> part;                  // per-block gradient partials
    // per-vertex state, sized for the largest level seen
    i64 cap = 0;
    bool hcap = false;
    Buf<u64> pk;
    Buf<u16> in, h;
    Buf<float> s, t, logit, r, ds;
    Buf<i32> stamp, pruned;  // sweep of the last move / of the prune decision (0: none)
    Buf<i64> blk;            // per-block partial counts

    void init(Ctx& c);
    void ensure(Ctx& c, i64 n, bool train);
    void set_params(Ctx& c, const float* host);  // also resets Adam
    void reset_counters(Ctx& c);
};

// Predictor pass before a sweep.  Computes the features from the state
// (part, dest, gain, stamps, vol), runs both layers for the active vertices
// and, when `prune`, caches NO_MOVE for those with logit < tau.  `cur` is the
// sweep number within the call (stamps equal to cur-1 mark last sweep's
// movers); `first` zeroes the gain feature (gains are not yet computed).
void gnn_predict(Ctx& c, Gnn& G, LevelPlan& lp, const DGraph& g, const i32* part, i32* dest, const float* gain,
                 const i64* vol, i32 cur, bool first, bool prune, bool keep_hidden);
// start of a local-move call on a level of n vertices: no vertex has moved
// or been pruned yet
void gnn_begin(Ctx& c, Gnn& G, i64 n, bool train);
// marks the accepted moves of sweep `cur`
void gnn_stamp(Ctx& c, Gnn& G, const i32* moves, i64 k, i32 cur);
// restores pruned vertices to the active set; returns false if none waited
// (reads one count back from the device)
bool gnn_restore(Ctx& c, Gnn& G, i64 n, i32* dest);
// loss and gradient of the last gnn_predict (run with keep_hidden) against
// labels (nullable: stamps equal to cur), then an Adam step when `update`
void gnn_backward(Ctx& c, Gnn& G, LevelPlan& lp, const DGraph& g, const i32* labels, i32 cur, bool update);

}  // namespace pggc
