/*
 * pggc.h -- C ABI of the MI355X-native PGGC Louvain/Leiden hot path.
 *
 * Replaces the reference's C++ template boundary
 *   jet_community::clustering_methods::leiden_part / louvain_part
 *   (interface/clustering_methods.h:20-24)
 * and its internal seams (src/header/local_mover.h:21-25,
 * src/header/contract.h:20-24, src/header/leidenR.h:26-31,
 * src/header/ordering.h:21), with plain pointers and sizes.
 *
 * Conventions
 *  - Every array argument is DEVICE memory (HBM of the context's GPU) unless
 *    its comment says "host".
 *  - Every function returns PGGC_OK (0) or a negative status; the message of
 *    the last failure is available from pggc_last_error().
 *  - A context owns one HIP stream and a caching device allocator.  It is not
 *    thread-safe; different contexts are independent.
 *  - Graphs are symmetric CSR without self loops and without repeated
 *    columns in a row (as the reference's readers produce them,
 *    app/io/io_mtx.cpp:71-93,127-129): row_ptr int64[n+1], col int32[nnz], at
 *    most one of w (int32[nnz]) / w64 (int64[nnz]) set (neither = unit
 *    weights), vtx_w int64[n].
 *  - Results are deterministic for a given seed.
 */
#ifndef PGGC_H
#define PGGC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGGC_OK 0
#define PGGC_EINVAL -1   /* bad argument */
#define PGGC_ENOMEM -2   /* device allocation failed */
#define PGGC_EHIP -3     /* HIP runtime error */
#define PGGC_ECAPACITY -4/* output or label capacity exceeded */
#define PGGC_EINTERNAL -5

/* objectives (app/io/parse_args.cpp:10-16) */
#define PGGC_OBJ_MOD 0   /* modularity, unit edge weights, degree vertex weights */
#define PGGC_OBJ_WMOD 1  /* weighted modularity, weighted-degree vertex weights */
#define PGGC_OBJ_CPM 2   /* constant Potts model, unit vertex weights */
#define PGGC_OBJ_NLCC 3  /* normalized LambdaCC */
#define PGGC_OBJ_LCC 4   /* base LambdaCC */

/* algorithms (app/CMakeLists.txt:3-7: plouvain, pleiden, pleiden+) */
#define PGGC_LOUVAIN 0
#define PGGC_LEIDEN 1
#define PGGC_LEIDEN_PLUS 2

typedef struct pggc_ctx pggc_ctx;

/* weighted_graph (interface/weighted_graph.h:10-19) */
typedef struct pggc_graph {
    int64_t n;
    int64_t nnz;
    const int64_t* row_ptr;
    const int32_t* col;
    const int32_t* w;    /* int32 weights or NULL */
    const int64_t* w64;  /* int64 weights or NULL */
    const int64_t* vtx_w;
} pggc_graph;

/* cluster_data (interface/cluster_data.hpp:18-28) plus the objective's
 * normalisation: objective() = obj * scale + offset.  vol is a device array
 * of vol_cap >= n entries owned by the caller. */
typedef struct pggc_state {
    int64_t* vol;
    int64_t vol_cap;
    int64_t uncut;
    int64_t top_nnz;
    double obj;
    double lambda;
    double scale;
    double offset;
    int32_t label_count;
    int32_t reserved;
} pggc_state;

int pggc_version(void);
const char* pggc_last_error(void);

/* stream: a hipStream_t (NULL = default stream) */
int pggc_ctx_create(pggc_ctx** ctx, void* stream);
int pggc_ctx_destroy(pggc_ctx* ctx);
/* seed stream used by the Leiden refinement (stands in for the reference's
 * std::random_device, src/leidenR.cpp:289,291,399,479) */
int pggc_ctx_set_seed(pggc_ctx* ctx, uint64_t seed, uint64_t draws);
int pggc_ctx_get_draws(pggc_ctx* ctx, uint64_t* draws);
/* level > 0: pggc_cluster prints one line per outer iteration to stderr, as
 * the reference CLI does (app/cluster_main.cpp:137,156); level >= 2 also
 * prints the device memory the context holds at each stage of a clustering
 * call (local moves, refinement, contraction) */
int pggc_ctx_set_verbose(pggc_ctx* ctx, int level);
/* Kernel timing (measurement hook, no reference counterpart): when on, HIP
 * events are recorded on the context stream around the launches of each slot
 * below, together with the algorithmic bytes of each launch.  set_timing
 * resets the totals; kernel_stats waits for the stream and returns the
 * launches, summed milliseconds and summed algorithmic bytes of one slot. */
#define PGGC_KS_SCORE_GROUP 0    /* neighbour-cluster scoring, d <= 64 */
#define PGGC_KS_SCORE_TABLE 1    /* scoring with LDS tables, 64 < d <= 2048 */
#define PGGC_KS_SCORE_HUB 2      /* hub scoring (HBM table build or scan), d > 2048 */
#define PGGC_KS_AFTERBURNER 3    /* conflict filter */
#define PGGC_KS_MOVES 4          /* hub-table updates, exact uncut delta, volume moves */
#define PGGC_KS_CONTRACT_AGG 5   /* coarse-row merge in LDS */
#define PGGC_KS_CONTRACT_BIG 6   /* coarse-row merge in HBM tables */
#define PGGC_KS_CONTRACT_COPY 7  /* coarse-row copy-out */
#define PGGC_KS_EVAL 8           /* evaluator edge pass */
#define PGGC_KS_LEIDEN 9         /* leidenR refinement */
#define PGGC_KS_ORDER 10         /* ordering pass */
#define PGGC_KS_COUNT 11
int pggc_ctx_set_timing(pggc_ctx* ctx, int on);
int pggc_ctx_kernel_stats(pggc_ctx* ctx, int slot, int64_t* launches, double* ms, double* bytes);

/* Per-outer-iteration callback of pggc_cluster, called on the host after
 * the first clustering and after every ex_iters iteration with the wall
 * seconds of that iteration, the edge cut, get_objective() and the label
 * count -- what the reference CLI prints and logs per iteration
 * (app/cluster_main.cpp:137,156; interface/ExperimentLoggerUtil.hpp:135-141).
 * NULL disables it. */
typedef void (*pggc_progress_fn)(void* user, int64_t iteration, double seconds, int64_t cut, double objective,
                                 int32_t labels);
int pggc_ctx_set_progress(pggc_ctx* ctx, pggc_progress_fn fn, void* user);
/* Per-sweep trace of the local mover.  No counterpart in the reference,
 * whose logger keeps per-level totals only (interface/ExperimentLoggerUtil.hpp:76-98):
 * called on the host after every sweep with rec[0..7] = level vertex
 * count, level nnz, phase (0 local move from singletons, 1 local move from
 * a clustering, 2 strict local move), sweep index within the call, active
 * vertices (scored this sweep), candidates (positive gain), accepted moves
 * and the sweep's microseconds.  It synchronises the stream every sweep;
 * NULL disables it. */
typedef void (*pggc_sweep_fn)(void* user, const int64_t* rec, int nrec);
int pggc_ctx_set_sweep_trace(pggc_ctx* ctx, pggc_sweep_fn fn, void* user);

/* Tuning knobs of the HIP path (no counterpart in the reference, whose
 * equivalents are compile-time constants: src/header/ordering.h:17-19,
 * src/local_mover.cpp:985,1312).
 *   PGGC_KNOB_HUB_UPDATE_DIV: the hash tables of the highest-degree vertices
 *     are kept current from the movers while a sweep moves at most
 *     n / value vertices and rebuilt otherwise; 0 rebuilds every sweep.
 *     Default 32.
 *   PGGC_KNOB_GNN_MAX_SWEEPS: the active-set predictor runs on at most this
 *     many sweeps of a local-move call; 0 (default) runs it on every
 *     eligible sweep.  (It changes which vertices are pruned, so with
 *     pruning on it changes results.)
 * Execution shape of the scoring kernels; clustering results are identical
 * for every value, only time changes (values other than the listed ones are
 * rejected):
 *   PGGC_KNOB_TABLE_THREADS: threads per vertex of the LDS-hash-table
 *     scoring of vertices with 64 < degree <= 512: 64 (default), 128, 256.
 *   PGGC_KNOB_TABLE_SLOTS_PER_EDGE: LDS hash-table slots per edge of a degree
 *     class's bound, 64 < degree <= 2048: 2 (default) or 4.
 *   PGGC_KNOB_HUB_LDS_MAX_DEGREE: vertices of degree 2049..value are scored
 *     from an 8192-slot LDS table instead of an HBM hub table: 2048
 *     (default, none) or 4096.
 *   PGGC_KNOB_HUB_EDGES_PER_ITEM: edges per hub-table build work item:
 *     512, 1024 (default), 2048.
 *   PGGC_KNOB_HUB_SLOTS_PER_ITEM: hub-table slots per scan work item:
 *     2048, 4096 (default), 8192.
 *   PGGC_KNOB_RAW_SWEEPS: 1 (default) scores the first sweep of a local move
 *     from singletons on the raw adjacency (src/local_mover.cpp:1391-1395),
 *     0 with the neighbour-cluster tables (the same result, more work).
 *   PGGC_KNOB_HUB_BUILD: how HBM hub tables are built: 0 (default) one LDS
 *     slice of a table per block, written out whole (the largest tables by
 *     edge work items); 1 every table by edge work items adding into the HBM
 *     table with atomics; 2 the tables of odd hub index by edge items, the
 *     rest by slices (the same tables' contents in every mode).
 *   PGGC_KNOB_RENUMBER: where the local mover of the top level runs: on an
 *     internal copy of the graph with vertices in descending degree order
 *     (rows keep their entry order; labels, priorities and new cluster ids
 *     follow the caller's ids, so results are identical) or on the caller's
 *     graph.  1: always the copy; 2 (default): the copy when the call makes
 *     more than one top-level pass (pggc_cluster with ex_iters != 0), which
 *     pays for building it; 0: never.  Not used with the active-set
 *     predictor on.
 *   PGGC_KNOB_CONTRACT_BATCH: the contraction merges its largest coarse rows
 *     in HBM hash tables and sorts them in batches of rows holding at most
 *     this many table slots (default 2^30) and a quarter as many sort
 *     entries, which bounds its memory on graphs of billions of edges; the
 *     coarse graph is the same for every value >= 1. */
#define PGGC_KNOB_HUB_UPDATE_DIV 0
#define PGGC_KNOB_GNN_MAX_SWEEPS 1
#define PGGC_KNOB_TABLE_THREADS 2
#define PGGC_KNOB_TABLE_SLOTS_PER_EDGE 3
#define PGGC_KNOB_HUB_LDS_MAX_DEGREE 4
#define PGGC_KNOB_HUB_EDGES_PER_ITEM 5
#define PGGC_KNOB_HUB_SLOTS_PER_ITEM 6
#define PGGC_KNOB_RAW_SWEEPS 7
#define PGGC_KNOB_HUB_BUILD 8
#define PGGC_KNOB_RENUMBER 9
#define PGGC_KNOB_CONTRACT_BATCH 10
#define PGGC_KNOB_COUNT 11
int pggc_ctx_set_knob(pggc_ctx* ctx, int knob, int64_t value);
int pggc_ctx_get_knob(pggc_ctx* ctx, int knob, int64_t* value);
/* Autotuner of the execution-shape knobs (no counterpart in the reference):
 * coordinate descent in the order hub LDS max degree, table threads, table
 * slots per edge, hub edges per item, hub slots per item, hub update divisor;
 * each candidate value is timed on one pLouvain clustering of g
 * (louvain_part, src/clustering_methods.cpp:139-222; best of reps runs after
 * an untimed one), and the fastest value is kept on ctx.  report (host, nullable): up to cap rows of
 * (knob, value, microseconds); *nrows (nullable) receives the rows timed.
 * The active-set predictor must be off. */
int pggc_autotune(pggc_ctx* ctx, const pggc_graph* g, int objective, int reps, int64_t* report, int cap,
                  int* nrows);

/* ---------------------------------------------------------------------
 * Active-set predictor.  No counterpart in the reference: a two-layer graph
 * network (hidden width 32, bf16 on MFMA with fp32 accumulation) that, before
 * a sweep of local_move (src/local_mover.cpp:1360-1417), predicts which
 * active vertices will move and caches NO_MOVE for the others -- the same
 * dest-cache the reference uses to skip vertices until a neighbour moves
 * (:477-479,717,798).  A sweep that finds nothing to move among the vertices
 * let through is repeated with the pruned ones restored, so a call that
 * stops for lack of moves stops under the plain method's condition; a call
 * that stops at its sweep limit leaves pruned vertices for the next call,
 * which starts with every vertex active.  The predictor runs on a sweep only
 * while the previous sweep left at least n/16 candidates.  Applies to
 * pggc_cluster, pggc_louvain and pggc_leiden on levels with at least
 * min_nnz nonzeros.
 *
 *   PGGC_GNN_TRAIN: sweeps run plain; after each, one Adam step (rate lr) on
 *                   the weighted cross-entropy of "moved in this sweep" over
 *                   the sweep's active vertices (moved class weighted by
 *                   pos_weight).  Starts from fixed initial parameters if none
 *                   were set.
 *   PGGC_GNN_PRUNE: active vertices with logit < tau are pruned; with
 *                   full_every = k > 0 every k-th sweep of a call restores
 *                   and scores all of them.  Needs parameters.
 * Parameters are PGGC_GNN_PARAMS floats: W1[16][32] (input feature major),
 * b1[32], w2s[32], w2n[32], b2. */
#define PGGC_GNN_OFF 0
#define PGGC_GNN_TRAIN 1
#define PGGC_GNN_PRUNE 2
#define PGGC_GNN_PARAMS 609
int pggc_gnn_configure(pggc_ctx* ctx, int mode, double tau, int64_t min_nnz, int full_every, double lr,
                       double pos_weight);
int pggc_gnn_set_params(pggc_ctx* ctx, const float* params, int count);  /* host array; resets Adam */
int pggc_gnn_get_params(pggc_ctx* ctx, float* params, int count);        /* host array */
/* counts[9] = predictor sweeps, active vertices seen, prune decisions,
 * restore sweeps, training steps, and over training sweeps the confusion
 * counts of (logit >= tau) against moved: tp, fp, fn, tn; *loss_sum = summed
 * per-sweep mean loss of the training steps */
int pggc_gnn_stats(pggc_ctx* ctx, int64_t* counts, double* loss_sum);
int pggc_gnn_reset_stats(pggc_ctx* ctx);
/* One predictor pass from a given state, for tests and offline evaluation.
 * Device inputs: part[n], dest[n] (NO_MOVE = -3 marks inactive vertices),
 * gain[n], stamp[n] (stamp == cur-1 marks last sweep's movers), vol[labels];
 * first != 0 zeroes the gain feature.  Device outputs (nullable): rows_out
 * [n*16] bf16 input rows, logit_out[n] (0 for inactive vertices).  With
 * labels[n] (device, nonzero = moved): *loss_out (host) and grad_out[609]
 * (host) of the mean loss, without a parameter update. */
int pggc_gnn_eval(pggc_ctx* ctx, const pggc_graph* g, const int32_t* part, const int32_t* dest, const float* gain,
                  const int32_t* stamp, int32_t cur, int first, const int64_t* vol, const int32_t* labels,
                  uint16_t* rows_out, float* logit_out, double* loss_out, float* grad_out);
/* wall seconds spent in contraction and in Leiden refinement during the last
 * pggc_cluster call (the reference logs them as the coarsen-contract and
 * leiden-refine measurements, src/clustering_methods.cpp:89,117,220) */
int pggc_ctx_last_times(pggc_ctx* ctx, double* contract_s, double* refine_s);

/* bytes held by the context allocator (host out) */
int pggc_ctx_memory(pggc_ctx* ctx, int64_t* live, int64_t* cached);

/* get_vtx_weights (app/io/vertex_weighting.cpp:27-44): degree (MOD),
 * weighted degree (WMOD), unit otherwise. */
int pggc_vertex_weights(pggc_ctx* ctx, const pggc_graph* g, int objective, int64_t* vtx_w_out);

/* get_objective (app/objective_helpers.hpp:9-22): sets lambda and the
 * normalisation, vol = vtx_w, singleton clustering, obj updated. */
int pggc_state_init(pggc_ctx* ctx, const pggc_graph* g, int objective, double lambda_mult, pggc_state* st);

/* generate_orderings (src/ordering.cpp:131-164).  offs (host) receives
 * {offset_large, offset_large2, offset_mid, offset_mid2}. */
int pggc_orderings(pggc_ctx* ctx, const pggc_graph* g, int32_t* order1, int32_t* order2, int64_t* offs);

/* modularity evaluator (app/verify_objective.cpp:64-89 with
 * interface/cluster_data.hpp:76-87): labels in [0,L); vol_out[L] (device);
 * uncut_out and sq_out[2] = {high, low} words of sum(vol^2) on the host. */
int pggc_evaluate(pggc_ctx* ctx, const pggc_graph* g, const int32_t* labels, int64_t L, int64_t* vol_out,
                  int64_t* uncut_out, uint64_t* sq_out);

/* pggc_evaluate without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as pggc_evaluate.  Its scratch (per-chunk
 * partial sums) lives in a static device array of
 * pggc_evaluate_noalloc_limits()->scratch_bytes (about 8 MiB) that each GPU
 * holds until the process ends; graphs with more than max_nnz (2^32)
 * nonzeros are rejected with PGGC_ECAPACITY.  All calls on one GPU share that
 * array: calls that overlap on one GPU (several streams or host threads)
 * race. */
int pggc_evaluate_noalloc(pggc_ctx* ctx, const pggc_graph* g, const int32_t* labels, int64_t L, int64_t* vol_out,
                          int64_t* uncut_out, uint64_t* sq_out);
int pggc_evaluate_noalloc_limits(int64_t* max_nnz, int64_t* scratch_bytes);  /* host outputs, nullable */

/* own-cluster weight of every vertex (the pvals_clone contract,
 * src/local_mover.cpp:1354-1358) */
int pggc_own_weight(pggc_ctx* ctx, const pggc_graph* g, const int32_t* labels, int64_t* out);

/* build_coarse_graph (src/contract.cpp:234-336): coarse CSR of the labelling
 * vcmap[n] in [0,nc), self loops dropped, rows sorted by column.  out_rp has
 * nc+1 entries; out_col/out_w hold cap entries.  *nnz_out receives the coarse
 * nnz; PGGC_ECAPACITY if it exceeds cap (nothing written to out_col/out_w). */
int pggc_contract(pggc_ctx* ctx, const pggc_graph* g, const int32_t* vcmap, int64_t nc, int64_t* out_rp,
                  int32_t* out_col, int64_t* out_w, int64_t cap, int64_t* nnz_out);

/* One local-move sweep from a fixed state (candidates_and_destinations +
 * afterburner_filter, src/local_mover.cpp:567-634,374-433): dest and gain are
 * the in/out destination cache and gains; raw != 0 scores the raw adjacency
 * (first sweep of an initial call); constraint may be NULL.  moves_out (n)
 * receives the accepted moves in order1 order; counts on the host. */
int pggc_sweep(pggc_ctx* ctx, const pggc_graph* g, const int32_t* part, const int64_t* vol, int64_t L,
               double lambda, float ratio, int raw, const int32_t* constraint, int32_t* dest, float* gain,
               int32_t* moves_out, int64_t* nmoves, int64_t* ncand);

/* local_move / local_move_strict (src/local_mover.cpp:1360-1483) on one level,
 * unconstrained: part_inout is relabelled contiguously; st in/out. */
int pggc_local_move(pggc_ctx* ctx, const pggc_graph* g, int32_t* part_inout, pggc_state* st, int is_initial,
                    int strict, int simulated_annealing);

/* coarsen_leidenR (src/leidenR.cpp:263-500) of the Louvain labels louv with
 * volumes vol[L]; uniform != 0 uses the unit-weight neighbour rule.  The two
 * seeds are draws draw0 and draw0+1 of the stream `seed`. */
int pggc_leiden_refine(pggc_ctx* ctx, const pggc_graph* g, const int32_t* louv, const int64_t* vol, int64_t L,
                       double lambda, int uniform, uint64_t seed, uint64_t draw0, int32_t* vcmap_out,
                       int64_t* nc_out);

/* louvain_part<constrained> (interface/clustering_methods.h:23-24): constraint
 * (n entries, mutated like the reference) is used when constrained != 0. */
int pggc_louvain(pggc_ctx* ctx, const pggc_graph* g, pggc_state* st, int constrained, int32_t* constraint,
                 int32_t* part_out);

/* leiden_part<plus, improve> (interface/clustering_methods.h:20-21) */
int pggc_leiden(pggc_ctx* ctx, const pggc_graph* g, pggc_state* st, int plus, int improve, const int32_t* input,
                int32_t* part_out);

/* run_clustering (app/cluster_main.cpp:116-165) for one trial: the first
 * clustering plus the ex_iters loop (-1 = until the objective stops rising).
 * objective_out (host) = get_objective(); info (host, 8 entries) =
 * {label_count, uncut, levels, sweeps, moves, outer iterations, candidates, total nnz of the
 * last call's level graphs}. */
int pggc_cluster(pggc_ctx* ctx, const pggc_graph* g, int algo, int objective, double lambda_mult, int ex_iters,
                 uint64_t seed, int32_t* part_out, double* objective_out, int64_t* info);

/* Batched runs: `count` independent pggc_cluster calls (different graphs, or
 * seeded trials of one graph as in run_clustering's trial loop,
 * app/cluster_main.cpp:189-204, or meme population members) in one call.
 * Each instance gives exactly what pggc_cluster gives for it alone.
 * pLouvain with ex_iters 0 and the predictor off runs the instances as one
 * disjoint union graph: every kernel is launched once over all instances,
 * and each instance keeps its own lambda, sweep schedule, stopping rules,
 * best state and labels (`streams` is unused).  Other algorithms use up to
 * `streams` worker contexts, each with a host thread of its own and all on
 * the calling context's stream, taking instances in order: the host work of
 * small instances (launches, readbacks) overlaps while their kernels run in
 * submission order.  graphs[count] (host array of device graphs), seeds[count]
 * (host, nullable: instance index), parts_out[count] (host array of device
 * pointers), objectives_out[count] and info_out[8*count] (host, nullable).
 * The call first waits for the context's stream, so inputs produced on it
 * are complete, and returns when every instance has finished.
 * The context's knobs and pruning predictor (not training) apply to every
 * instance; progress, verbose and trace callbacks do not.  The first failing
 * instance stops the batch and is named in the error message. */
int pggc_cluster_batch(pggc_ctx* ctx, int count, const pggc_graph* graphs, int algo, int objective,
                       double lambda_mult, int ex_iters, const uint64_t* seeds, int32_t* const* parts_out,
                       double* objectives_out, int64_t* info_out, int streams);

/* Memetic loop building blocks (app/memetic_clustering.cpp).
 * gen_signature (:71-106): bit j of sig_out[(nnz+31)/32] (device) is set when
 * CSR edge j joins two clusters of part. */
int pggc_cut_signature(pggc_ctx* ctx, const pggc_graph* g, const int32_t* part, uint32_t* sig_out);
/* get_cut_diff (:108-116): number of differing bits of two signatures of
 * `words` 32-bit words (device); *diff_out on the host */
int pggc_signature_diff(pggc_ctx* ctx, const uint32_t* a, const uint32_t* b, int64_t words, int64_t* diff_out);
/* intersection_cluster (:27-62): out[i] (device) are equal iff both c1 and c2
 * (c2 with l2 labels) agree on i; labels numbered in increasing (c1, c2)
 * order, *labels_out (host) = their count */
int pggc_intersect(pggc_ctx* ctx, int64_t n, const int32_t* c1, const int32_t* c2, int64_t l2, int32_t* out,
                   int64_t* labels_out);

/* ---------------------------------------------------------------------
 * Device-side graph construction.
 *
 * pggc_graph_from_edges replaces construct_from_edgelist
 * (app/io/io_mtx.cpp:95-201): from an edge list in device memory (src[m],
 * dst[m] int32 in [0, n), w[m] int32 >= 0 or NULL) it drops self loops and
 * zero weights, merges repeated undirected pairs (keeping the weight of the
 * first occurrence in input order), stores both directions and sorts every
 * row.  Negative weights or ids outside [0, n) give PGGC_EINVAL.
 *
 * pggc_rmat builds the R-MAT graph of the synthetic workloads on the device:
 * edge_factor * 2^scale draws, each of `scale` quadrant choices taken from a
 * Philox4x32-10 draw (counter = edge index lo, hi, block, 0; key = seed;
 * level l uses word l % 4 of block l / 4; a draw r picks (0,0) when r <
 * floor(a 2^32), (0,1) when r < floor((a+b) 2^32), (1,0) when r < floor((a+b+c)
 * 2^32), else (1,1), first level = most significant bit), then the
 * construction above.  With wmax > 0 each undirected pair lo < hi weighs
 * wmin + Philox(lo, hi, 0, 1; seed) word 0 mod (wmax - wmin + 1).  The
 * reference reads its graphs from files; no generator of its own replaces
 * this.
 *
 * The graph is owned by the returned handle until pggc_devgraph_free; its
 * view (vtx_w = NULL) stays valid until then. */
typedef struct pggc_devgraph pggc_devgraph;
int pggc_graph_from_edges(pggc_ctx* ctx, int64_t n, int64_t m, const int32_t* src, const int32_t* dst,
                          const int32_t* w, pggc_devgraph** out);
int pggc_rmat(pggc_ctx* ctx, int scale, int edge_factor, double a, double b, double c, uint64_t seed, int32_t wmin,
              int32_t wmax, pggc_devgraph** out);
/* pggc_lfr builds an LFR-style benchmark graph (Lancichinetti, Fortunato,
 * Radicchi, Phys. Rev. E 78, 046110, 2008) on the device, with no
 * counterpart in the reference (which reads graphs from files):
 *  - degrees: the integer power law floor(X), X with density ~ x^-tau1 on
 *    [x0, kmax + 1), x0 solved so that the mean degree is kmean;
 *  - community sizes: floor(X), X ~ x^-tau2 on [cmin, cmax + 1), drawn in
 *    sequence until they cover n (the last takes what is left and joins the
 *    one before when that is below cmin); vertices fill them in order;
 *  - internal degree round((1 - mu) degree), at most the community size - 1;
 *  - internal stubs paired after a stable sort by (community, draw), external
 *    stubs after a stable sort by draw over the whole graph; pairs of one
 *    vertex and repeated pairs vanish in the construction above;
 *  - vertex ids permuted by a stable sort on one draw per vertex;
 *  - with wmax > 0, per-pair weights as for pggc_rmat.
 * Every draw is a u32 Philox4x32-10 word compared against integer tables
 * built on the host, so a CPU implementation of the same steps builds the
 * same graph.  gt_out (device, n entries, nullable) receives each vertex's
 * planted community. */
int pggc_lfr(pggc_ctx* ctx, int64_t n, double tau1, double kmean, int32_t kmax, double tau2, int32_t cmin,
             int32_t cmax, double mu, int32_t wmin, int32_t wmax, uint64_t seed, int32_t* gt_out,
             pggc_devgraph** out);
int pggc_devgraph_view(const pggc_devgraph* g, pggc_graph* view);
/* device-to-device copy into caller buffers (row_ptr[n+1], col[nnz], w[nnz]
 * when weighted) on the context's stream */
int pggc_devgraph_copy(pggc_ctx* ctx, const pggc_devgraph* g, int64_t* row_ptr, int32_t* col, int32_t* w);
int pggc_devgraph_free(pggc_devgraph* g);
/* out[i] = Philox4x32-10(counter i lo, i hi, 0, 2; key seed) word 0 mod L:
 * the fixed random labellings of the contraction workloads */
int pggc_random_labels(pggc_ctx* ctx, int64_t n, int64_t L, uint64_t seed, int32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* PGGC_H */
