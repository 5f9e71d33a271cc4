/* Drop-in C++ facade for the reference's clustering entry points
 * (interface/clustering_methods.h:20-24): jet_community::clustering_methods::
 * louvain_part<constrained> and leiden_part<plus, improve>, over the C ABI of
 * include/pggc.h.  Kokkos is not needed: a graph is a struct of device
 * pointers (weighted_graph, interface/weighted_graph.h:10-19, with int64
 * row offsets and optional int32 or int64 edge weights), and the cluster
 * state is pggc_state (cluster_data, interface/cluster_data.hpp:18-28).
 * Errors throw std::runtime_error with the library's message. */
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

#include "pggc.h"

namespace jet_community {
namespace clustering_methods {

inline void check(int rc) {
    if (rc != PGGC_OK) throw std::runtime_error(std::string("pggc: ") + pggc_last_error());
}

/* weighted_graph without Kokkos.  At most one of w / w64 is set (neither:
 * unit weights); vtx_w holds the objective's vertex weights
 * (pggc_vertex_weights). */
struct Graph {
    int64_t n = 0, nnz = 0;
    const int64_t* row_ptr = nullptr;  // device, n + 1
    const int32_t* col = nullptr;      // device, nnz
    const int32_t* w = nullptr;        // device int32 weights or nullptr
    const int64_t* w64 = nullptr;      // device int64 weights or nullptr
    const int64_t* vtx_w = nullptr;    // device, n
};

inline pggc_graph view(const Graph& g) {
    pggc_graph v;
    v.n = g.n;
    v.nnz = g.nnz;
    v.row_ptr = g.row_ptr;
    v.col = g.col;
    v.w = g.w;
    v.w64 = g.w64;
    v.vtx_w = g.vtx_w;
    return v;
}

/* get_vtx_weights + get_objective (app/io/vertex_weighting.cpp:27-44,
 * app/objective_helpers.hpp:9-22): fills vtx_w_out (device, n) and the
 * state of the all-singletons clustering; vol (device, vol_cap >= n) is
 * owned by the caller. */
inline pggc_state make_state(pggc_ctx* ctx, Graph& g, int objective, double lambda_mult, int64_t* vtx_w_out,
                             int64_t* vol, int64_t vol_cap) {
    pggc_graph v = view(g);
    check(pggc_vertex_weights(ctx, &v, objective, vtx_w_out));
    g.vtx_w = vtx_w_out;
    v.vtx_w = vtx_w_out;
    pggc_state st{};
    st.vol = vol;
    st.vol_cap = vol_cap;
    check(pggc_state_init(ctx, &v, objective, lambda_mult, &st));
    return st;
}

/* louvain_part<constrained> (src/clustering_methods.cpp:139-222) */
template <bool constrained>
void louvain_part(pggc_ctx* ctx, const Graph& top, pggc_state& rfd, int32_t* constraint, int32_t* part_out) {
    pggc_graph g = view(top);
    check(pggc_louvain(ctx, &g, &rfd, constrained ? 1 : 0, constraint, part_out));
}

/* leiden_part<plus, improve> (src/clustering_methods.cpp:47-137) */
template <bool plus, bool improve>
void leiden_part(pggc_ctx* ctx, const Graph& top, pggc_state& rfd, const int32_t* input, int32_t* part_out) {
    pggc_graph g = view(top);
    check(pggc_leiden(ctx, &g, &rfd, plus ? 1 : 0, improve ? 1 : 0, input, part_out));
}

}  // namespace clustering_methods
}  // namespace jet_community
