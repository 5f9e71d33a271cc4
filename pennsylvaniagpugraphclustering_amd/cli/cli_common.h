// Shared parts of the command-line tools: argument parsing with the
// reference's option names and messages (app/io/parse_args.cpp), the
// configuration printout, and loading a graph file into device memory with
// its vertex weights (app/io/vertex_weighting.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "pggc.h"
#include "pggc_io.h"

namespace cli {

enum class Objective { Modularity, WModularity, CPM, NLCC, Base };

struct BaseArgs {
    std::string graph_file, output_file, vw_file;
    Objective obj = Objective::Modularity;
    double lambda_multiplier = 1.0;
    bool valid = true;
};
struct ClusterArgs : BaseArgs {
    int n_trials = 1;
    int n_successive_iterations = 0;
    std::string metrics_file;
};
struct VerifyArgs : BaseArgs {
    std::string cluster_file;
};
struct MemeArgs : BaseArgs {
    int pop_size = 10;
    int time_limit = 10;
};

ClusterArgs parse_cluster_args(int argc, char** argv);
VerifyArgs parse_verify_args(int argc, char** argv);
MemeArgs parse_meme_args(int argc, char** argv);

int objective_code(Objective o);      // PGGC_OBJ_*
bool uses_unit_edges(Objective o);    // app/objective_helpers.hpp:24-29
// the label cluster_data::print uses (interface/cluster_data.hpp:89-91,127-129,148-150,171-173)
const char* objective_label(Objective o);

// A graph resident in device memory, with vertex weights.
struct DeviceGraph {
    pggc_graph g{};
    int64_t* rp = nullptr;
    int32_t* col = nullptr;
    int32_t* w = nullptr;
    int64_t* vw = nullptr;
    bool weighted = false;
    pggc_devgraph* gen = nullptr;  // a generated graph (owns the arrays g points at)
    ~DeviceGraph();
};

// Reads the graph file, or builds the graph on the GPU when the name is
// rmat:SCALE:EDGE_FACTOR:SEED[:WMIN:WMAX] (R-MAT with a,b,c = .57,.19,.19,
// per-pair weights uniform in [WMIN, WMAX] when given), with unit edges if
// the objective says so; prints the
// vertex / undirected edge counts, uploads it and sets the vertex weights
// (a -vtx_weights file, else degree / weighted degree / unit by objective).
// Returns false after printing a FATAL ERROR line.
bool load_device_graph(pggc_ctx* ctx, const BaseArgs& args, DeviceGraph& dg);

// "FATAL ERROR: <what>: <pggc_last_error()>" on stderr when rc != PGGC_OK
bool check(int rc, const char* what);

}  // namespace cli
