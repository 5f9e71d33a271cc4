// Argument parsing, configuration printout and device graph loading for the
// command-line tools.  Option names, defaults, clamps and messages follow
// app/io/parse_args.cpp; vertex weighting follows app/io/vertex_weighting.cpp.
#include "cli_common.h"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <iostream>
#include <string>
#include <unordered_set>
#include <vector>

namespace cli {

static const std::unordered_map<std::string, Objective> kObjectives = {
    {"Mod", Objective::Modularity},
    {"WMod", Objective::WModularity},
    {"CPM", Objective::CPM},
    {"NLCC", Objective::NLCC},
    {"LCC", Objective::Base},
};

int objective_code(Objective o) {
    switch (o) {
        case Objective::Modularity: return PGGC_OBJ_MOD;
        case Objective::WModularity: return PGGC_OBJ_WMOD;
        case Objective::CPM: return PGGC_OBJ_CPM;
        case Objective::NLCC: return PGGC_OBJ_NLCC;
        default: return PGGC_OBJ_LCC;
    }
}

bool uses_unit_edges(Objective o) { return o == Objective::Modularity || o == Objective::CPM; }

const char* objective_label(Objective o) {
    switch (o) {
        case Objective::Modularity:
        case Objective::WModularity: return "Modularity";
        case Objective::CPM: return "Constant-Potts";
        case Objective::NLCC: return "Normalized LambdaCC";
        default: return "Objective";
    }
}

// print_configuration_base (app/io/parse_args.cpp:19-54)
static void print_base(const BaseArgs& a) {
    std::cout << "Program configuration: " << std::endl;
    std::cout << "-- Input Graph File: " << a.graph_file << std::endl;
    if (!a.vw_file.empty()) {
        std::cout << "-- Input Vertex Weights File: " << a.vw_file << std::endl;
    } else {
        std::cout << "-- Vertex Weighting Method: ";
        if (a.obj == Objective::Modularity) std::cout << "Degree" << std::endl;
        else if (a.obj == Objective::WModularity) std::cout << "Weighted Degree" << std::endl;
        else std::cout << "Unit Uniform" << std::endl;
    }
    std::cout << "-- Objective Type: ";
    switch (a.obj) {
        case Objective::Modularity: std::cout << "Modularity" << std::endl; break;
        case Objective::WModularity: std::cout << "Weighted Modularity" << std::endl; break;
        case Objective::CPM: std::cout << "Constant-Potts Model" << std::endl; break;
        case Objective::NLCC: std::cout << "Normalized LambdaCC" << std::endl; break;
        case Objective::Base: std::cout << "Base LambdaCC" << std::endl; break;
    }
    std::cout << "-- Lambda Multiplier: " << a.lambda_multiplier << std::endl;
    if (!a.output_file.empty()) std::cout << "-- Output Clusters File: " << a.output_file << std::endl;
}

// parse_base_args (app/io/parse_args.cpp:75-144)
static std::unordered_map<std::string, std::string> parse_base(int argc, char** argv,
                                                               const std::vector<std::string>& extra, BaseArgs& a) {
    std::unordered_set<std::string> valid = {"-i", "-o", "-objective", "-lambda", "-vtx_weights"};
    valid.insert(extra.begin(), extra.end());
    std::unordered_map<std::string, std::string> given;
    a.valid = true;
    std::string name;
    bool expect_name = true;
    for (int k = 1; k < argc; k++) {
        const std::string opt = argv[k];
        if (expect_name) {
            if (!valid.count(opt)) {
                std::cerr << "FATAL ERROR: Invalid option name: " << opt << std::endl;
                std::cerr << "Valid Options:" << std::endl;
                for (const auto& v : valid) std::cerr << v << std::endl;
                a.valid = false;
                return given;
            }
            name = opt;
            expect_name = false;
        } else {
            given[name] = opt;
            expect_name = true;
        }
    }
    if (!expect_name) {
        std::cerr << "FATAL ERROR: Value not provided for option '" << name << "'" << std::endl;
        a.valid = false;
        return given;
    }
    if (!given.count("-i")) {
        std::cerr << "FATAL ERROR: Required graph input file not given. Please specify with argument '-i <graph file>'"
                  << std::endl;
        a.valid = false;
    }
    a.graph_file = given["-i"];
    if (given.count("-o")) a.output_file = given["-o"];
    if (given.count("-objective")) {
        const std::string o = given["-objective"];
        auto it = kObjectives.find(o);
        if (it != kObjectives.end()) {
            a.obj = it->second;
        } else {
            std::cerr << "FATAL ERROR: Invalid objective name '" << o << "'!" << std::endl;
            a.valid = false;
        }
    }
    if (given.count("-lambda")) a.lambda_multiplier = std::stod(given["-lambda"]);
    if (given.count("-vtx_weights")) a.vw_file = given["-vtx_weights"];
    return given;
}

// verify_config (app/io/parse_args.cpp:146-153)
static void verify_config(BaseArgs& a) {
    if (!a.vw_file.empty() && a.obj != Objective::NLCC && a.obj != Objective::Base) {
        a.valid = false;
        std::cerr << "FATAL ERROR: Vertex weights file is only supported with objectives Normalized LambdaCC and "
                     "Base LambdaCC!"
                  << std::endl;
    }
}

ClusterArgs parse_cluster_args(int argc, char** argv) {
    ClusterArgs a;
    auto given = parse_base(argc, argv, {"-trials", "-ex_iters", "-metrics"}, a);
    if (!a.valid) return a;
    if (given.count("-trials")) a.n_trials = std::stoi(given["-trials"]);
    if (a.n_trials < 1) a.n_trials = 1;
    if (given.count("-ex_iters")) a.n_successive_iterations = std::stoi(given["-ex_iters"]);
    if (a.n_successive_iterations < -1) a.n_successive_iterations = -1;
    if (given.count("-metrics")) a.metrics_file = given["-metrics"];
    print_base(a);
    std::cout << "-- Successive Iteration Count: " << a.n_successive_iterations << std::endl;
    std::cout << "-- Total trial count: " << a.n_trials << std::endl;
    if (!a.metrics_file.empty()) std::cout << "-- Metrics Output File: " << a.metrics_file << std::endl;
    verify_config(a);
    return a;
}

// parse_meme_args (app/io/parse_args.cpp:174-197)
MemeArgs parse_meme_args(int argc, char** argv) {
    MemeArgs a;
    auto given = parse_base(argc, argv, {"-pop_size", "-time_limit"}, a);
    if (!a.valid) return a;
    if (given.count("-pop_size")) a.pop_size = std::stoi(given["-pop_size"]);
    if (a.pop_size < 10) {
        std::cout << "WARNING: Population size given as " << a.pop_size
                  << " < 10. Setting population size to 10." << std::endl;
        a.pop_size = 10;
    }
    if (given.count("-time_limit")) a.time_limit = std::stoi(given["-time_limit"]);
    if (a.time_limit < 10) {
        std::cout << "WARNING: Time limit given as " << a.time_limit << "s < 10s. Setting time limit to 10s."
                  << std::endl;
        a.time_limit = 10;
    }
    print_base(a);
    std::cout << "-- Population Size: " << a.pop_size << std::endl;
    std::cout << "-- Time Limit Seconds: " << a.time_limit << std::endl;
    verify_config(a);
    return a;
}

VerifyArgs parse_verify_args(int argc, char** argv) {
    VerifyArgs a;
    auto given = parse_base(argc, argv, {"-clusters"}, a);
    if (!a.valid) return a;
    if (!given.count("-clusters")) {
        std::cerr << "FATAL ERROR: Required clusters file not given! Please specify with argument '-clusters <cluster "
                     "file>'"
                  << std::endl;
        a.valid = false;
        return a;
    }
    a.cluster_file = given["-clusters"];
    print_base(a);
    std::cout << "-- Clusters File: " << a.cluster_file << std::endl;
    verify_config(a);
    return a;
}

bool check(int rc, const char* what) {
    if (rc == PGGC_OK) return true;
    std::cerr << "FATAL ERROR: " << what << ": " << pggc_last_error() << std::endl;
    return false;
}

static bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::cerr << "FATAL ERROR: " << what << ": " << hipGetErrorString(e) << std::endl;
    return false;
}

DeviceGraph::~DeviceGraph() {
    (void)hipFree(rp);
    (void)hipFree(col);
    (void)hipFree(w);
    (void)hipFree(vw);
    if (gen) (void)pggc_devgraph_free(gen);
}

// rmat:SCALE:EDGE_FACTOR:SEED[:WMIN:WMAX] -> a graph built on the device
static bool generate_device_graph(pggc_ctx* ctx, const BaseArgs& args, DeviceGraph& dg) {
    std::vector<long long> f;
    std::string spec = args.graph_file.substr(5);
    size_t at = 0;
    while (at <= spec.size()) {
        const size_t nx = spec.find(':', at);
        const std::string tok = spec.substr(at, nx == std::string::npos ? std::string::npos : nx - at);
        char* end = nullptr;
        const long long v = std::strtoll(tok.c_str(), &end, 10);
        if (tok.empty() || *end != '\0') {
            std::cerr << "FATAL ERROR: graph spec must be rmat:SCALE:EDGE_FACTOR:SEED[:WMIN:WMAX]" << std::endl;
            return false;
        }
        f.push_back(v);
        if (nx == std::string::npos) break;
        at = nx + 1;
    }
    if (f.size() != 3 && f.size() != 5) {
        std::cerr << "FATAL ERROR: graph spec must be rmat:SCALE:EDGE_FACTOR:SEED[:WMIN:WMAX]" << std::endl;
        return false;
    }
    const int32_t wmin = f.size() == 5 ? (int32_t)f[3] : 0, wmax = f.size() == 5 ? (int32_t)f[4] : 0;
    if (!check(pggc_rmat(ctx, (int)f[0], (int)f[1], 0.57, 0.19, 0.19, (uint64_t)f[2], wmin, wmax, &dg.gen),
               "generating the graph"))
        return false;
    pggc_graph v{};
    if (!check(pggc_devgraph_view(dg.gen, &v), "generating the graph")) return false;
    std::cout << "Vertex Count: " << v.n << "; Undirected Edge Count: " << v.nnz / 2 << std::endl;
    std::cout << std::endl;
    dg.weighted = v.w != nullptr && !uses_unit_edges(args.obj);
    if (!hip_ok(hipMalloc((void**)&dg.vw, (v.n > 0 ? v.n : 1) * sizeof(int64_t)), "device allocation")) return false;
    dg.g = pggc_graph{v.n, v.nnz, v.row_ptr, v.col, dg.weighted ? v.w : nullptr, nullptr, dg.vw};
    if (!args.vw_file.empty()) {
        std::vector<int64_t> vw(v.n);
        if (!check(pggc_read_ints(args.vw_file.c_str(), v.n, vw.data()), "reading the vertex weights")) return false;
        return hip_ok(hipMemcpy(dg.vw, vw.data(), v.n * sizeof(int64_t), hipMemcpyHostToDevice), "upload");
    }
    return check(pggc_vertex_weights(ctx, &dg.g, objective_code(args.obj), dg.vw), "vertex weights");
}

bool load_device_graph(pggc_ctx* ctx, const BaseArgs& args, DeviceGraph& dg) {
    if (args.graph_file.rfind("rmat:", 0) == 0) return generate_device_graph(ctx, args, dg);
    pggc_host_graph* hg = nullptr;
    int weighted = 0;
    if (!check(pggc_read_graph(args.graph_file.c_str(), uses_unit_edges(args.obj) ? 1 : 0, &hg, &weighted),
               "reading the graph"))
        return false;
    int64_t n = 0, nnz = 0;
    pggc_host_graph_info(hg, &n, &nnz);
    std::cout << "Vertex Count: " << n << "; Undirected Edge Count: " << nnz / 2 << std::endl;
    std::cout << std::endl;
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> col(nnz), w(weighted ? nnz : 0);
    pggc_host_graph_copy(hg, rp.data(), col.data(), weighted ? w.data() : nullptr);
    pggc_host_graph_free(hg);
    if (!hip_ok(hipMalloc((void**)&dg.rp, (n + 1) * sizeof(int64_t)), "device allocation") ||
        !hip_ok(hipMalloc((void**)&dg.col, (nnz > 0 ? nnz : 1) * sizeof(int32_t)), "device allocation") ||
        !hip_ok(hipMalloc((void**)&dg.vw, (n > 0 ? n : 1) * sizeof(int64_t)), "device allocation"))
        return false;
    if (!hip_ok(hipMemcpy(dg.rp, rp.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice), "upload") ||
        !hip_ok(hipMemcpy(dg.col, col.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice), "upload"))
        return false;
    if (weighted) {
        if (!hip_ok(hipMalloc((void**)&dg.w, (nnz > 0 ? nnz : 1) * sizeof(int32_t)), "device allocation") ||
            !hip_ok(hipMemcpy(dg.w, w.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice), "upload"))
            return false;
    }
    dg.weighted = weighted != 0;
    dg.g = pggc_graph{n, nnz, dg.rp, dg.col, dg.w, nullptr, dg.vw};
    if (!args.vw_file.empty()) {
        std::vector<int64_t> vw(n);
        if (!check(pggc_read_ints(args.vw_file.c_str(), n, vw.data()), "reading the vertex weights")) return false;
        if (!hip_ok(hipMemcpy(dg.vw, vw.data(), n * sizeof(int64_t), hipMemcpyHostToDevice), "upload")) return false;
    } else {
        if (!check(pggc_vertex_weights(ctx, &dg.g, objective_code(args.obj), dg.vw), "vertex weights")) return false;
    }
    return true;
}

}  // namespace cli
