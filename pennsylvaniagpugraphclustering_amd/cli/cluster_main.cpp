// plouvain / pleiden / pleiden+ (app/cluster_main.cpp, app/CMakeLists.txt:3-5):
// the same options, stdout lines, partition file and -metrics JSON as the
// reference CLI, with every clustering run on the GPU through the C ABI.
// PGGC_ALGO selects the method at compile time (0 Louvain, 1 Leiden,
// 2 Leiden+).  The reference seeds Leiden refinement from std::random_device;
// here trial t uses seed PGGC_SEED + t (environment, default 0).
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <vector>

#include <hip/hip_runtime.h>

#include "cli_common.h"

#ifndef PGGC_ALGO
#define PGGC_ALGO 0
#endif

namespace {

struct TrialLog {
    const char* label;
    std::vector<double> times, objs;
};

// "Cluster time: t; Cut: c; <label>: q; Labels: l" (app/cluster_main.cpp:137,156 with
// interface/cluster_data.hpp:98-103), 6 significant digits
void on_iteration(void* user, int64_t, double seconds, int64_t cut, double objective, int32_t labels) {
    auto* log = static_cast<TrialLog*>(user);
    std::cout << std::setprecision(6) << "Cluster time: " << seconds << "; Cut: " << cut << "; " << log->label << ": "
              << objective << "; Labels: " << labels << std::endl;
    log->times.push_back(seconds);
    log->objs.push_back(objective);
}

// ExperimentLoggerUtil::log (interface/ExperimentLoggerUtil.hpp:151-186)
void write_metrics(const std::string& path, bool first, bool last, int64_t cut, double objective, int64_t t_nnz,
                   int64_t levels, double contract_s, double refine_s, double total_s, const TrialLog& log) {
    std::ofstream f(path, std::ios::app);
    if (!f.is_open()) {
        std::cerr << "Could not open " << path << std::endl;
        return;
    }
    if (first) f << "[";
    f << "{";
    f << "\"edge-cut\":" << std::fixed << cut << ",";
    f << "\"modularity\":" << objective << ",";
    f << "\"total-nnz\":" << t_nnz << ",";
    f << "\"level-count\":" << levels << ",";
    const char* names[5] = {"coarsen-contract", "coarsen", "leiden-refine", "free-graph", "total"};
    const double vals[5] = {contract_s, 0.0, refine_s, 0.0, total_s};
    for (int k = 0; k < 5; k++) f << "\"" << names[k] << "-duration-seconds\":" << vals[k] << ",";
    for (size_t k = 0; k < log.times.size(); k++) f << "\"iter-times-" << k << "\":" << log.times[k] << ",";
    for (size_t k = 0; k < log.objs.size(); k++) f << "\"iter-obj-" << k << "\":" << log.objs[k] << ",";
    f << "\"number-coarse-levels\":" << 0;
    f << "}";
    f << (last ? "]" : ",");
}

}  // namespace

int main(int argc, char** argv) {
    const cli::ClusterArgs args = cli::parse_cluster_args(argc, argv);
    if (!args.valid) return -1;
    pggc_ctx* ctx = nullptr;
    if (!cli::check(pggc_ctx_create(&ctx, nullptr), "creating the GPU context")) return -1;
    int rc = 0;
    {
        cli::DeviceGraph dg;
        if (!cli::load_device_graph(ctx, args, dg)) {
            pggc_ctx_destroy(ctx);
            return -1;
        }
        const int64_t n = dg.g.n;
        int32_t* part_d = nullptr;
        if (hipMalloc((void**)&part_d, (n > 0 ? n : 1) * sizeof(int32_t)) != hipSuccess) {
            std::cerr << "FATAL ERROR: device allocation failed" << std::endl;
            pggc_ctx_destroy(ctx);
            return -1;
        }
        std::vector<int32_t> best(n), cur(n);
        double best_obj = -std::numeric_limits<double>::infinity();
        const char* seed_env = std::getenv("PGGC_SEED");
        const uint64_t seed0 = seed_env ? std::strtoull(seed_env, nullptr, 10) : 0;
        for (int t = 0; t < args.n_trials && rc == 0; t++) {
            TrialLog log{cli::objective_label(args.obj), {}, {}};
            pggc_ctx_set_progress(ctx, on_iteration, &log);
            auto t0 = std::chrono::steady_clock::now();
            double obj = 0.0;
            int64_t info[8] = {0};
            if (!cli::check(pggc_cluster(ctx, &dg.g, PGGC_ALGO, cli::objective_code(args.obj), args.lambda_multiplier,
                                         args.n_successive_iterations, seed0 + (uint64_t)t, part_d, &obj, info),
                            "clustering")) {
                rc = -1;
                break;
            }
            const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::cout << "Total time: " << total << std::endl;
            std::cout << std::endl;
            if (obj > best_obj) {
                best_obj = obj;
                if (hipMemcpy(best.data(), part_d, n * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
                    std::cerr << "FATAL ERROR: download failed" << std::endl;
                    rc = -1;
                    break;
                }
            }
            if (!args.metrics_file.empty()) {
                double contract_s = 0.0, refine_s = 0.0;
                pggc_ctx_last_times(ctx, &contract_s, &refine_s);
                const int64_t cut = (dg.g.nnz - info[1]) / 2;
                write_metrics(args.metrics_file, t == 0, t + 1 == args.n_trials, cut, obj, info[7], info[2],
                              contract_s, refine_s, total, log);
            }
        }
        pggc_ctx_set_progress(ctx, nullptr, nullptr);
        (void)hipFree(part_d);
        if (rc == 0) {
            std::cout << std::setprecision(9) << "Best objective found: " << best_obj << std::endl;
            if (!args.output_file.empty()) {
                std::cout << "Writing best clustering to " << args.output_file << std::endl;
                if (!cli::check(pggc_write_part(args.output_file.c_str(), best.data(), n), "writing the clustering"))
                    rc = -1;
            }
        }
    }
    pggc_ctx_destroy(ctx);
    return rc;
}
