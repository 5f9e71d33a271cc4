// verify_obj (app/verify_objective.cpp): reads a graph and a clusters file
// (one label per line, labels below the vertex count) and prints the
// objective of that clustering, "<label>: <value>" with 9 significant digits.
// The volumes, intra-cluster weight and exact sum of squared volumes come
// from the GPU evaluator; lambda and the normalisation from pggc_state_init.
#include <cmath>
#include <iomanip>
#include <iostream>
#include <vector>

#include <hip/hip_runtime.h>

#include "cli_common.h"

int main(int argc, char** argv) {
    const cli::VerifyArgs args = cli::parse_verify_args(argc, argv);
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
        std::vector<int64_t> raw(n);
        std::vector<int32_t> labels(n);
        int64_t* vol = nullptr;
        int32_t* lab_d = nullptr;
        if (!cli::check(pggc_read_ints(args.cluster_file.c_str(), n, raw.data()), "reading the clusters")) rc = -1;
        for (int64_t i = 0; rc == 0 && i < n; i++) {
            if (raw[i] < 0 || raw[i] >= n) {
                std::cerr << "FATAL ERROR: cluster label " << raw[i] << " of vertex " << i << " is outside [0, " << n
                          << ")" << std::endl;
                rc = -1;
            }
            labels[i] = (int32_t)raw[i];
        }
        if (rc == 0 && (hipMalloc((void**)&vol, (n + 1) * sizeof(int64_t)) != hipSuccess ||
                        hipMalloc((void**)&lab_d, (n > 0 ? n : 1) * sizeof(int32_t)) != hipSuccess ||
                        hipMemcpy(lab_d, labels.data(), n * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)) {
            std::cerr << "FATAL ERROR: device allocation or upload failed" << std::endl;
            rc = -1;
        }
        pggc_state st{};
        int64_t uncut = 0;
        uint64_t sq[2] = {0, 0};
        if (rc == 0) {
            st.vol = vol;
            st.vol_cap = n;
            if (!cli::check(pggc_state_init(ctx, &dg.g, cli::objective_code(args.obj), args.lambda_multiplier, &st),
                            "objective setup") ||
                !cli::check(pggc_evaluate(ctx, &dg.g, lab_d, n, vol, &uncut, sq), "evaluation"))
                rc = -1;
        }
        if (rc == 0) {
            // cluster_data::update_objective (interface/cluster_data.hpp:76-87) with the exact sum
            const unsigned __int128 s = ((unsigned __int128)sq[0] << 64) | (unsigned __int128)sq[1];
            double obj = (double)uncut;
            obj -= st.lambda * (double)s;
            const double value = obj * st.scale + st.offset;
            std::cout << std::setprecision(9) << cli::objective_label(args.obj) << ": " << value << std::endl;
        }
        (void)hipFree(vol);
        (void)hipFree(lab_d);
    }
    pggc_ctx_destroy(ctx);
    return rc;
}
