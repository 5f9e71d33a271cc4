// meme (app/memetic_clustering.cpp:118-249): a population of pLeiden+
// clusterings evolved for a time limit.  Each step picks two parents by
// binary tournament, clusters again under the intersection of the parents
// (constrained pLouvain), improves the offspring with five pLeiden+ passes
// and lets it replace the worse member whose cut-edge signature is closest.
// Every clustering runs on the GPU through the C ABI; the population
// (clusterings and signatures) stays in device memory.  The reference seeds
// its generator and the Leiden refinement from std::random_device; here both
// come from PGGC_SEED (environment, default 0).
#include <chrono>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <limits>
#include <random>
#include <vector>

#include <hip/hip_runtime.h>

#include "cli_common.h"

namespace {

struct Member {
    int32_t* clusters = nullptr;
    uint32_t* signature = nullptr;
    double obj = 0.0;
    int labels = 0;
};

double objective_of(const pggc_state& st) { return st.obj * st.scale + st.offset; }

bool dev_alloc(void** p, size_t bytes) {
    if (hipMalloc(p, bytes > 0 ? bytes : 4) == hipSuccess) return true;
    std::cerr << "FATAL ERROR: device allocation of " << bytes << " bytes failed" << std::endl;
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    const cli::MemeArgs args = cli::parse_meme_args(argc, argv);
    if (!args.valid) return -1;
    pggc_ctx* ctx = nullptr;
    if (!cli::check(pggc_ctx_create(&ctx, nullptr), "creating the GPU context")) return -1;
    const char* seed_env = std::getenv("PGGC_SEED");
    // PGGC_PROGRESS=1: a line on stderr after each improvement pass (long
    // passes on large graphs otherwise print nothing for minutes)
    const char* progress_env = std::getenv("PGGC_PROGRESS");
    const bool progress = progress_env && progress_env[0] == '1';
    const uint64_t seed = seed_env ? std::strtoull(seed_env, nullptr, 10) : 0;
    pggc_ctx_set_seed(ctx, seed, 0);
    int rc = 0;
    std::vector<Member> pop;
    int64_t* vol = nullptr;
    int32_t *cons = nullptr, *c3 = nullptr, *c3b = nullptr;
    uint32_t* sig3 = nullptr;
    {
        cli::DeviceGraph dg;
        if (!cli::load_device_graph(ctx, args, dg)) {
            pggc_ctx_destroy(ctx);
            return -1;
        }
        const pggc_graph* g = &dg.g;
        const int64_t n = g->n;
        const int64_t words = (g->nnz + 31) / 32;
        const int obj = cli::objective_code(args.obj);
        pggc_state st{};
        auto reset = [&]() {  // rfd.reset (interface/cluster_data.hpp:49-55)
            return cli::check(pggc_state_init(ctx, g, obj, args.lambda_multiplier, &st), "objective setup");
        };
        if (!dev_alloc((void**)&vol, (n + 1) * sizeof(int64_t)) || !dev_alloc((void**)&cons, n * sizeof(int32_t)) ||
            !dev_alloc((void**)&c3, n * sizeof(int32_t)) || !dev_alloc((void**)&c3b, n * sizeof(int32_t)) ||
            !dev_alloc((void**)&sig3, words * sizeof(uint32_t)))
            rc = -1;
        st.vol = vol;
        st.vol_cap = n;
        if (rc == 0 && !reset()) rc = -1;
        std::cout << std::setprecision(9) << std::fixed;
        // initial population: independent pLeiden+ runs
        for (int i = 0; rc == 0 && i < args.pop_size; i++) {
            Member m;
            if (!dev_alloc((void**)&m.clusters, n * sizeof(int32_t)) ||
                !dev_alloc((void**)&m.signature, words * sizeof(uint32_t))) {
                rc = -1;
                break;
            }
            pop.push_back(m);
            Member& y = pop.back();
            if (!cli::check(pggc_leiden(ctx, g, &st, 1, 0, nullptr, y.clusters), "pLeiden+") ||
                !cli::check(pggc_cut_signature(ctx, g, y.clusters, y.signature), "signature")) {
                rc = -1;
                break;
            }
            y.obj = objective_of(st);
            y.labels = st.label_count;
            std::cout << "Adding clustering with obj: " << y.obj << std::endl;
            if (!reset()) rc = -1;
        }
        std::default_random_engine e1((unsigned)seed);
        std::uniform_int_distribution<int> d1(0, args.pop_size - 1);
        std::uniform_int_distribution<int> d2(0, args.pop_size - 2);
        double best = 0.0;
        int epoch = 0, iter = 0;
        const auto t0 = std::chrono::steady_clock::now();
        auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
        while (rc == 0 && elapsed() < args.time_limit) {
            auto tournament = [&]() {
                int a = d1(e1);
                int b = (a + d2(e1)) % args.pop_size;
                return pop[a].obj > pop[b].obj ? a : b;
            };
            const int p1 = tournament();
            int p2 = p1;
            while (p2 == p1) p2 = tournament();
            int64_t L = 0;
            if (!cli::check(pggc_intersect(ctx, n, pop[p1].clusters, pop[p2].clusters, pop[p2].labels, cons, &L),
                            "intersection") ||
                !cli::check(pggc_louvain(ctx, g, &st, 1, cons, c3), "constrained pLouvain")) {
                rc = -1;
                break;
            }
            std::cout << "Parent 1 obj: " << pop[p1].obj << "; Parent 2 obj: " << pop[p2].obj
                      << "; Offspring obj: " << objective_of(st) << std::flush;
            for (int x = 0; x < 5 && rc == 0; x++) {
                if (!cli::check(pggc_leiden(ctx, g, &st, 1, 1, c3, c3b), "pLeiden+ improvement")) rc = -1;
                std::swap(c3, c3b);
                if (progress) std::cerr << "meme: improvement pass " << x + 1 << " of 5: " << objective_of(st) << std::endl;
            }
            if (rc) break;
            const double o3 = objective_of(st);
            std::cout << "; Post leiden obj: " << o3;
            if (o3 > best) best = o3;
            if (!cli::check(pggc_cut_signature(ctx, g, c3, sig3), "signature")) {
                rc = -1;
                break;
            }
            int am = -1;
            int64_t min_diff = std::numeric_limits<int64_t>::max();
            for (int p = 0; p < args.pop_size; p++) {
                if (pop[p].obj < o3) {
                    int64_t diff = 0;
                    if (!cli::check(pggc_signature_diff(ctx, sig3, pop[p].signature, words, &diff), "signature")) {
                        rc = -1;
                        break;
                    }
                    if (diff < min_diff) {
                        min_diff = diff;
                        am = p;
                    }
                }
            }
            if (rc) break;
            if (am != -1) {
                std::swap(pop[am].clusters, c3);
                std::swap(pop[am].signature, sig3);
                pop[am].obj = o3;
                pop[am].labels = st.label_count;
            } else {
                min_diff = 0;
            }
            std::cout << "; Min diff: " << min_diff << std::endl;
            if (!reset()) {
                rc = -1;
                break;
            }
            if (++iter >= args.pop_size) {
                std::cout << "Epoch " << epoch << " best objective: " << best << std::endl;
                epoch++;
                iter = 0;
            }
        }
        if (rc == 0 && !pop.empty()) {
            int am = 0;
            for (int p = 1; p < (int)pop.size(); p++)
                if (pop[p].obj > pop[am].obj) am = p;
            std::cout << "Final objective: " << pop[am].obj << std::endl;
            if (!args.output_file.empty()) {
                std::vector<int32_t> h(n);
                std::cout << "Writing best clustering to " << args.output_file << std::endl;
                if (hipMemcpy(h.data(), pop[am].clusters, n * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
                    !cli::check(pggc_write_part(args.output_file.c_str(), h.data(), n), "writing the clustering"))
                    rc = -1;
            }
        }
        for (auto& m : pop) {
            (void)hipFree(m.clusters);
            (void)hipFree(m.signature);
        }
        (void)hipFree(vol);
        (void)hipFree(cons);
        (void)hipFree(c3);
        (void)hipFree(c3b);
        (void)hipFree(sig3);
    }
    pggc_ctx_destroy(ctx);
    return rc;
}
