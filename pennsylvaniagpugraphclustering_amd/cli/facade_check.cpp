// Exercises the C++ facade include/pggc_clustering_methods.hpp as a caller
// of the reference's clustering_methods would: a graph held as plain device
// pointers (here an R-MAT graph from the library's generator, optionally
// with its int32 weights widened to int64), vertex weights and the initial
// state from make_state, then louvain_part<false> and leiden_part<true,false>.
// Prints one JSON line per call: objective (exact, as a hex double), label
// count and an FNV-1a hash of the partition.
//
//   facade_check <scale> <edge_factor> <seed> <weights: none|int32|int64>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pggc_clustering_methods.hpp"

namespace cm = jet_community::clustering_methods;

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

static unsigned long long fnv1a(const std::vector<int32_t>& v) {
    unsigned long long h = 1469598103934665603ull;
    const unsigned char* p = reinterpret_cast<const unsigned char*>(v.data());
    for (size_t i = 0; i < v.size() * sizeof(int32_t); i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

static void report(const char* call, const pggc_state& st, const int32_t* part_dev, int64_t n) {
    std::vector<int32_t> h((size_t)n);
    hip_check(hipMemcpy(h.data(), part_dev, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost), "readback");
    std::printf("{\"call\": \"%s\", \"objective\": \"%a\", \"labels\": %d, \"fnv\": %llu}\n", call,
                st.obj * st.scale + st.offset, st.label_count, fnv1a(h));
    std::fflush(stdout);
}

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: facade_check <scale> <edge_factor> <seed> <none|int32|int64>\n");
        return 2;
    }
    const int scale = std::atoi(argv[1]), ef = std::atoi(argv[2]);
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    const std::string wk = argv[4];
    try {
        pggc_ctx* ctx = nullptr;
        cm::check(pggc_ctx_create(&ctx, nullptr));
        pggc_devgraph* dg = nullptr;
        const bool weighted = wk != "none";
        cm::check(pggc_rmat(ctx, scale, ef, 0.57, 0.19, 0.19, seed, weighted ? 1 : 0, weighted ? 100 : 0, &dg));
        pggc_graph v;
        cm::check(pggc_devgraph_view(dg, &v));
        cm::Graph g;
        g.n = v.n;
        g.nnz = v.nnz;
        g.row_ptr = v.row_ptr;
        g.col = v.col;
        int64_t* w64 = nullptr;
        if (wk == "int32") {
            g.w = v.w;
        } else if (wk == "int64") {  // the caller holds 64-bit weights
            std::vector<int32_t> h32((size_t)v.nnz);
            hip_check(hipMemcpy(h32.data(), v.w, (size_t)v.nnz * sizeof(int32_t), hipMemcpyDeviceToHost), "w");
            std::vector<int64_t> h64(h32.begin(), h32.end());
            hip_check(hipMalloc(&w64, (size_t)v.nnz * sizeof(int64_t)), "alloc w64");
            hip_check(hipMemcpy(w64, h64.data(), (size_t)v.nnz * sizeof(int64_t), hipMemcpyHostToDevice), "w64");
            g.w64 = w64;
        }
        const int objective = weighted ? PGGC_OBJ_WMOD : PGGC_OBJ_MOD;
        int64_t *vtx_w = nullptr, *vol = nullptr;
        int32_t *part = nullptr, *part2 = nullptr;
        hip_check(hipMalloc(&vtx_w, (size_t)g.n * sizeof(int64_t)), "alloc");
        hip_check(hipMalloc(&vol, (size_t)(g.n + 1) * sizeof(int64_t)), "alloc");
        hip_check(hipMalloc(&part, (size_t)g.n * sizeof(int32_t)), "alloc");
        hip_check(hipMalloc(&part2, (size_t)g.n * sizeof(int32_t)), "alloc");

        pggc_state st = cm::make_state(ctx, g, objective, 1.0, vtx_w, vol, g.n);
        cm::louvain_part<false>(ctx, g, st, nullptr, part);
        report("louvain_part", st, part, g.n);

        cm::check(pggc_ctx_set_seed(ctx, 7, 0));
        st = cm::make_state(ctx, g, objective, 1.0, vtx_w, vol, g.n);
        cm::leiden_part<true, false>(ctx, g, st, nullptr, part2);
        report("leiden_part_plus", st, part2, g.n);

        (void)hipFree(part2);
        (void)hipFree(part);
        (void)hipFree(vol);
        (void)hipFree(vtx_w);
        if (w64) (void)hipFree(w64);
        cm::check(pggc_devgraph_free(dg));
        cm::check(pggc_ctx_destroy(ctx));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
