// Device-side graph construction and synthetic generators (gen.hip).
#pragma once
#include <vector>

#include "ctx.h"

namespace pggc {

// Symmetric CSR in device memory owned by its holder (hipMalloc'd; release()
// frees it): rp int64[n+1], col int32[nnz], w int32[nnz] or null.
struct DevCSR {
    i64 n = 0, nnz = 0;
    i64* rp = nullptr;
    i32* col = nullptr;
    i32* w = nullptr;
    void release();
};

void rmat_thresholds(double a, double b, double c, u64 t[3]);
// R-MAT graph of edge_factor * 2^scale Philox draws; weights uniform in
// [wmin, wmax] per undirected pair when wmax > 0.
void rmat_csr(Ctx& c, int scale, int edge_factor, double a, double b, double cc, u64 seed, i32 wmin, i32 wmax,
              DevCSR& out);
// construct_from_edgelist (app/io/io_mtx.cpp:95-201) of a device edge list
void edges_csr(Ctx& c, i64 n, i64 m, const i32* src, const i32* dst, const i32* w, DevCSR& out);
// LFR-style benchmark graph (Lancichinetti, Fortunato, Radicchi, Phys. Rev. E
// 78, 046110, 2008): power-law degrees (exponent tau1, mean kmean, at most
// kmax) and community sizes (exponent tau2, in [cmin, cmax]), a fraction mu
// of each vertex's edges leaving its community, configuration-model stub
// matching inside each community and across the graph, randomly permuted
// vertex ids, weights uniform in [wmin, wmax] per undirected pair when
// wmax > 0.  gt (device, nullable, n entries) receives each vertex's planted
// community.
struct LfrParams {
    i64 n = 0;
    double tau1 = 2.5, kmean = 20, tau2 = 1.5, mu = 0.4;
    i32 kmax = 1000, cmin = 20, cmax = 2000;
    i32 wmin = 0, wmax = 0;
    u64 seed = 0;
};
// Integer values of a power law: X with density proportional to x^-tau on
// [x0, top + 1), value floor(X); draw r (u32) maps to k0 + the first i with
// r < thr[i].  x0 is solved so that the mean value is `mean` when mean > 0,
// else x0 = lo.
struct PowerLawTable {
    i32 k0 = 0;
    std::vector<u64> thr;
};
PowerLawTable power_law_table(double tau, i32 lo, i32 top, double mean);
// community sizes of an LFR graph on n vertices (host)
std::vector<i64> lfr_community_sizes(const LfrParams& p);
void lfr_csr(Ctx& c, const LfrParams& p, DevCSR& out, i32* gt);
// out[i] = Philox(counter (i, 0, 0, 2), key seed) word 0 mod L
void random_labels(Ctx& c, i64 n, i64 L, u64 seed, i32* out);

}  // namespace pggc
