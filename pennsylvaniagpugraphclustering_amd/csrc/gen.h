// Device-side graph construction and synthetic generators (gen.hip).
#pragma once
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
// out[i] = Philox(counter (i, 0, 0, 2), key seed) word 0 mod L
void random_labels(Ctx& c, i64 n, i64 L, u64 seed, i32* out);

}  // namespace pggc
