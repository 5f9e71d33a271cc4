/*
 * pggc_io.h -- host-side graph and vector files of the PGGC command-line
 * tools (no GPU calls).
 *
 * Replaces the reference's readers and writers
 *   load_graph / load_metis_graph / load_mtx_graph  (app/io/io_mtx.cpp:337-346,
 *                                                    app/io/io_metis.cpp:66-181,
 *                                                    app/io/io_mtx.cpp:221-335)
 *   write_part / load_view                          (app/io/header/io_views.hpp:11-41)
 * with plain C entry points over host memory.  Status codes are those of
 * pggc.h; pggc_last_error() gives the message.  Informational lines go to
 * stdout with the reference's wording.
 */
#ifndef PGGC_IO_H
#define PGGC_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pggc_host_graph pggc_host_graph;

/* Reads a graph file: a name ending in ".mtx" is Matrix Market, anything
 * else METIS.  unit_edges != 0 ignores edge weights (objectives Mod and CPM,
 * app/objective_helpers.hpp:24-29); *weighted_out tells whether the graph
 * kept them.  Matrix Market entries become a symmetric CSR without self
 * loops, zero-weight edges or repeated columns, rows sorted by column
 * (app/io/io_mtx.cpp:86-192); METIS rows are taken in file order
 * (app/io/io_metis.cpp:128-165), self loops and repeated columns are dropped
 * with a warning (the clustering kernels assume neither). */
int pggc_read_graph(const char* path, int unit_edges, pggc_host_graph** out, int* weighted_out);
/* sizes of a read graph */
int pggc_host_graph_info(const pggc_host_graph* g, int64_t* n, int64_t* nnz);
/* copies the CSR into caller buffers: row_ptr[n+1], col[nnz], w[nnz] (w may
 * be NULL; unit weights read as 1) */
int pggc_host_graph_copy(const pggc_host_graph* g, int64_t* row_ptr, int32_t* col, int32_t* w);
int pggc_host_graph_free(pggc_host_graph* g);

/* write_part: one label per line */
int pggc_write_part(const char* path, const int32_t* part, int64_t n);
/* load_view: n whitespace-separated integers (cluster or vertex-weight files) */
int pggc_read_ints(const char* path, int64_t n, int64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* PGGC_IO_H */
