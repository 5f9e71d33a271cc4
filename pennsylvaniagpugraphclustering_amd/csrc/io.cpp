// Host-side graph and vector files of the command-line tools: METIS and
// Matrix Market readers (app/io/io_metis.cpp, app/io/io_mtx.cpp), the
// one-label-per-line partition writer and the integer-vector reader
// (app/io/header/io_views.hpp).  Plain C++; no GPU calls.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "pggc.h"
#include "pggc_io.h"

void set_last_error(const std::string& msg);

struct pggc_host_graph {
    int64_t n = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> col;
    std::vector<int32_t> w;  // empty = unit weights
};

namespace {

struct Fail {
    int code;
    std::string msg;
};

template <class F>
int guarded(F f) {
    try {
        f();
        return PGGC_OK;
    } catch (const Fail& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return PGGC_ENOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return PGGC_EINTERNAL;
    }
}

// Whole file plus a trailing newline, so that parsers can always stop at one.
std::string slurp(const std::string& path, const char* what) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw Fail{PGGC_EINVAL, std::string("Could not open ") + what + " file " + path};
    in.seekg(0, std::ios::end);
    const std::streamoff sz = in.tellg();
    in.seekg(0, std::ios::beg);
    std::string s((size_t)sz, '\0');
    in.read(&s[0], sz);
    std::cout << "Reading " << sz << " bytes from " << path << std::endl;
    s.push_back('\n');
    return s;
}

void skip_line(const char*& f, const char* end) {
    while (f < end && *f != '\n') f++;
    if (f < end) f++;
}

uint64_t read_uint(const char*& f, const char* end) {
    uint64_t v = 0;
    while (f < end && std::isdigit((unsigned char)*f)) v = v * 10 + (uint64_t)(*f++ - '0');
    return v;
}

// Reads up to `cnt` numbers from the current line (the header line of both
// formats) and leaves f after its newline.
int header_numbers(const char*& f, const char* end, uint64_t* out, int cnt) {
    while (f < end && *f == '%') skip_line(f, end);
    while (f < end && !std::isdigit((unsigned char)*f)) f++;
    int k = 0;
    while (f < end && *f != '\n') {
        if (std::isdigit((unsigned char)*f)) {
            const uint64_t v = read_uint(f, end);
            if (k < cnt) out[k] = v;
            k++;
        } else {
            f++;
        }
    }
    if (f < end) f++;
    return k;
}

// Rows taken as given by the METIS reader: drop self loops and repeated
// columns (keeping the first), which the clustering kernels do not accept.
void clean_rows(pggc_host_graph& g) {
    int64_t loops = 0, dups = 0, o = 0;
    std::vector<int64_t> rp(g.n + 1, 0);
    std::unordered_set<int32_t> seen;
    for (int64_t i = 0; i < g.n; i++) {
        seen.clear();
        for (int64_t j = g.rp[i]; j < g.rp[i + 1]; j++) {
            const int32_t v = g.col[j];
            if (v == i) {
                loops++;
                continue;
            }
            if (!seen.insert(v).second) {
                dups++;
                continue;
            }
            g.col[o] = v;
            if (!g.w.empty()) g.w[o] = g.w[j];
            o++;
        }
        rp[i + 1] = o;
    }
    if (loops) std::cout << "WARNING: Ignoring " << loops << " self loop entries." << std::endl;
    if (dups) std::cout << "WARNING: Ignoring " << dups << " repeated entries." << std::endl;
    g.rp.swap(rp);
    g.col.resize(o);
    if (!g.w.empty()) g.w.resize(o);
}

// load_metis_graph (app/io/io_metis.cpp:66-181)
void read_metis(const std::string& path, bool& unit_edges, pggc_host_graph& g) {
    const std::string s = slurp(path, "metis graph");
    std::cout << "Parsing file as metis format" << std::endl;
    const char* f = s.data();
    const char* end = f + s.size();
    uint64_t hdr[4] = {0, 0, 0, 0};
    header_numbers(f, end, hdr, 4);
    const int64_t n = (int64_t)hdr[0], m = (int64_t)hdr[1];
    const uint64_t fmt = hdr[2], ncon = hdr[3];
    if (fmt != 0 && fmt != 1)
        throw Fail{PGGC_EINVAL, "Unsupported format flags " + std::to_string(fmt) +
                                    " (the graph parser does not support vertex weights)"};
    if (ncon != 0)
        throw Fail{PGGC_EINVAL, "Unsupported ncon " + std::to_string(ncon) +
                                    " (the graph parser does not support vertex weights)"};
    const bool has_ew = fmt == 1;
    const bool keep_w = has_ew && !unit_edges;
    if (has_ew && unit_edges)
        std::cout << "INFO: This file has edge weights. The given objective function treats edge weights as unit "
                     "uniform. Please see Objectives.md for more information."
                  << std::endl;
    if (!has_ew) unit_edges = true;
    g.n = n;
    g.rp.assign(n + 1, 0);
    g.col.resize(2 * m);
    if (keep_w) g.w.resize(2 * m);
    int64_t rows = 0, nz = 0;
    bool value_next = false;
    while (f < end) {
        while (f < end && !std::isdigit((unsigned char)*f)) {
            if (*f == '%') {
                skip_line(f, end);
                continue;
            }
            if (*f == '\n' && rows < n) g.rp[++rows] = nz;
            if (*f == '-') throw Fail{PGGC_EINVAL, "Negative value in metis file (edge weights must be >= 0)"};
            f++;
        }
        if (f >= end) break;
        const uint64_t x = read_uint(f, end);
        if (nz >= 2 * m) throw Fail{PGGC_EINVAL, "More nonzeros than the header's " + std::to_string(2 * m)};
        if (!value_next) {
            if (x < 1 || (int64_t)x > n) throw Fail{PGGC_EINVAL, "Vertex id " + std::to_string(x) + " out of range"};
            g.col[nz] = (int32_t)(x - 1);
        } else if (keep_w) {
            if (x > (uint64_t)INT32_MAX)
                throw Fail{PGGC_EINVAL, "Edge weight " + std::to_string(x) + " exceeds 2147483647"};
            g.w[nz] = (int32_t)x;
        }
        if (has_ew) {
            if (value_next) nz++;
            value_next = !value_next;
        } else {
            nz++;
        }
    }
    if (rows != n || nz != 2 * m) {
        std::ostringstream e;
        e << "Mismatch between expected and actual line/nonzero count in metis file: read " << rows << " lines and "
          << nz << " nonzeros; expected " << n << " lines and " << 2 * m << " nonzeros";
        throw Fail{PGGC_EINVAL, e.str()};
    }
    clean_rows(g);
}

// load_mtx_graph and construct_from_edgelist (app/io/io_mtx.cpp:86-335)
void read_mtx(const std::string& path, bool& unit_edges, pggc_host_graph& g) {
    const std::string s = slurp(path, "Matrix Market graph");
    std::cout << "Parsing file as matrix market format" << std::endl;
    const char* f = s.data();
    const char* end = f + s.size();
    // banner: %%MatrixMarket matrix <format> <field> <symmetry>
    std::vector<std::string> fields;
    {
        const char* b = f;
        while (f < end && *f != '\n') f++;
        std::string line(b, f);
        std::string tok;
        for (char ch : line) {
            if (std::isalnum((unsigned char)ch)) tok.push_back(ch);
            else if (!tok.empty()) {
                fields.push_back(tok);
                tok.clear();
            }
        }
        if (!tok.empty()) fields.push_back(tok);
        if (f < end) f++;
    }
    auto field = [&](size_t k) { return k < fields.size() ? fields[k] : std::string(); };
    const bool coordinate = field(2) == "coordinate";
    const bool pattern = field(3) == "pattern";
    const bool integer = field(3) == "integer";
    const bool symmetric = field(4) == "symmetric";
    if (!coordinate)
        throw Fail{PGGC_EINVAL, "Importing of non-coordinate format matrix market files is not currently supported!"};
    if (!pattern) {
        if (integer) {
            if (unit_edges)
                std::cout << "INFO: This file has edge weights. The given objective function treats edge weights as "
                             "unit uniform. Please see Objectives.md for more information."
                          << std::endl;
        } else if (!unit_edges) {
            unit_edges = true;
            std::cout << "WARNING: Non-integer edge weights given in matrix market file. This is not currently "
                         "supported. The given weights will be ignored."
                      << std::endl;
        }
    } else {
        unit_edges = true;
    }
    if (!symmetric) {
        std::cout << "INFO: Non-symmetric matrix market file detected. The graph will be automatically symmetrized."
                  << std::endl;
        if (!unit_edges) {
            unit_edges = true;
            std::cout << "WARNING: Edge weights given in matrix market file. This is not supported for non-symmetric "
                         "file formats. The given weights will be ignored."
                      << std::endl;
        }
    }
    uint64_t hdr[3] = {0, 0, 0};
    header_numbers(f, end, hdr, 3);
    if (hdr[0] != hdr[1])
        throw Fail{PGGC_EINVAL, "Bipartite file format detected. This format is not currently supported!"};
    const int64_t n = (int64_t)hdr[0], m = (int64_t)hdr[2];
    std::vector<int32_t> src(m), dst(m), val(unit_edges ? 0 : m);
    int64_t read = 0;
    int state = 0;  // 0 source, 1 destination, 2 weight
    while (f < end) {
        while (f < end && !std::isdigit((unsigned char)*f)) {
            if (*f == '%') {
                skip_line(f, end);
                continue;
            }
            if (*f == '\n' && state != 0) throw Fail{PGGC_EINVAL, "Line with exactly one value detected"};
            if (*f == '-') throw Fail{PGGC_EINVAL, "Negative value in matrix market file (edge weights must be >= 0)"};
            f++;
        }
        if (f >= end) break;
        if (read >= m) throw Fail{PGGC_EINVAL, "More entries than the header's " + std::to_string(m)};
        if (state == 2) {
            const uint64_t x = read_uint(f, end);
            if (x > (uint64_t)INT32_MAX)
                throw Fail{PGGC_EINVAL, "Edge weight " + std::to_string(x) + " exceeds 2147483647"};
            val[read++] = (int32_t)x;
            state = 0;
            skip_line(f, end);
            continue;
        }
        const uint64_t x = read_uint(f, end);
        if (x < 1 || (int64_t)x > n) throw Fail{PGGC_EINVAL, "Vertex id " + std::to_string(x) + " out of range"};
        if (state == 0) {
            src[read] = (int32_t)(x - 1);
            state = 1;
        } else {
            dst[read] = (int32_t)(x - 1);
            if (unit_edges) {
                read++;
                state = 0;
                skip_line(f, end);
            } else {
                state = 2;
            }
        }
    }
    if (read != m)
        throw Fail{PGGC_EINVAL, "Mismatch between expected and actual nonzero count in matrix market file: read " +
                                    std::to_string(read) + ", expected " + std::to_string(m)};
    std::cout << "Finished reading data from " << path << std::endl;
    std::cout << "Beginning conversion to CSR format" << std::endl;
    // undirected edges: no self loops, no zero weights, each unordered pair once
    // (the first occurrence wins; non-symmetric files list a pair in both
    // directions, symmetric files once)
    std::vector<uint64_t> keys;  // unordered pair of each kept entry
    std::vector<int64_t> pos;    // its position in the file
    keys.reserve(m);
    pos.reserve(m);
    for (int64_t k = 0; k < m; k++) {
        const int32_t u = src[k], v = dst[k];
        if (u == v) continue;
        if (!unit_edges && val[k] == 0) continue;
        const uint64_t lo = (uint64_t)std::min(u, v), hi = (uint64_t)std::max(u, v);
        keys.push_back(lo << 32 | hi);
        pos.push_back(k);
    }
    // keep the first occurrence of each pair: stable sort by pair
    std::vector<int64_t> idx(pos.size());
    for (size_t k = 0; k < idx.size(); k++) idx[k] = (int64_t)k;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
    std::vector<int64_t> keep;
    keep.reserve(idx.size());
    int64_t dups = 0;
    for (size_t k = 0; k < idx.size(); k++) {
        if (k > 0 && keys[idx[k]] == keys[idx[k - 1]]) {
            dups++;
            continue;
        }
        keep.push_back(pos[idx[k]]);
    }
    if (dups && symmetric) std::cout << "WARNING: Ignoring " << dups << " repeated entries." << std::endl;
    g.n = n;
    g.rp.assign(n + 1, 0);
    for (int64_t k : keep) {
        g.rp[src[k] + 1]++;
        g.rp[dst[k] + 1]++;
    }
    for (int64_t i = 0; i < n; i++) g.rp[i + 1] += g.rp[i];
    const int64_t nnz = g.rp[n];
    g.col.resize(nnz);
    if (!unit_edges) g.w.resize(nnz);
    std::vector<int64_t> fill(g.rp.begin(), g.rp.end() - 1);
    for (int64_t k : keep) {
        const int32_t u = src[k], v = dst[k];
        const int64_t a = fill[u]++, b = fill[v]++;
        g.col[a] = v;
        g.col[b] = u;
        if (!unit_edges) g.w[a] = g.w[b] = val[k];
    }
    // rows sorted by column (sort_crs_matrix in the reference)
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int64_t i = 0; i < n; i++) {
        const int64_t a = g.rp[i], b = g.rp[i + 1];
        row.clear();
        for (int64_t j = a; j < b; j++) row.emplace_back(g.col[j], unit_edges ? 1 : g.w[j]);
        std::sort(row.begin(), row.end());
        for (int64_t j = a; j < b; j++) {
            g.col[j] = row[j - a].first;
            if (!unit_edges) g.w[j] = row[j - a].second;
        }
    }
}

bool ends_with(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

}  // namespace

extern "C" {

int pggc_read_graph(const char* path, int unit_edges, pggc_host_graph** out, int* weighted_out) {
    return guarded([&] {
        if (!path || !out) throw Fail{PGGC_EINVAL, "null argument"};
        auto t0 = std::chrono::steady_clock::now();
        auto* g = new pggc_host_graph();
        bool unit = unit_edges != 0;
        try {
            const std::string p(path);
            if (ends_with(p, ".mtx")) read_mtx(p, unit, *g);
            else read_metis(p, unit, *g);
        } catch (...) {
            delete g;
            throw;
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Read graph from " << path << " in " << std::setprecision(3) << secs << "s" << std::endl;
        std::cout << std::setprecision(6);
        *out = g;
        if (weighted_out) *weighted_out = g->w.empty() ? 0 : 1;
    });
}

int pggc_host_graph_info(const pggc_host_graph* g, int64_t* n, int64_t* nnz) {
    return guarded([&] {
        if (!g) throw Fail{PGGC_EINVAL, "null graph"};
        if (n) *n = g->n;
        if (nnz) *nnz = g->rp.empty() ? 0 : g->rp.back();
    });
}

int pggc_host_graph_copy(const pggc_host_graph* g, int64_t* row_ptr, int32_t* col, int32_t* w) {
    return guarded([&] {
        if (!g) throw Fail{PGGC_EINVAL, "null graph"};
        if (row_ptr) std::copy(g->rp.begin(), g->rp.end(), row_ptr);
        if (col) std::copy(g->col.begin(), g->col.end(), col);
        if (w) {
            if (g->w.empty()) std::fill(w, w + g->col.size(), 1);
            else std::copy(g->w.begin(), g->w.end(), w);
        }
    });
}

int pggc_host_graph_free(pggc_host_graph* g) {
    delete g;
    return PGGC_OK;
}

int pggc_write_part(const char* path, const int32_t* part, int64_t n) {
    return guarded([&] {
        if (!path || (!part && n > 0)) throw Fail{PGGC_EINVAL, "null argument"};
        std::ofstream out(path);
        if (!out.is_open()) throw Fail{PGGC_EINVAL, std::string("Could not open ") + path};
        std::string buf;
        buf.reserve((size_t)n * 8);
        char tmp[16];
        for (int64_t i = 0; i < n; i++) {
            const int len = snprintf(tmp, sizeof(tmp), "%d\n", part[i]);
            buf.append(tmp, (size_t)len);
        }
        out << buf;
    });
}

int pggc_read_ints(const char* path, int64_t n, int64_t* out) {
    return guarded([&] {
        if (!path || (!out && n > 0)) throw Fail{PGGC_EINVAL, "null argument"};
        std::ifstream in(path);
        if (!in.is_open()) throw Fail{PGGC_EINVAL, std::string("Could not open ") + path};
        for (int64_t i = 0; i < n; i++) {
            if (!(in >> out[i]))
                throw Fail{PGGC_EINVAL, std::string(path) + " holds fewer than " + std::to_string(n) + " values"};
        }
    });
}

}  // extern "C"
