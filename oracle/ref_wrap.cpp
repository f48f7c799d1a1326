// ref_wrap.cpp -- gives C linkage to functions of the reference's own
// non-Kokkos sgpar.h (cmg/sgpar.h, serial branch: no _KOKKOS, no _OPENMP) so
// the tests can pin the CPU oracle against reference code.  The header is
// compiled where it lies under the reference tree (see Makefile); no
// reference source is copied here.  TEST INFRASTRUCTURE ONLY.
//
// Functions reached:
//   sgp_pcg32_random_r           sgpar.h:84-92
//   heap_deduplicate             sgpar.h:811-908
//   hashmap_deduplicate          sgpar.h:911-938
//   sgp_build_coarse_graph_msd   sgpar.h:1223-1345 (serial)
//   sgp_power_iter               sgpar.h:1790-1948 (serial)
//   sgp_power_iter_eigenvalue_log sgpar.h:1758-1788 (serial, prints only)
//   sgp_compute_partition        sgpar.h:1969-2087 (shared with the Kokkos build)
//   sgp_load_graph / sgp_free_graph / sgp_load_config / sgp_use_partition /
//   compute_partition_edit_distance   sgpar.h:2262-2361, 2666-2787
//
// The serial power iteration reports its iteration count only on stdout
// ("number of iterations: N"), so ref_power_iter captures stdout around the
// call and parses the last such line.
#define SGPAR_IMPLEMENTATION
#include "sgpar.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>

using namespace sgpar;

extern "C" {

uint32_t ref_pcg32(uint64_t* state, uint64_t inc)
{
    sgp_pcg32_random_t r;
    r.state = *state;
    r.inc = inc;
    uint32_t x = sgp_pcg32_random_r(&r);
    *state = r.state;
    return x;
}

void ref_heap_dedupe(uint32_t* offset_bottom, uint32_t* dest, uint32_t* wgt,
                     uint32_t* edges_per_source, uint32_t* total)
{
    heap_deduplicate(offset_bottom, dest, wgt, edges_per_source, total);
}

void ref_hashmap_dedupe(uint32_t* offset_bottom, uint32_t* dest, uint32_t* wgt,
                        uint32_t* edges_per_source, uint32_t* total)
{
    hashmap_deduplicate(offset_bottom, dest, wgt, edges_per_source, total);
}

static sgp_graph_t make_graph(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                              const uint32_t* w, std::vector<uint32_t>& eps,
                              std::vector<uint32_t>& wdeg, std::vector<uint32_t>& ew)
{
    sgp_graph_t g;
    std::memset(&g, 0, sizeof(g));
    g.nvertices = n;
    g.nedges = rowptr[n] / 2;
    g.source_offsets = const_cast<uint32_t*>(rowptr);
    g.destination_indices = const_cast<uint32_t*>(col);
    eps.resize(n + 1);
    wdeg.resize(n + 1);
    ew.resize(rowptr[n] + 1);
    for (uint32_t i = 0; i < n; i++) {
        eps[i] = rowptr[i + 1] - rowptr[i];
        uint32_t d = 0;
        for (uint32_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            ew[e] = w ? w[e] : 1u;
            d += ew[e];
        }
        wdeg[i] = d;
    }
    g.edges_per_source = eps.data();
    g.weighted_degree = wdeg.data();
    g.eweights = ew.data();
    return g;
}

// Serial reference power iteration.  final = 0 uses the weighted graph
// (weighted degree, edge weights); final = 1 the topology-only finest level.
// Returns the iteration count; u is updated in place.
long ref_power_iter(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                    const uint32_t* w, double* u, int normLap, int final_level, double tol)
{
    std::vector<uint32_t> eps, wdeg, ew;
    sgp_graph_t g = make_graph(n, rowptr, col, w, eps, wdeg, ew);
    change_tol(tol);
    std::fflush(stdout);
    FILE* tmp = std::tmpfile();
    if (!tmp) return -1;
    int saved = dup(1);
    dup2(fileno(tmp), 1);
    sgp_power_iter(u, g, normLap, final_level);
    std::fflush(stdout);
    dup2(saved, 1);
    close(saved);
    std::rewind(tmp);
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), tmp)) > 0) text.append(buf, k);
    std::fclose(tmp);
    const std::string key = "number of iterations: ";
    size_t at = text.rfind(key);
    if (at == std::string::npos) return -1;
    return std::strtol(text.c_str() + at + key.size(), nullptr, 10);
}

// Serial reference eigenvalue log (sgpar.h:1758-1788, topology only); it
// only prints, so stdout is captured and out = {eigenvalue, min, max} is
// parsed from its line (9 decimals).  Returns 0 on success.
int ref_eigenvalue_log(uint32_t n, const uint32_t* rowptr, const uint32_t* col, const double* u, double* out)
{
    std::vector<uint32_t> eps, wdeg, ew;
    sgp_graph_t g = make_graph(n, rowptr, col, nullptr, eps, wdeg, ew);
    std::vector<double> uu(u, u + n);
    std::fflush(stdout);
    FILE* tmp = std::tmpfile();
    if (!tmp) return -1;
    int saved = dup(1);
    dup2(fileno(tmp), 1);
    sgp_power_iter_eigenvalue_log(uu.data(), g);
    std::fflush(stdout);
    dup2(saved, 1);
    close(saved);
    std::rewind(tmp);
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), tmp)) > 0) text.append(buf, k);
    std::fclose(tmp);
    const size_t at = text.rfind("eigenvalue = ");
    if (at == std::string::npos) return -1;
    return std::sscanf(text.c_str() + at, "eigenvalue = %lf (%lf %lf)", &out[0], &out[1], &out[2]) == 3 ? 0 : -1;
}

// Serial reference coarse build for a level > 1 (weighted) with hashmap
// dedupe on every row.  Writes the coarse graph compacted into
// out_rowptr[nc+1], out_col, out_w (capacity rowptr[n]); returns m_c.
long ref_build_coarse(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                      const uint32_t* w, uint32_t* vcmap, uint32_t nc,
                      uint32_t* out_rowptr, uint32_t* out_col, uint32_t* out_w)
{
    std::vector<uint32_t> eps, wdeg, ew;
    sgp_graph_t g = make_graph(n, rowptr, col, w, eps, wdeg, ew);
    sgp_graph_t gc;
    std::memset(&gc, 0, sizeof(gc));
    gc.nvertices = nc;
    double times[8] = {0};
    if (sgp_build_coarse_graph_msd(&gc, vcmap, g, 2, times, 0) != 0) return -1;
    long mc = 0;
    out_rowptr[0] = 0;
    for (uint32_t c = 0; c < nc; c++) {
        for (uint32_t k = 0; k < gc.edges_per_source[c]; k++) {
            out_col[mc] = gc.destination_indices[gc.source_offsets[c] + k];
            out_w[mc] = gc.eweights[gc.source_offsets[c] + k];
            mc++;
        }
        out_rowptr[c + 1] = (uint32_t)mc;
    }
    sgp_free_graph(&gc);
    return mc;
}

int ref_compute_partition(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                          const double* evec, int num_partitions, int perc_imbalance,
                          uint32_t* part, uint32_t* edgecut)
{
    sgp_graph_t g;
    std::memset(&g, 0, sizeof(g));
    g.nvertices = n;
    g.nedges = rowptr[n] / 2;
    g.source_offsets = const_cast<uint32_t*>(rowptr);
    g.destination_indices = const_cast<uint32_t*>(col);
    return sgp_compute_partition(part, num_partitions, edgecut, perc_imbalance, 0,
                                 const_cast<double*>(evec), g);
}

// Loads a binary CSR with the reference loader; returns an opaque graph.
void* ref_load_graph(const char* path)
{
    sgp_graph_t* g = new sgp_graph_t;
    std::memset(g, 0, sizeof(*g));
    if (sgp_load_graph(g, const_cast<char*>(path)) != 0) {
        delete g;
        return nullptr;
    }
    return g;
}

void ref_graph_info(void* h, uint32_t* n, uint32_t* nedges, uint32_t** rowptr, uint32_t** col)
{
    sgp_graph_t* g = static_cast<sgp_graph_t*>(h);
    *n = g->nvertices;
    *nedges = g->nedges;
    *rowptr = g->source_offsets;
    *col = g->destination_indices;
}

void ref_graph_free(void* h)
{
    sgp_graph_t* g = static_cast<sgp_graph_t*>(h);
    sgp_free_graph(g);
    delete g;
}

int ref_load_config(const char* path, int* ints5, double* tol)
{
    config_t c;
    std::memset(&c, 0, sizeof(c));
    int rc = sgp_load_config(path, &c);
    ints5[0] = c.coarsening_alg;
    ints5[1] = c.refine_alg;
    ints5[2] = c.local_search_alg;
    ints5[3] = c.num_partitions;
    ints5[4] = c.num_iter;
    *tol = c.tol;
    return rc;
}

// Splits g by part with the reference routine; returns two opaque graphs.
int ref_use_partition(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                      uint32_t* part, void** g1, void** g2)
{
    sgp_graph_t g;
    std::memset(&g, 0, sizeof(g));
    g.nvertices = n;
    g.nedges = rowptr[n] / 2;
    g.source_offsets = const_cast<uint32_t*>(rowptr);
    g.destination_indices = const_cast<uint32_t*>(col);
    sgp_graph_t* a = new sgp_graph_t;
    sgp_graph_t* b = new sgp_graph_t;
    std::memset(a, 0, sizeof(*a));
    std::memset(b, 0, sizeof(*b));
    int rc = sgp_use_partition(part, g, a, b);
    *g1 = a;
    *g2 = b;
    return rc;
}

int ref_edit_distance(const uint32_t* p1, const uint32_t* p2, uint32_t n, uint32_t* diff)
{
    return compute_partition_edit_distance(p1, p2, n, diff);
}

}  // extern "C"
