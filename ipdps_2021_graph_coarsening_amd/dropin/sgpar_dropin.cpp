// sgpar_dropin.cpp -- the reference's C++ API (namespace sgpar, the
// declarations of cmg/sgpar.h:2230-2244 with EXPERIMENT defined) implemented
// over libsgpar_mi.so, so that a caller written against the reference's
// header links this library in place of the Kokkos build.
//
// Compiled against the reference's own cmg/sgpar.h and
// cmg/ExperimentLoggerUtil.cpp where they lie (-I <reference>/cmg,
// -DEXPERIMENT, without _KOKKOS and without SGPAR_IMPLEMENTATION): the header
// then declares the API with the reference's types (sgp_graph_t, config_t,
// sgp_pcg32_random_t, ExperimentLoggerUtil) and defines none of the
// functions below, which this file defines.
//
//   sgp_partition_graph      sgpar.h:2363-2431 (Kokkos branch): coarsen,
//                            eigensolve, split -> mi_partition_graph on the
//                            GPU, metrics into the caller's
//                            ExperimentLoggerUtil as the Kokkos path records
//                            them (eigensolve_kokkos.h:198, coarsen_kokkos.h
//                            :1151,1159, heuristics.h:956,999,1008,
//                            sgpar.h:2398-2402)
//   sgp_load_graph           sgpar.h:2262-2289
//   sgp_free_graph           sgpar.h:2291-2319
//   sgp_load_partition       sgpar.h:2321-2340
//   sgp_use_partition        sgpar.h:2666-2752 -> mi_use_partition
//   sgp_load_config          sgpar.h:2754-2787
//   compute_partition_edit_distance  sgpar.h:2343-2361
//
// Differences from the reference, all stated where they occur:
// - the power-iteration tolerance is config->tol (sgp_load_config sets it
//   there and in the reference's global SGPAR_POWERITER_TOL alike; a global
//   in the caller's translation unit is not visible to a library);
// - the HEC pool seed and the reference's std::time(nullptr) (heuristics.h
//   :947) come from one clock read per call, which sgpar_mi_fix_clock() can
//   pin (the reference pins time() itself for reproducible runs);
// - the graph upload and every device buffer belong to one library context
//   on HIP device 0 (SGPAR_MI_DEVICE selects another), made on first use.
//
// Compiled with -DHEC_V3 (the hec3_* binaries) the context coarsens with the
// v3 map, with -DHEC_V2 (the hec2_* binaries) with the v2 map, with -DPUREMATCH (the match_* binaries) with heavy-edge matching,
// with -DMTMETIS (the mtmetis_* binaries) with that matching and its passes,
// with -DMIS (the mis_* binaries) with MIS-2, with -DGOSH (the
// gosh_* binaries) with GOSH centres, with -DGOSH_V2 (the gosh2_* binaries)
// with GOSH v2, with -DHEC_SERIAL (hec_spec_serial) with the serial HEC
// (mi_ctx_set_coarsening); with -DHASHMAP
// (hec_spec_hashmap) its coarse rows keep the HASHMAP build's order
// (mi_ctx_set_builder).
// Compiled with -DFM (as the reference's *_fm binaries are,
// cmg/CMakeLists.txt:50) the context refines by FM (mi_ctx_set_refinement)
// and the metrics record what the reference's FM path records: no coarse
// level per power iteration (that path runs none) and the FM time under
// fm-refine (eigensolve_kokkos.h:1236; here the whole FM eigensolve,
// initial partition and prolongations included).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <vector>

#include "sgpar.h"
#include "sgpar_mi.h"

namespace {

std::mutex g_mu;
mi_ctx* g_ctx = nullptr;  // never destroyed: the HIP runtime tears itself down at exit
int64_t g_fixed_clock = -1;

mi_ctx* context()
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ctx) {
        const char* dev = std::getenv("SGPAR_MI_DEVICE");
        if (mi_ctx_create(dev ? std::atoi(dev) : 0, &g_ctx) != 0) g_ctx = nullptr;
#ifdef FM
        if (g_ctx && mi_ctx_set_refinement(g_ctx, 1) != 0) {
            mi_ctx_destroy(g_ctx);
            g_ctx = nullptr;
        }
#endif
#if defined HEC_V3 || defined HEC_V2 || defined MIS || defined GOSH || defined GOSH_V2 || defined PUREMATCH || \
    defined MTMETIS || defined HEC_SERIAL
#if defined HEC_SERIAL
        const int variant = 8;
#elif defined HEC_V2
        const int variant = 4;
#elif defined GOSH_V2
        const int variant = 7;
#elif defined MIS
        const int variant = 5;
#elif defined PUREMATCH
        const int variant = 1;
#elif defined MTMETIS
        const int variant = 2;
#elif defined GOSH
        const int variant = 6;
#else
        const int variant = 3;
#endif
        if (g_ctx && mi_ctx_set_coarsening(g_ctx, variant) != 0) {
            mi_ctx_destroy(g_ctx);
            g_ctx = nullptr;
        }
#endif
#ifdef HASHMAP
        if (g_ctx && mi_ctx_set_builder(g_ctx, 1) != 0) {
            mi_ctx_destroy(g_ctx);
            g_ctx = nullptr;
        }
#endif
    }
    return g_ctx;
}

uint64_t clock_seed()
{
    std::lock_guard<std::mutex> lock(g_mu);
    return g_fixed_clock >= 0 ? (uint64_t)g_fixed_clock : (uint64_t)std::time(nullptr);
}

int fail(mi_ctx* ctx, const char* what, int rc)
{
    std::printf("sgpar Error: %s returned %d (%s)\n", what, rc, ctx ? mi_last_error(ctx) : "no HIP context");
    return EXIT_FAILURE;
}

}  // namespace

// Pins the value that stands for std::time(nullptr) (t >= 0), or returns to
// the real clock (t < 0).  No counterpart in the reference, whose runs pin
// time() from outside the process.
extern "C" void sgpar_mi_fix_clock(int64_t t)
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_fixed_clock = t;
}

namespace sgpar {

int sgp_load_graph(sgp_graph_t* g, char* csr_filename)
{
    FILE* f = std::fopen(csr_filename, "rb");
    if (!f) {
        std::printf("Error: Could not open input file. Exiting ...\n");
        return EXIT_FAILURE;
    }
    int64_t head[6];
    if (std::fread(head, sizeof(int64_t), 6, f) != 6 || head[0] < 0 || head[1] < 0 || head[0] > UINT32_MAX ||
        head[1] > UINT32_MAX) {
        std::fclose(f);
        return EXIT_FAILURE;
    }
    g->nvertices = (sgp_vid_t)head[0];
    g->nedges = (sgp_eid_t)(head[1] / 2);
    const size_t nrow = (size_t)g->nvertices + 1, nent = 2 * (size_t)g->nedges;
    g->source_offsets = (sgp_eid_t*)std::malloc(nrow * sizeof(sgp_eid_t));
    g->destination_indices = (sgp_vid_t*)std::malloc((nent ? nent : 1) * sizeof(sgp_vid_t));
    g->eweights = nullptr;
    g->weighted_degree = nullptr;
    g->edges_per_source = nullptr;
    const bool ok = g->source_offsets && g->destination_indices &&
                    std::fread(g->source_offsets, sizeof(sgp_eid_t), nrow, f) == nrow &&
                    std::fread(g->destination_indices, sizeof(sgp_vid_t), nent, f) == nent;
    std::fclose(f);
    if (!ok) {
        std::free(g->source_offsets);
        std::free(g->destination_indices);
        g->source_offsets = nullptr;
        g->destination_indices = nullptr;
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

int sgp_free_graph(sgp_graph_t* g)
{
    std::free(g->source_offsets);
    std::free(g->destination_indices);
    std::free(g->eweights);
    std::free(g->weighted_degree);
    std::free(g->edges_per_source);
    g->source_offsets = nullptr;
    g->destination_indices = nullptr;
    g->eweights = nullptr;
    g->weighted_degree = nullptr;
    g->edges_per_source = nullptr;
    return EXIT_SUCCESS;
}

int sgp_load_partition(sgp_vid_t* part, sgp_vid_t size, char* part_filename)
{
    FILE* f = std::fopen(part_filename, "r");
    if (!f) {
        std::printf("Error: Could not open partition file %s. Exiting ...\n", part_filename);
        return EXIT_FAILURE;
    }
    for (sgp_vid_t i = 0; i < size; i++) {
        int v = 0;
        if (std::fscanf(f, "%i", &v) == 0) {
            std::fclose(f);
            return EXIT_FAILURE;
        }
        part[i] = (sgp_vid_t)v;
    }
    return std::fclose(f) == 0 ? EXIT_SUCCESS : EXIT_FAILURE;
}

// Half the number of vertices whose labels differ, under the better of the
// two labellings of a bisection.
int compute_partition_edit_distance(const sgp_vid_t* part1, const sgp_vid_t* part2, sgp_vid_t size,
                                    sgp_vid_t* diff)
{
    sgp_vid_t differ = 0;
    for (sgp_vid_t i = 0; i < size; i++) differ += part1[i] != part2[i] ? 1u : 0u;
    const sgp_vid_t same = size - differ;
    *diff = (differ < same ? differ : same) / 2;
    return EXIT_SUCCESS;
}

// The six scalars in the file's order: coarsening_alg, refine_alg,
// local_search_alg, num_partitions, num_iter, tol.
int sgp_load_config(const char* config_f, config_t* c)
{
    FILE* f = std::fopen(config_f, "r");
    if (!f) {
        std::printf("Error: Could not open config file %s. Exiting ...\n", config_f);
        return EXIT_FAILURE;
    }
    const bool ok = std::fscanf(f, "%i", &c->coarsening_alg) != 0 && std::fscanf(f, "%i", &c->refine_alg) != 0 &&
                    std::fscanf(f, "%i", &c->local_search_alg) != 0 &&
                    std::fscanf(f, "%i", &c->num_partitions) != 0 && std::fscanf(f, "%i", &c->num_iter) != 0 &&
                    std::fscanf(f, "%lf", &c->tol) != 0;
    if (std::fclose(f) != 0 || !ok) return EXIT_FAILURE;
    if (c->tol != 1e-10) std::printf("Using non-default tolerance: %lf\n", c->tol);
    return EXIT_SUCCESS;
}

// Induced subgraphs of the two sides on the GPU (mi_use_partition), copied
// back into malloc'd arrays as the reference returns them: unit weights
// dropped, nedges = the stored entry count of the side.
int sgp_use_partition(sgp_vid_t* part, const sgp_graph_t g, sgp_graph_t* g1, sgp_graph_t* g2)
{
    mi_ctx* ctx = context();
    if (!ctx) return fail(ctx, "mi_ctx_create", -1);
    const uint32_t n = g.nvertices;
    mi_graph* dg = nullptr;
    int rc = mi_graph_create(ctx, n, g.source_offsets, g.destination_indices, nullptr, &dg);
    if (rc) return fail(ctx, "mi_graph_create", rc);
    mi_graph* sides[2] = {nullptr, nullptr};
    rc = mi_use_partition(ctx, dg, part, 0, &sides[0], &sides[1]);
    mi_graph_destroy(dg);
    if (rc) return fail(ctx, "mi_use_partition", rc);
    sgp_graph_t* out[2] = {g1, g2};
    for (int s = 0; s < 2; s++) {
        uint32_t sn = 0, sm = 0;
        mi_graph_size(sides[s], &sn, &sm);
        sgp_graph_t& h = *out[s];
        h.nvertices = sn;
        h.nedges = sm;
        h.source_offsets = (sgp_eid_t*)std::malloc(((size_t)sn + 1) * sizeof(sgp_eid_t));
        h.destination_indices = (sgp_vid_t*)std::malloc((sm ? sm : 1) * sizeof(sgp_vid_t));
        h.eweights = nullptr;
        h.edges_per_source = nullptr;
        h.weighted_degree = nullptr;
        if (!h.source_offsets || !h.destination_indices) rc = -1;
        if (!rc) rc = mi_graph_copy(ctx, sides[s], h.source_offsets, h.destination_indices, nullptr);
    }
    mi_graph_destroy(sides[0]);
    mi_graph_destroy(sides[1]);
    if (rc) return fail(ctx, "mi_graph_copy", rc);
    return EXIT_SUCCESS;
}

int sgp_partition_graph(sgp_vid_t* part, sgp_eid_t* edge_cut, config_t* config, const int perc_imbalance_allowed,
                        const sgp_graph_t g, ExperimentLoggerUtil& experiment, sgp_pcg32_random_t* rng)
{
    std::printf("sgpar settings: %d %lu %.16f\n", SGPAR_COARSENING_VTX_CUTOFF, (uint64_t)SGPAR_POWERITER_ITER,
                config->tol);
    mi_ctx* ctx = context();
    if (!ctx) return fail(ctx, "mi_ctx_create", -1);
    const double start = sgp_timer();
    const uint32_t n = g.nvertices;
    mi_graph* dg = nullptr;
    int rc = mi_graph_create(ctx, n, g.source_offsets, g.destination_indices, nullptr, &dg);
    if (rc) return fail(ctx, "mi_graph_create", rc);
    const double uploaded = sgp_timer();
    mi_config cfg;
    cfg.coarsening_alg = config->coarsening_alg;
    cfg.refine_alg = config->refine_alg;
    cfg.local_search_alg = config->local_search_alg;
    cfg.num_iter = config->num_iter;
    cfg.tol = config->tol;
    cfg.num_partitions = config->num_partitions;
    mi_stats st;
    uint64_t state = rng->state;
    uint32_t cut = 0;
    rc = mi_partition_graph(ctx, dg, &cfg, perc_imbalance_allowed, clock_seed(), &state, rng->inc, part, 0, &cut,
                            &st);
    mi_graph_destroy(dg);
    if (rc) return fail(ctx, "mi_partition_graph", rc);
    rng->state = state;
    *edge_cut = cut;

    // what the Kokkos path records: one coarse level per power iteration,
    // coarsest first (eigensolve_kokkos.h:198); map and build times
    // (coarsen_kokkos.h:1151,1159) with the map's permutation, heavy
    // neighbour and construction parts (heuristics.h:956,999,1008)
    using M = ExperimentLoggerUtil::Measurement;
#ifdef FM
    experiment.addMeasurement(M::FMRefine, st.eigen_ms * 1e-3);
#else
    for (int l = st.nlevels - 1; l >= 0; l--)
        experiment.addCoarseLevel((int)st.level_iters[l], st.level_max_reached[l] != 0, st.level_n[l]);
#endif
    experiment.addMeasurement(M::Map, (st.permute_ms + st.heavy_ms + st.map_ms) * 1e-3);
    experiment.addMeasurement(M::Permute, st.permute_ms * 1e-3);
    experiment.addMeasurement(M::Heavy, st.heavy_ms * 1e-3);
    experiment.addMeasurement(M::MapConstruct, st.map_ms * 1e-3);
    experiment.addMeasurement(M::Build, st.build_ms * 1e-3);
    const double upload_s = uploaded - start;
    const double coarsen_s = upload_s + st.coarsen_ms * 1e-3;  // the reference's coarsening includes the upload
    const double refine_s = st.eigen_ms * 1e-3;
    experiment.setTotalDurationSeconds(coarsen_s + refine_s);
    experiment.setCoarsenDurationSeconds(coarsen_s);
    experiment.setRefineDurationSeconds(refine_s);
    experiment.setFinestEdgeCut(*edge_cut);
    experiment.modifyCoarseLevelEC(0, *edge_cut);
    std::printf("Total: %3.3lf s, coarsening %3.3lf %3.0lf%% (sort %3.3lf %3.0lf%%), refine %3.3lf s\n",
                coarsen_s + refine_s, coarsen_s, 100.0 * coarsen_s / (coarsen_s + refine_s), st.build_ms * 1e-3,
                100.0 * st.build_ms * 1e-3 / (coarsen_s + refine_s), refine_s);
    return EXIT_SUCCESS;
}

}  // namespace sgpar
