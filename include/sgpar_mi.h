/*
 * sgpar_mi.h -- C ABI of the MI355X (gfx950) multilevel spectral bisection
 * library libsgpar_mi.so.
 *
 * The library replaces the hot path of the reference's sgpar partitioner:
 * heavy-edge coarsening, the coarse-graph build P^T A P, the multilevel
 * Fiedler power iteration and the sort-and-split bisection.  Plain pointers
 * and sizes only; every call returns an int status (0 = success, see
 * MI_STATUS_*) and never throws.  mi_last_error() describes the last failure
 * of a context.
 *
 * Types follow the reference's default build (definitions_kokkos.h:61-75):
 * vertex ids, edge offsets and weights are uint32_t, vectors are double.
 *
 * Ownership: host arrays are caller-owned and only read/written during the
 * call.  Device arrays passed with the *_device entry points are
 * caller-owned device memory on the context's device.  Objects (mi_ctx,
 * mi_graph, mi_hier) are owned by the caller and released with the matching
 * *_destroy call.  One context per host thread; all work of a context runs
 * in order on its own HIP stream.
 */
#ifndef SGPAR_MI_H_
#define SGPAR_MI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_STATUS_OK 0
#define MI_STATUS_ARG 1      /* invalid argument or malformed CSR */
#define MI_STATUS_HIP 2      /* HIP runtime failure (no GPU, launch error) */
#define MI_STATUS_ALLOC 3    /* device memory exhausted */
#define MI_STATUS_STATE 4    /* object used in the wrong state */
#define MI_STATUS_INTERNAL 5 /* internal consistency check failed */

#define MI_MAX_LEVELS 202 /* the reference stops after 200 graphs, coarsen_kokkos.h:1218 */

typedef struct mi_ctx mi_ctx;
typedef struct mi_graph mi_graph;
typedef struct mi_hier mi_hier;

/* Mirrors config_t, sgpar.h:186-193 (same fields, same order). */
typedef struct {
    int coarsening_alg;   /* ignored on this path, as in the reference's Kokkos branch */
    int refine_alg;       /* 0: Laplacian, 1: normalised Laplacian (normLap) */
    int local_search_alg; /* 0 only (the reference returns before local search); FM refinement is
                             the context's mi_ctx_set_refinement, as it is a build flag there */
    int num_iter;         /* trials; the library runs one trial per call */
    double tol;           /* power-iteration tolerance, default 1e-10 */
    int num_partitions;   /* 2 or 4 (4 splits in two, like sgp_compute_partition) */
} mi_config;

/* Per-call measurements, levels indexed 0 = finest. */
typedef struct {
    int nlevels;
    uint32_t level_n[MI_MAX_LEVELS];
    uint32_t level_m[MI_MAX_LEVELS];          /* stored nonzeros */
    uint64_t level_iters[MI_MAX_LEVELS];      /* power iterations (FM: fm_refine calls) */
    int level_max_reached[MI_MAX_LEVELS];     /* 1e14/n cap hit */
    double level_eigen_ms[MI_MAX_LEVELS];
    double coarsen_ms, eigen_ms, partition_ms, total_ms;
    double permute_ms, heavy_ms, map_ms, build_ms;
    uint32_t edge_cut;
    /* Final eigenvalue log of the finest level, refine_alg 0 only
       (sgp_power_iter_eigenvalue_log, eigensolve_kokkos.h:60-100):
       eigenvalue = sum of t_i^2 with t = len .* u - A u (len = row lengths),
       eigen_min / eigen_max = min / max of t_i / u_i.  The reference prints
       these with "edge cut lb" = eigenvalue * n / 4 and
       "gap ratio" = ceil(1 / (1 - eigenvalue)).  has_eigen_log = 0 when not
       computed (refine_alg 1, or a call without a final level). */
    int has_eigen_log;
    double eigenvalue, eigen_min, eigen_max;
    int mixed;  /* levels whose power iteration ran with FP32 vectors */
} mi_stats;

/* ---------------------------------------------------------------- context */
/* Creates a context on HIP device `device` (its stream, pool, workspace). */
int mi_ctx_create(int device, mi_ctx** out);
/* Ends the caller's use of ctx.  Graphs and hierarchies made on it stay
   valid for mi_graph_destroy / mi_hier_destroy, in any order; the context's
   memory goes when the last of them does. */
void mi_ctx_destroy(mi_ctx* ctx);
/* Text of the last failure on this context ("" if none). */
const char* mi_last_error(const mi_ctx* ctx);
/* Returns cached pool blocks to the runtime. */
int mi_ctx_trim(mi_ctx* ctx);
/* enable != 0: power iterations are replayed from a HIP graph captured once
   per level; 0 (the default, measured 1.5 % faster here): every kernel is
   launched from the host, a chunk ahead of the device.  Results are
   identical either way.  No counterpart in the reference. */
int mi_ctx_set_graphs(mi_ctx* ctx, int enable);
/* mixed != 0: mixed-precision eigensolve.  On every level above the finest
   with at least min_entries stored entries (0 selects 2^20: smaller levels
   are bound by launch and reduction latency, not by bytes, and stay FP64)
   the power iteration keeps
   its vectors, row partials and products in FP32 while the weighted degrees
   (integers), the shift, the five sums behind the deflation, the norm and
   the dot product (the dot normalised by the norm of the rounded x), and the
   stopping test stay FP64; it stops at the configured tolerance or after
   20000 iterations.  The finest level is always solved in FP64: the final
   refinement.  mi_power_iter (one level) runs in FP32 under this setting
   whatever the level's size.  mixed == 0 (the default): FP64 throughout, as
   the reference.  No counterpart in the reference. */
int mi_ctx_set_precision(mi_ctx* ctx, int mixed, uint32_t min_entries);
/* Levels of at most 16384 vertices and at most max_entries stored entries
   (default 131072; 0: none) run their whole FP64 power iteration in one
   1024-thread workgroup, chunk iterations per launch at most (0: 8192),
   instead of two kernel launches per iteration.  Same stop rule and
   result bookkeeping; the sums are added in another order.  No counterpart
   in the reference. */
int mi_ctx_set_small_levels(mi_ctx* ctx, uint32_t max_entries, uint32_t chunk);
/* enable != 0 (the default): a level above the finest with at least
   min_vertices vertices (0: 65536) whose power step gathers x from global
   memory (too many vertices for the LDS column blocks) is solved with its
   vertices in the order of their smallest finest-level vertex, which keeps
   the gathers of x local on meshes and geometric graphs (HEC numbers coarse
   vertices without locality); the level's vector goes back to the HEC
   numbering afterwards, so every output keeps it.  It applies only when
   the finest level's own order has locality (sampled mean |row - column|
   below 2 % of n; a random order gives about 33 %).  Sums run in another
   order, so results agree with enable = 0 to rounding.  No counterpart in
   the reference. */
int mi_ctx_set_locality(mi_ctx* ctx, int enable, uint32_t min_vertices);

/* Coarse builds (P^T A P with duplicate merge) of levels with at least
   min_entries stored entries take the bucketed form: rows grouped by coarse
   vertex, merged by sorts and dense accumulators in LDS; smaller levels take
   one radix sort over all entries.  Default 0xFFFFFFFF (the radix form on
   every level); 0 puts every level on the bucketed form.  Results are
   identical either way.  No counterpart in the reference. */
int mi_ctx_set_bucketed_build(mi_ctx* ctx, uint32_t min_entries);
/*
 * Refinement of the bisections of this context (mi_partition_graph,
 * mi_bisect_batch, mi_partition_4way): fm = 0 (default) the multilevel
 * Fiedler power iteration of the reference's spectral binaries; fm = 1 FM
 * refinement as the reference's *_fm binaries (hec_fm, the documented
 * default, build_and_usage_instructions.txt:80; compiled with -DFM,
 * cmg/CMakeLists.txt:50): the FM branch of sgp_eigensolve
 * (eigensolve_kokkos.h:1506-1587, see mi_fm_eigensolve) whose 0/1
 * partition goes to the sort-and-split as the eigenvector.  With fm = 1
 * the pcg state is not consumed, the coarsest level's init_spectral (more
 * than 200 vertices) is seeded with hec_seed, and mi_stats.level_iters
 * counts fm_refine calls.
 */
int mi_ctx_set_refinement(mi_ctx* ctx, int fm);
/*
 * Coarsening map of this context's hierarchies (mi_hierarchy_build, every
 * bisection, mi_hec_level): variant 0 (default) the HEC map of the hec_*
 * binaries (heuristics.h:633-693); variant 1 the heavy-edge matching of the
 * match_* binaries (-DPUREMATCH, cmg/CMakeLists.txt:45,53; sgp_coarsen_match,
 * heuristics.h:1584-1735, 1890-1895): HEC's permutation and level-1 random
 * neighbours, the first heaviest neighbour above level 1, rounds in
 * permutation order with the ordering condition alternating in which a
 * vertex pairs with its heavy neighbour when both are unmatched at its turn
 * (ids in order of pairing), unmatched vertices moving on with their last
 * unmatched neighbour of the largest position (level 1) or weight, the rest
 * singletons in index order -- a level with a vertex without entries, or
 * with only self-loops left, has no map in the reference and gives
 * MI_STATUS_ARG; variant 2 the mtmetis_* binaries' (-DMTMETIS,
 * cmg/CMakeLists.txt:46,54): that matching, then, each while more than a
 * quarter of the vertices is unmapped, the leaf, twin and relative passes
 * of heuristics.h:1737-1888 (leaves of a mapped vertex paired in row order;
 * vertices with equal sums of their neighbours' MurmurHash3 digests paired
 * along Kokkos::BinSort's order; unmapped neighbours of mapped vertices
 * naming each other, with one matching round per naming); variant 4 the map
 * of the hec2_* binaries (-DHEC_V2, cmg/CMakeLists.txt:41,51):
 * parallel_map_construct_v2 (heuristics.h:695-764) on HEC's heavy
 * neighbours as one thread runs it -- rounds in which the first list vertex
 * naming each heavy-neighbour slot claims it, slots take pseudo-locks in
 * order and pair their vertex with its heavy neighbour (a later pair may
 * re-map a vertex, which can leave ids without members, as there), and
 * unmapped vertices join their heavy neighbour's id; a round that maps
 * nothing gives MI_STATUS_ARG; variant 3 the map of the hec3_* binaries
 * (compiled with -DHEC_V3, cmg/CMakeLists.txt:42,52): parallel_map_construct_v3
 * (heuristics.h:766-826) on the same heavy neighbours -- mutual heavy pairs
 * keyed by their smaller end, heavy targets keyed by themselves, every
 * other vertex joining its heavy target's key, keys numbered in increasing
 * order; variant 5 the MIS-2 coarsening of the mis_* binaries (-DMIS,
 * cmg/CMakeLists.txt:47,55): mis_2 (heuristics.h:8-100), a maximal
 * distance-2 independent set from the XorShift64 pool of the level's seed,
 * one aggregate per set vertex in index order with its neighbours, the
 * rest joined by sgp_coarsen_mis_2's index-order sweep (:218-289); variant
 * 6 the GOSH coarsening of the gosh_* binaries (-DGOSH,
 * cmg/CMakeLists.txt:48,56): GOSH_clusters (heuristics.h:102-217), centres
 * by rounds of (state, entries, index) neighbour maxima that skip edges
 * whose ends both have more than nnz/n entries, one aggregate per centre in
 * index order, every other vertex joining its centre neighbour with the
 * most entries, the first of equals in its row (sgp_coarsen_GOSH,
 * :312-377); variant 7 the GOSH v2 coarsening of the gosh2_* binaries
 * (-DGOSH_V2, cmg/CMakeLists.txt:49,57; sgp_coarsen_GOSH_v2,
 * heuristics.h:444-601): seeds of more than max(nnz/n, 50) entries numbered
 * in index order, three index-order sweeps that see their own writes
 * (unmapped vertices join the mapped neighbour of the largest weight, then
 * of the most entries; mapped vertices hand their id to unmapped
 * neighbours), then the remaining vertices choose a heavy neighbour by
 * (weight, XorShift64 draw of the level's seed, one per entry) and go
 * through parallel_map_construct_prefilled (:378-442) -- a remaining
 * vertex without entries gives MI_STATUS_ARG; variant 8 the serial HEC of
 * hec_spec_serial (-DHEC_SERIAL, cmg/CMakeLists.txt:16,40;
 * sgp_coarsen_HEC_serial, heuristics.h:828-928): the list is a
 * Fisher-Yates shuffle by the XorShift64 pool of the level's seed (step i =
 * n .. 2 swaps positions draw % i and i - 1), the heavy neighbour a random
 * entry at level 1 (the next n draws) and the first heaviest entry above
 * it, and the map the greedy walk in list order (an unmapped vertex makes a
 * new id for its heavy neighbour when that one is unmapped, else joins it)
 * -- a vertex without entries gives MI_STATUS_ARG.  With variants 5 to 7
 * mi_hec_level leaves vperm and hn untouched.  Other values:
 * MI_STATUS_ARG.
 */
int mi_ctx_set_coarsening(mi_ctx* ctx, int variant);

/*
 * Coarse-graph build of this context's hierarchies (every level build and
 * mi_build_coarse): builder 0 (default) the RADIX build of the reference's
 * binaries (coarsen_kokkos.h:750-762), rows sorted by column; builder 1 the
 * HASHMAP build of hec_spec_hashmap (-DHASHMAP, cmg/CMakeLists.txt:44;
 * coarsen_kokkos.h:658-749 with the accumulator of :520-645) as it runs on
 * one thread: each coarse row lists its columns in order of first
 * occurrence among the fine entries (fine vertices ascending, each row in
 * CSR order), weights summed.  The same graph as a set; the order decides
 * the heavy-neighbour ties of the next level and the summation order of the
 * products.  Other values: MI_STATUS_ARG.
 */
int mi_ctx_set_builder(mi_ctx* ctx, int builder);

/* ---------------------------------------------------------------- graphs */
/*
 * Uploads a CSR graph (host arrays) to HBM.  Replaces the host->device copy
 * of sgp_generate_coarse_graphs, coarsen_kokkos.h:1169-1195.  w == NULL means
 * unit weights, which is what the reference's sgp_partition_graph always uses
 * (coarsen_kokkos.h:1182); a non-NULL w gives the weighted form.  The CSR must
 * be symmetric with no self-loops (the reference input format, mtx2csr.cpp).
 */
int mi_graph_create(mi_ctx* ctx, uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                    const uint32_t* w, mi_graph** out);
/* Builds a graph from an edge list (host arrays u, v, and w or NULL, of
   nedges edges, every vertex id below n) the way the reference's converter
   mtx2csr.cpp:116-289 turns a MatrixMarket file into its binary CSR, on the
   device: every edge stored in both directions, self-loops dropped,
   duplicates merged (their weights summed; w == NULL gives unit weights,
   the reference's pattern graphs), rows sorted by column, and with lcc != 0
   only the largest connected component, its vertices renumbered in their
   order (a tie between components of equal size goes to the one with the
   smallest vertex, which the reference's BFS in vertex order finds first).
   mi_graph_size gives the resulting n and m. */
int mi_graph_from_edges(mi_ctx* ctx, uint32_t n, uint64_t nedges, const uint32_t* u, const uint32_t* v,
                        const uint32_t* w, int lcc, mi_graph** out);
/* Wraps device-resident CSR arrays (not copied; must outlive the graph). */
int mi_graph_create_device(mi_ctx* ctx, uint32_t n, uint32_t m, const uint32_t* d_rowptr,
                           const uint32_t* d_col, const uint32_t* d_w, mi_graph** out);
void mi_graph_destroy(mi_graph* g);
int mi_graph_size(const mi_graph* g, uint32_t* n, uint32_t* m);
/* Copies a graph to host arrays (rowptr n+1, col m, w m; any may be NULL;
   w receives 1s for a unit-weight graph). */
int mi_graph_copy(mi_ctx* ctx, const mi_graph* g, uint32_t* rowptr, uint32_t* col, uint32_t* w);

/* ---------------------------------------------------------------- coarsening */
/*
 * Builds the multilevel hierarchy: replaces sgpar_kokkos::sgp_generate_coarse_graphs,
 * coarsen_kokkos.h:1165-1234 (HEC map heuristics.h:930-1031 and coarse build
 * coarsen_kokkos.h:1061-1127 per level; coarsen while n > 50, at most 200
 * graphs, drop a last level under 10 vertices).  `seed` plays the role of
 * std::time(nullptr), which the reference uses to seed the XorShift pool at
 * every level (heuristics.h:947): with one seed for all levels the result
 * equals the reference's Serial-backend run with time() fixed to `seed`.
 */
int mi_hierarchy_build(mi_ctx* ctx, const mi_graph* g, uint64_t seed, mi_hier** out,
                       mi_stats* stats);
/* Wraps a hierarchy given as host CSR levels (level 0 finest; vcmap[l] maps
   level l to l+1, vcmap[nlevels-1] unused).  w[l] == NULL means unit. */
int mi_hier_create_from_host(mi_ctx* ctx, int nlevels, const uint32_t* n,
                             const uint32_t* const* rowptr, const uint32_t* const* col,
                             const uint32_t* const* w, const uint32_t* const* vcmap,
                             mi_hier** out);
void mi_hier_destroy(mi_hier* h);
int mi_hier_nlevels(const mi_hier* h);
int mi_hier_level_size(const mi_hier* h, int level, uint32_t* n, uint32_t* m);
/* Copies level `level` to host; any pointer may be NULL.  w receives 1s for a
   unit-weight level; vcmap (n entries) is only defined below the coarsest. */
int mi_hier_level_copy(mi_ctx* ctx, const mi_hier* h, int level, uint32_t* rowptr,
                       uint32_t* col, uint32_t* w, uint32_t* vcmap, uint32_t* vtx_w);

/* ---------------------------------------------------------------- eigensolve */
/*
 * Multilevel Fiedler vector: replaces sgpar_kokkos::sgp_eigensolve,
 * eigensolve_kokkos.h:1493-1595 (non-FM branch).  The coarsest guess is
 * pcg32()/UINT32_MAX drawn from (*pcg_state, pcg_inc), which is advanced like
 * the reference's rng (:1500-1505); each level runs sgp_power_iter
 * (:102-206) then prolongs (:1556-1558).  normLap = config.refine_alg.
 * evec (host, n0 doubles) may be NULL.  stats may be NULL.
 */
int mi_fiedler(mi_ctx* ctx, const mi_hier* h, uint64_t* pcg_state, uint64_t pcg_inc,
               int normLap, double tol, double* evec, mi_stats* stats);
/*
 * The FM branch of sgp_eigensolve, eigensolve_kokkos.h:1506-1587 (without
 * RECOARSEN): the coarsest level's partition from init_gggp (at most 200
 * vertices, :1242-1320) or init_spectral (:1322-1368, its random start from
 * the XorShift64 pool of `seed` where the reference uses std::time); at
 * each level, coarsest first, fm_refine repeated until it returns the cut
 * it returned before (the last cut carried over from the coarser level);
 * the partition prolonged between levels.  evec (host, n0 doubles, may be
 * NULL) receives the finest partition as 0.0 / 1.0, like the reference's
 * eigenvec; *last_cut (may be NULL) the last fm_refine result (twice the
 * cut weight).  stats.level_iters = fm_refine calls per level.
 */
int mi_fm_eigensolve(mi_ctx* ctx, const mi_hier* h, uint64_t seed, double tol, double* evec,
                     uint32_t* last_cut, mi_stats* stats);
/*
 * One fm_refine call (eigensolve_kokkos.h:973-1240, with the gains and
 * bucket lists of fm_create_ds, :208-369) on graph g with vertex weights
 * vtx_w (host, n; NULL = all 1): part (host, n, in/out; 0 = side 0,
 * anything else side 1) receives the refined sides, *min_cut the value
 * fm_refine returns (twice the cut weight).  info (host, 4, may be NULL):
 * moves made, index of the last kept move (-1 = none kept), |balance| of
 * the result, the largest weighted degree.  The largest weighted degree
 * must be below 2^30.
 */
int mi_fm_refine(mi_ctx* ctx, const mi_graph* g, const uint32_t* vtx_w, uint32_t* part,
                 uint32_t* min_cut, int64_t* info);
/*
 * The FM path's initial partition of a coarsest graph, chosen by size as
 * eigensolve_kokkos.h:1510-1514 does: init_gggp (:1242-1320) for at most
 * 200 vertices, else init_spectral (:1322-1368) with the XorShift64 pool of
 * `seed` and power-iteration tolerance tol.  part (host, n) receives 0/1.
 */
int mi_fm_init(mi_ctx* ctx, const mi_graph* g, const uint32_t* vtx_w, uint64_t seed, double tol,
               uint32_t* part);
/* One reference power iteration (eigensolve_kokkos.h:102-206) on graph g,
   in place on host vector u (n); *iters receives the iteration count. */
int mi_power_iter(mi_ctx* ctx, const mi_graph* g, double* u, int normLap, double tol,
                  uint64_t* iters, int* max_reached);

/* ---------------------------------------------------------------- partition */
/*
 * Sort-and-split bisection of g by the host eigenvector evec: replaces
 * sgp_compute_partition, sgpar.h:1969-2087 with local_search_alg = 0.  part
 * (host, n) receives 0/1, *edge_cut the number of cut edges.
 */
int mi_compute_partition(mi_ctx* ctx, const mi_graph* g, const double* evec,
                         int num_partitions, int perc_imbalance, uint32_t* part,
                         uint32_t* edge_cut);

/*
 * mi_compute_partition without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under stream capture, serving loops, harnesses that forbid hipMalloc).
 * d_evec (n doubles) and d_part (n words) are device arrays on the
 * context's device; the result is mi_compute_partition's, bit for bit.
 * Every temporary lives in one fixed device array of the library (about
 * 49 MiB per device, held until the process ends), so g may have at most
 * MI_NOALLOC_MAX_VERTICES vertices (MI_STATUS_ARG beyond).  Calls that
 * overlap on one device (other threads or contexts) race on that array.
 */
#define MI_NOALLOC_MAX_VERTICES 2097152u
int mi_compute_partition_noalloc(mi_ctx* ctx, const mi_graph* g, const double* d_evec,
                                 int num_partitions, int perc_imbalance, uint32_t* d_part,
                                 uint32_t* edge_cut);

/*
 * The whole bisection: replaces sgp_partition_graph, sgpar.h:2363-2431
 * (Kokkos branch): coarsen, eigensolve, split.  part is a host array of n, or
 * a device array when part_on_device != 0.  hec_seed as in
 * mi_hierarchy_build; (*pcg_state, pcg_inc) as in mi_fiedler.
 */
int mi_partition_graph(mi_ctx* ctx, const mi_graph* g, const mi_config* config,
                       int perc_imbalance, uint64_t hec_seed, uint64_t* pcg_state,
                       uint64_t pcg_inc, uint32_t* part, int part_on_device,
                       uint32_t* edge_cut, mi_stats* stats);

/*
 * Bisections of count independent graphs: the same results, graph by graph
 * and bit for bit, as mi_partition_graph(ctx, graphs[k], config,
 * perc_imbalance, hec_seeds[k], &pcg_states[k], pcg_inc, parts[k], ...) for
 * k = 0 .. count-1.  The reference has no batched form (it bisects one
 * graph per call).  The hierarchies are built graph by graph; then the
 * power iterations of the small levels (those that run in one workgroup
 * with every vector in LDS),
 * from each graph's coarsest level down, run for all graphs together, one
 * workgroup per graph and level, level by level; the larger levels and the
 * splits follow graph by graph.  parts[k] are host arrays of graphs[k]'s n
 * (device arrays when parts_on_device != 0).  stats (count entries, may be
 * NULL) receive the sizes, iteration counts, cuts and eigenvalue logs but
 * no times; phase_ms (3 doubles, may be NULL) receives the times of the
 * coarsening, the eigensolves and the splits of the whole batch.
 */
int mi_bisect_batch(mi_ctx* ctx, int count, const mi_graph* const* graphs,
                    const mi_config* config, int perc_imbalance, const uint64_t* hec_seeds,
                    uint64_t* pcg_states, uint64_t pcg_inc, uint32_t* const* parts,
                    int parts_on_device, uint32_t* edge_cuts, mi_stats* stats,
                    double* phase_ms);

/*
 * Induced subgraphs of the two sides of part (n entries, 0 = first side,
 * anything else = second; host or device memory): sgp_use_partition,
 * sgpar.h:2666-2752.  Vertices are relabelled in order within their side,
 * edges between the sides are dropped, the rest keep their CSR order, and
 * every edge weight is 1 whatever g's weights are.  *g1 and *g2 are new
 * graphs on ctx (mi_graph_destroy them); a side may have no vertices.
 */
int mi_use_partition(mi_ctx* ctx, const mi_graph* g, const uint32_t* part, int part_on_device,
                     mi_graph** g1, mi_graph** g2);

/*
 * Recursive 4-way partition, the driver of cmg/sgpar.cpp:53-89 run on the
 * device: bisect g (as mi_partition_graph), take the induced subgraphs of
 * the two sides (sgp_use_partition, sgpar.h:2666-2752: each side's vertices
 * relabelled in order, edges between the sides dropped, every edge weight 1
 * whatever g's weights are), then bisect each
 * side with the same config, hec_seed and the continuing pcg state.  part
 * (n entries, host or device) receives 2*side + sub-side in {0,1,2,3};
 * cuts[0..2] the three bisection cuts; *edge_cut_4way their sum (what the
 * reference logs as the 4-way cut, ExperimentLoggerUtil::setEdgeCut4Way).
 * A side with fewer than 2 vertices is not split (cut 0).
 */
int mi_partition_4way(mi_ctx* ctx, const mi_graph* g, const mi_config* config,
                      int perc_imbalance, uint64_t hec_seed, uint64_t* pcg_state,
                      uint64_t pcg_inc, uint32_t* part, int part_on_device, uint32_t* cuts,
                      uint32_t* edge_cut_4way);

/* ---------------------------------------------------------------- building blocks
 * Single steps of the path, exposed for parity tests and measurement. */

/* urand64 draws of the XorShift64 pool of `seed` at stream positions
   [first, first+count) (Kokkos_Random.hpp:702-720, 863-898). */
int mi_xs_draws(mi_ctx* ctx, uint64_t seed, uint64_t first, uint64_t count, uint64_t* out);
/* One HEC level (heuristics.h:930-1031) on g; host outputs, any may be NULL
   except vcmap. */
int mi_hec_level(mi_ctx* ctx, const mi_graph* g, int level, uint64_t seed, uint32_t* vcmap,
                 uint32_t* vperm, uint32_t* hn, uint32_t* nc);
/* Coarse graph of g under vcmap (coarsen_kokkos.h:1061-1127); returns a new
   graph (weighted) and the coarse vertex weights (host, nc; may be NULL). */
int mi_build_coarse(mi_ctx* ctx, const mi_graph* g, const uint32_t* vcmap, uint32_t nc,
                    mi_graph** out, uint32_t* c_vtx_w);
/* y = A x + (gb - wdeg) .* x with gb = max(2 wdeg, 2) (eigensolve_kokkos.h:161-179);
   host x, y (n). */
int mi_spmv_shift(mi_ctx* ctx, const mi_graph* g, const double* x, double* y);
/* Measurement of one power-iteration step (the FP64 shifted product with its
   reductions, eigensolve_kokkos.h:161-189) on g, timed with HIP events on the
   context stream over `reps` steps, and of its streaming kernel alone over
   `reps` launches.  Algorithmic bytes count col (+ w), the vectors x and y,
   wdeg (and rowptr for the step); impl_bytes adds the per-entry row ids and
   the row sums the streaming kernel moves.  column_blocks: 0 when the level
   gathers x from global memory, else the number of LDS column blocks.  No
   counterpart in the reference. */
typedef struct {
    double step_ms, stream_ms;
    double step_bytes, stream_bytes, impl_bytes;
    uint32_t column_blocks;
} mi_step_bench;
int mi_bench_step(mi_ctx* ctx, const mi_graph* g, int reps, mi_step_bench* out);
/* Stable LSD radix sort of (key, value) pairs on bits [begin_bit, end_bit), host arrays. */
int mi_radix_sort_pairs(mi_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, int begin_bit,
                        int end_bit);
/* The permutation sort of generate_permutation (heuristics.h:614-631: a
   BinSort of 64-bit random draws): stable order of all 64 key bits, taken as
   four radix passes on the high halves plus the ordering of runs of equal
   high halves.  Exact for any keys; fast when the keys are random. */
int mi_sort_random_keys(mi_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n);
/* Exclusive prefix sum of u32, host arrays; *total receives the sum. */
int mi_scan_u32(mi_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total);

#ifdef __cplusplus
}
#endif

#endif /* SGPAR_MI_H_ */
