/* Compiled and linked by tests/test_abi.py against include/sgpar_mi.h and
 * libsgpar_mi.so with a plain C compiler: every entry point is bound through
 * a function pointer of the declared type, so the header is usable from C
 * and the library resolves every declared symbol.  Run without a GPU, it
 * checks the error behaviour of the entry points that need no device: a
 * context cannot be created, and calls on a null context fail with
 * MI_STATUS_ARG instead of crashing. */
#include <stdio.h>
#include <string.h>

#include "sgpar_mi.h"

#define BIND(f) (void)(fp_##f = f)

int main(void)
{
    int (*fp_mi_ctx_create)(int, mi_ctx**);
    void (*fp_mi_ctx_destroy)(mi_ctx*);
    const char* (*fp_mi_last_error)(const mi_ctx*);
    int (*fp_mi_ctx_trim)(mi_ctx*);
    int (*fp_mi_ctx_set_graphs)(mi_ctx*, int);
    int (*fp_mi_ctx_set_precision)(mi_ctx*, int, uint32_t);
    int (*fp_mi_ctx_set_small_levels)(mi_ctx*, uint32_t, uint32_t);
    int (*fp_mi_ctx_set_locality)(mi_ctx*, int, uint32_t);
    int (*fp_mi_bisect_batch)(mi_ctx*, int, const mi_graph* const*, const mi_config*, int, const uint64_t*,
                              uint64_t*, uint64_t, uint32_t* const*, int, uint32_t*, mi_stats*, double*);
    int (*fp_mi_graph_create)(mi_ctx*, uint32_t, const uint32_t*, const uint32_t*, const uint32_t*, mi_graph**);
    int (*fp_mi_graph_from_edges)(mi_ctx*, uint32_t, uint64_t, const uint32_t*, const uint32_t*, const uint32_t*, int,
                                  mi_graph**);
    int (*fp_mi_graph_create_device)(mi_ctx*, uint32_t, uint32_t, const uint32_t*, const uint32_t*,
                                     const uint32_t*, mi_graph**);
    void (*fp_mi_graph_destroy)(mi_graph*);
    int (*fp_mi_graph_size)(const mi_graph*, uint32_t*, uint32_t*);
    int (*fp_mi_graph_copy)(mi_ctx*, const mi_graph*, uint32_t*, uint32_t*, uint32_t*);
    int (*fp_mi_hierarchy_build)(mi_ctx*, const mi_graph*, uint64_t, mi_hier**, mi_stats*);
    int (*fp_mi_hier_create_from_host)(mi_ctx*, int, const uint32_t*, const uint32_t* const*,
                                       const uint32_t* const*, const uint32_t* const*, const uint32_t* const*,
                                       mi_hier**);
    void (*fp_mi_hier_destroy)(mi_hier*);
    int (*fp_mi_hier_nlevels)(const mi_hier*);
    int (*fp_mi_hier_level_size)(const mi_hier*, int, uint32_t*, uint32_t*);
    int (*fp_mi_hier_level_copy)(mi_ctx*, const mi_hier*, int, uint32_t*, uint32_t*, uint32_t*, uint32_t*,
                                 uint32_t*);
    int (*fp_mi_fiedler)(mi_ctx*, const mi_hier*, uint64_t*, uint64_t, int, double, double*, mi_stats*);
    int (*fp_mi_power_iter)(mi_ctx*, const mi_graph*, double*, int, double, uint64_t*, int*);
    int (*fp_mi_compute_partition)(mi_ctx*, const mi_graph*, const double*, int, int, uint32_t*, uint32_t*);
    int (*fp_mi_ctx_set_bucketed_build)(mi_ctx*, uint32_t);
    int (*fp_mi_ctx_set_refinement)(mi_ctx*, int);
    int (*fp_mi_ctx_set_coarsening)(mi_ctx*, int);
    int (*fp_mi_ctx_set_builder)(mi_ctx*, int);
    int (*fp_mi_fm_eigensolve)(mi_ctx*, const mi_hier*, uint64_t, double, double*, uint32_t*, mi_stats*);
    int (*fp_mi_fm_refine)(mi_ctx*, const mi_graph*, const uint32_t*, uint32_t*, uint32_t*, int64_t*);
    int (*fp_mi_fm_init)(mi_ctx*, const mi_graph*, const uint32_t*, uint64_t, double, uint32_t*);
    int (*fp_mi_compute_partition_noalloc)(mi_ctx*, const mi_graph*, const double*, int, int, uint32_t*,
                                           uint32_t*);
    int (*fp_mi_partition_graph)(mi_ctx*, const mi_graph*, const mi_config*, int, uint64_t, uint64_t*, uint64_t,
                                 uint32_t*, int, uint32_t*, mi_stats*);
    int (*fp_mi_use_partition)(mi_ctx*, const mi_graph*, const uint32_t*, int, mi_graph**, mi_graph**);
    int (*fp_mi_partition_4way)(mi_ctx*, const mi_graph*, const mi_config*, int, uint64_t, uint64_t*, uint64_t,
                                uint32_t*, int, uint32_t*, uint32_t*);
    int (*fp_mi_xs_draws)(mi_ctx*, uint64_t, uint64_t, uint64_t, uint64_t*);
    int (*fp_mi_hec_level)(mi_ctx*, const mi_graph*, int, uint64_t, uint32_t*, uint32_t*, uint32_t*, uint32_t*);
    int (*fp_mi_build_coarse)(mi_ctx*, const mi_graph*, const uint32_t*, uint32_t, mi_graph**, uint32_t*);
    int (*fp_mi_spmv_shift)(mi_ctx*, const mi_graph*, const double*, double*);
    int (*fp_mi_bench_step)(mi_ctx*, const mi_graph*, int, mi_step_bench*);
    int (*fp_mi_radix_sort_pairs)(mi_ctx*, uint64_t*, uint32_t*, uint64_t, int, int);
    int (*fp_mi_sort_random_keys)(mi_ctx*, uint64_t*, uint32_t*, uint64_t);
    int (*fp_mi_scan_u32)(mi_ctx*, const uint32_t*, uint32_t*, uint64_t, uint32_t*);

    BIND(mi_ctx_create);
    BIND(mi_ctx_destroy);
    BIND(mi_last_error);
    BIND(mi_ctx_trim);
    BIND(mi_ctx_set_graphs);
    BIND(mi_ctx_set_precision);
    BIND(mi_ctx_set_small_levels);
    BIND(mi_ctx_set_locality);
    BIND(mi_bisect_batch);
    BIND(mi_graph_create);
    BIND(mi_graph_from_edges);
    BIND(mi_graph_create_device);
    BIND(mi_graph_destroy);
    BIND(mi_graph_size);
    BIND(mi_graph_copy);
    BIND(mi_hierarchy_build);
    BIND(mi_hier_create_from_host);
    BIND(mi_hier_destroy);
    BIND(mi_hier_nlevels);
    BIND(mi_hier_level_size);
    BIND(mi_hier_level_copy);
    BIND(mi_fiedler);
    BIND(mi_power_iter);
    BIND(mi_compute_partition);
    BIND(mi_compute_partition_noalloc);
    BIND(mi_ctx_set_bucketed_build);
    BIND(mi_ctx_set_refinement);
    BIND(mi_ctx_set_coarsening);
    BIND(mi_ctx_set_builder);
    BIND(mi_fm_eigensolve);
    BIND(mi_fm_refine);
    BIND(mi_fm_init);
    BIND(mi_partition_graph);
    BIND(mi_use_partition);
    BIND(mi_partition_4way);
    BIND(mi_xs_draws);
    BIND(mi_hec_level);
    BIND(mi_build_coarse);
    BIND(mi_spmv_shift);
    BIND(mi_bench_step);
    BIND(mi_radix_sort_pairs);
    BIND(mi_sort_random_keys);
    BIND(mi_scan_u32);

    int bad = 0;
    mi_config cfg = {4, 0, 0, 1, 1e-10, 2};
    uint32_t part[4], cut = 0;
    uint64_t st = 1;
    if (fp_mi_partition_graph(NULL, NULL, &cfg, 0, 0, &st, 1, part, 0, &cut, NULL) != MI_STATUS_ARG) bad |= 1;
    if (fp_mi_ctx_trim(NULL) != MI_STATUS_ARG) bad |= 2;
    if (strcmp(fp_mi_last_error(NULL), "null context") != 0) bad |= 4;
    if (fp_mi_hier_nlevels(NULL) >= 0) bad |= 8;
    mi_ctx* ctx = NULL;
    const int rc = fp_mi_ctx_create(0, &ctx);
    printf("ctx_create=%d sizeof(mi_config)=%zu sizeof(mi_stats)=%zu bad=%d\n", rc, sizeof(mi_config),
           sizeof(mi_stats), bad);
    if (ctx) fp_mi_ctx_destroy(ctx);
    return bad;
}
