// sgpar_dropin_main.cpp -- a command-line partitioner written only against
// the reference's API header (cmg/sgpar.h with EXPERIMENT, compiled where it
// lies) and linked with libsgpar_dropin.so: the flow of the reference's
// driver cmg/sgpar.cpp:19-107 (load config and graph, num_iter trials of
// sgp_partition_graph with one PCG stream, the 4-way step through
// sgp_use_partition and two more bisections, the metrics JSON of every
// trial through ExperimentLoggerUtil::log).
//
//   sgpar_dropin csr_file metrics_file config_file [best_partition_file [partition_out]]
//
// Additions of this driver: SGPAR_FIXED_TIME=<t> stands for a pinned
// time(): it seeds the PCG stream and, through sgpar_mi_fix_clock, the
// coarsening's pool; partition_out (or "-" in place of
// best_partition_file to skip the comparison) receives the last trial's
// partition, one label per line (0/1, or 0..3 for four parts).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include "sgpar.h"

extern "C" void sgpar_mi_fix_clock(int64_t t);

using namespace sgpar;

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "Usage: %s csr_filename metrics_filename config_file [best_partition|-] [partition_out]\n",
                     argv[0]);
        return EXIT_FAILURE;
    }
    char* metrics = argv[2];
    config_t config;
    CHECK_SGPAR(sgp_load_config(argv[3], &config));
    sgp_graph_t g;
    CHECK_SGPAR(sgp_load_graph(&g, argv[1]));
    std::printf("n: %lu, m: %lu\n", (unsigned long)g.nvertices, (unsigned long)g.nedges);

    std::vector<sgp_vid_t> part(g.nvertices ? g.nvertices : 1), best(g.nvertices ? g.nvertices : 1);
    const bool compare = argc > 4 && std::strcmp(argv[4], "-") != 0;
    if (compare) CHECK_SGPAR(sgp_load_partition(best.data(), g.nvertices, argv[4]));

    sgp_pcg32_random_t rng;
    rng.state = (uint64_t)std::time(nullptr);
    rng.inc = 1;
    if (const char* t = std::getenv("SGPAR_FIXED_TIME")) {
        const long long v = std::atoll(t);
        sgpar_mi_fix_clock(v);
        rng.state = (uint64_t)v;
    }

    long long best_cut = 1ll << 30;
    std::vector<sgp_vid_t> out4;
    for (int trial = 0; trial < config.num_iter; trial++) {
        sgp_eid_t cut = 0;
        ExperimentLoggerUtil experiment;
        CHECK_SGPAR(sgp_partition_graph(part.data(), &cut, &config, 0, g, experiment, &rng));
        sgp_vid_t diff = 0;
        if (compare) CHECK_SGPAR(compute_partition_edit_distance(part.data(), best.data(), g.nvertices, &diff));
        out4.assign(part.begin(), part.begin() + g.nvertices);
        if (config.num_partitions == 4) {
            sgp_graph_t g1, g2;
            CHECK_SGPAR(sgp_use_partition(part.data(), g, &g1, &g2));
            std::vector<sgp_vid_t> p1(g1.nvertices ? g1.nvertices : 1), p2(g2.nvertices ? g2.nvertices : 1);
            sgp_eid_t c1 = 0, c2 = 0;
            ExperimentLoggerUtil e1, e2;
            CHECK_SGPAR(sgp_partition_graph(p1.data(), &c1, &config, 0, g1, e1, &rng));
            CHECK_SGPAR(sgp_partition_graph(p2.data(), &c2, &config, 0, g2, e2, &rng));
            experiment.setEdgeCut4Way(cut + c1 + c2);
            sgp_vid_t k1 = 0, k2 = 0;
            for (sgp_vid_t i = 0; i < g.nvertices; i++)
                out4[i] = part[i] == 0 ? p1[k1++] : 2 + p2[k2++];
            sgp_free_graph(&g1);
            sgp_free_graph(&g2);
        }
        experiment.setPartitionDiff(diff);
        if ((long long)cut < best_cut) best_cut = cut;
        experiment.log(metrics, trial == 0, trial + 1 == config.num_iter);
    }
    std::printf("graph %s, min edgecut found is %lld\n", argv[1], best_cut);
    if (argc > 5) {
        FILE* f = std::fopen(argv[5], "w");
        if (!f) return EXIT_FAILURE;
        for (sgp_vid_t i = 0; i < g.nvertices; i++) std::fprintf(f, "%u\n", (unsigned)out4[i]);
        std::fclose(f);
    }
    CHECK_SGPAR(sgp_free_graph(&g));
    return EXIT_SUCCESS;
}
