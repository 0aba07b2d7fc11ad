// ORACLE — test infrastructure only.
//
// Harness linked against the reference's own CPU sources (graph.cpp, scd.cpp,
// wcc.cpp, common.cpp, compiled where they lie under the reference tree by
// oracle/Makefile).  It drives them in the order of src/main.cpp:38-93 and
// writes what the golden fixtures need to a binary file:
//   u32 n, u32 num_comms, u32 iterations, u32 printed_num_communities,
//   f64 cc, f64 initial_wcc, f64 final_wcc,
//   u32 labels[n]                       (initial partition)
//   i32 sizes/ext/int[num_comms]        (initial stats)
//   per iteration: u32 labels[n], i32 sizes/ext/int[num_comms], f64 wcc
// plus, on stdout, "REFDUMP refine_ms=<loop wall ms> init_ms=<ms>".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <sys/time.h>
#include <vector>

#include "common.hpp"
#include "graph.hpp"
#include "scd.hpp"
#include "wcc.hpp"

static double wall_ms() {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

static void put(FILE *f, const void *p, size_t bytes) {
    if (fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "write failed\n");
        exit(2);
    }
}

int main(int argc, char *argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <out.bin> [max_iters]\n", argv[0]);
        return 1;
    }
    const unsigned max_iters = argc >= 4 ? (unsigned)atoi(argv[3]) : 1000u;
    const double alpha = 1.0, threshold = 0.01;
    const Graph *g = Graph::read(argv[1]);
    if (!g) return 1;
    const unsigned n = g->n;
    const double t0 = wall_ms();
    Partition *p = create_initial_partition(g);
    double best_wcc = compute_wcc(g, p, alpha);
    const double init_ms = wall_ms() - t0;
    const unsigned num_comms = p->num_communities;

    std::vector<unsigned> history(p->labels, p->labels + n);
    std::vector<int> stats;
    stats.insert(stats.end(), p->sizes, p->sizes + num_comms);
    stats.insert(stats.end(), p->ext_edges, p->ext_edges + num_comms);
    stats.insert(stats.end(), p->int_edges, p->int_edges + num_comms);
    std::vector<double> wccs;
    std::vector<unsigned> best_labels(p->labels, p->labels + n);
    const double initial_wcc = best_wcc;

    const double before = wall_ms();
    unsigned iters = 0;
    while (iters < max_iters) {
        iters++;
        improve_partition(g, p, alpha);
        collect_stats(g, p);
        const double wcc = compute_wcc(g, p, alpha);
        if (wcc > best_wcc) std::copy(p->labels, p->labels + n, best_labels.begin());
        history.insert(history.end(), p->labels, p->labels + n);
        stats.insert(stats.end(), p->sizes, p->sizes + num_comms);
        stats.insert(stats.end(), p->ext_edges, p->ext_edges + num_comms);
        stats.insert(stats.end(), p->int_edges, p->int_edges + num_comms);
        wccs.push_back(wcc);
        if (wcc < best_wcc * (1 + threshold)) break;
        best_wcc = wcc;
    }
    const double refine_ms = wall_ms() - before;
    unsigned printed = 0;
    compress_labels(g, best_labels.data(), &printed);

    FILE *f = fopen(argv[2], "wb");
    if (!f) return 2;
    unsigned hdr[4] = {n, num_comms, iters, printed};
    put(f, hdr, sizeof(hdr));
    put(f, &g->clustering_coef, 8);
    put(f, &initial_wcc, 8);
    put(f, &best_wcc, 8);
    const size_t sc = (size_t)3 * num_comms;
    for (unsigned it = 0; it <= iters; it++) {
        put(f, &history[(size_t)it * n], sizeof(unsigned) * n);
        put(f, &stats[it * sc], sizeof(int) * sc);
        if (it > 0) put(f, &wccs[it - 1], 8);
    }
    fclose(f);
    printf("REFDUMP refine_ms=%.6f init_ms=%.6f iterations=%u\n", refine_ms, init_ms, iters);
    return 0;
}
