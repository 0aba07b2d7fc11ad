// main-gpu <graph> <device> [group size]
//
// The reference's single-GPU driver (src/main-gpu.cpp:30-113) on this engine:
// same arguments, same stdout lines and loop rule.  The initial partition and
// the initial WCC run on the device (the reference computes them on the host,
// src/main-gpu.cpp:54-55; the labels are identical).  Beyond the reference
// (which writes neither):
//   SCD_LABELS_OUT=<file>  the best labels as a labels file (scd_labels_write:
//                          [u32 n][u32 communities][u32 compact labels n])
//   SCD_RUN_LOG=<file>     a machine-readable run log, one JSON object per
//                          line: set-up, every iteration, the result.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stack>
#include <string>
#include <sys/time.h>
#include <utility>
#include <vector>

#include "scd_dropin.hpp"
#include "scd_graph_layout.hpp"
#include "scd_hip.h"
#include "scd_host.h"

static double now_ms() {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1000.0 + tv.tv_usec / 1000.0;
}

// Nested wall-clock timers printing "name: X ms" on exit (src/common.cpp:27-52).
static std::stack<std::pair<std::string, double>> g_timers;
static void timer_start(const char *name) { g_timers.push({name, now_ms()}); }
static void timer_end() {
    if (g_timers.empty()) return;
    auto t = g_timers.top();
    g_timers.pop();
    std::cout << t.first << ": " << (now_ms() - t.second) << " ms" << std::endl;
}

// JSON-lines run log (SCD_RUN_LOG); doubles with 17 significant digits.
static FILE *g_log = nullptr;

static void must(int rc, const char *what) {
    if (rc != SCD_OK) {
        std::cerr << what << ": " << scd_last_error() << std::endl;
        exit(EXIT_FAILURE);
    }
}

int main(int argc, char *argv[]) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <filename> <device> [group size]" << std::endl;
        return EXIT_FAILURE;
    }
    const double threshold = 0.01, alpha = 1.0;
    const int work_group_size = argc >= 4 ? atoi(argv[3]) : 256;
    if (!cuda_init(argc >= 3 ? atoi(argv[2]) : -1, work_group_size, true)) return EXIT_FAILURE;

    if (const char *lp = getenv("SCD_RUN_LOG")) {
        g_log = fopen(lp, "w");
        if (!g_log) {
            std::cerr << "failed to open " << lp << std::endl;
            return EXIT_FAILURE;
        }
    }
    double t_phase = now_ms();
    timer_start("load graph");
    scd_graph *hg = nullptr;
    if (scd_graph_read(argv[1], &hg) != SCD_OK) {
        std::cerr << scd_host_last_error() << std::endl;
        return EXIT_FAILURE;
    }
    timer_end();
    if (hg->n == 0) {
        std::cerr << "graph has no vertices" << std::endl;
        return EXIT_FAILURE;
    }
    const Graph g(hg->n, hg->m, scd_graph_clustering_coef(hg), hg->tri_deg, hg->tri, hg->indices, hg->adj);

    const double load_ms = now_ms() - t_phase;
    t_phase = now_ms();
    timer_start("transfer graph to device");
    if (!cuda_transfer_graph(&g, alpha)) return EXIT_FAILURE;
    timer_end();
    const double transfer_ms = now_ms() - t_phase;

    // the engine's context is the drop-in library's; reach it through the C ABI calls
    // that take the labels directly
    const unsigned n = g.n;
    std::vector<unsigned> best_labels(n);
    scd_ctx *ctx = scd_dropin_context();
    t_phase = now_ms();
    timer_start("create initial partition");
    uint32_t k = 0;
    must(scd_initial_partition(ctx, best_labels.data(), &k), "create initial partition");
    timer_end();
    const double ip_ms = now_ms() - t_phase;
    t_phase = now_ms();
    timer_start("compute WCC");
    double best_wcc = 0.0;
    must(scd_compute_wcc(ctx, &best_wcc, nullptr, nullptr, nullptr), "compute WCC");
    timer_end();
    const double wcc_ms = now_ms() - t_phase;
    if (g_log)
        fprintf(g_log,
                "{\"event\": \"setup\", \"graph\": \"%s\", \"nodes\": %u, \"edges\": %u, \"avg_cc\": %.17g, "
                "\"initial_labels\": %u, \"initial_wcc\": %.17g, \"load_ms\": %.3f, \"transfer_ms\": %.3f, "
                "\"initial_partition_ms\": %.3f, \"compute_wcc_ms\": %.3f}\n",
                argv[1], g.n, g.m, g.clustering_coef, k, best_wcc, load_ms, transfer_ms, ip_ms, wcc_ms);

    std::cout << std::endl;
    std::cout << "nodes: " << g.n << std::endl;
    std::cout << "edges: " << g.m << std::endl;
    std::cout << "avg. cc: " << g.clustering_coef << std::endl;
    std::cout << "initial WCC: " << best_wcc << std::endl;
    std::cout << "work group size: " << work_group_size << std::endl;
    std::cout << std::endl;

    const double before = now_ms();
    int iters = 0;
    while (true) {
        iters++;
        const double t_it = now_ms();
        timer_start("iteration");
        double wcc = 0;
        cuda_improve_labels(&wcc);
        const bool stored = wcc > best_wcc;
        if (stored) cuda_store_labels();
        timer_end();
        if (g_log)
            fprintf(g_log, "{\"event\": \"iteration\", \"iteration\": %d, \"wcc\": %.17g, \"stored\": %s, \"ms\": %.3f}\n",
                    iters, wcc, stored ? "true" : "false", now_ms() - t_it);
        std::cout << "WCC: " << wcc << std::endl << std::endl;
        if (wcc < best_wcc * (1 + threshold)) break;
        best_wcc = wcc;
    }
    const double after = now_ms();

    timer_start("transfer labels back");
    cuda_labels_from_device(best_labels.data());
    timer_end();

    // the count the reference prints after its compress_labels
    // (src/scd.cpp:32-55), and the real one
    const uint32_t printed = scd_labels_printed_count(n, best_labels.data());
    uint32_t communities = 0;
    {
        std::vector<uint32_t> compact(n);
        if (scd_labels_compact(n, best_labels.data(), compact.data(), &communities) != SCD_OK) {
            std::cerr << scd_host_last_error() << std::endl;
            return EXIT_FAILURE;
        }
    }
    if (const char *out = getenv("SCD_LABELS_OUT")) {
        if (scd_labels_write(out, n, best_labels.data()) != SCD_OK) {
            std::cerr << scd_host_last_error() << std::endl;
            return EXIT_FAILURE;
        }
    }
    if (g_log) {
        fprintf(g_log,
                "{\"event\": \"done\", \"best_wcc\": %.17g, \"iterations\": %d, \"time_per_iteration_ms\": %.3f, "
                "\"printed_communities\": %u, \"communities\": %u}\n",
                best_wcc, iters, (after - before) / iters, printed, communities);
        fclose(g_log);
    }
    std::cout << "----------------------------------------" << std::endl;
    std::cout << "num. communities: " << printed << std::endl;
    std::cout << "WCC: " << best_wcc << std::endl;
    std::cout << "iterations: " << iters << std::endl;
    std::cout << "time per iteration: " << (after - before) / iters << " ms" << std::endl;
    std::cout << "----------------------------------------" << std::endl;
    scd_graph_free(hg);
    return EXIT_SUCCESS;
}
